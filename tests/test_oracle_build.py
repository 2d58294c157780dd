"""The oracle's CSR build and transpose, pinned by the reference's own known
answers (graph/csr/graph_csr.rs:479-510, graph/csr/transpose.rs:124-140,
graph/edge_list.rs:396-410), by floating-point cases whose result depends on
the summation order, and by a brute-force build on small integer-weighted
edge lists."""
import numpy as np
import pytest

import oracle as O


def pairs(csr):
    out = {}
    for s in range(csr.n):
        for p in range(int(csr.offsets[s]), int(csr.offsets[s + 1])):
            out[(s, int(csr.targets[p]))] = None if csr.weights is None else float(csr.weights[p])
    return out


def test_parallel_edges_fold_summing_weights_and_self_loops_stay():
    # graph_csr.rs:479-492
    g, folded = O.csr_from_edge_list(3, [2, 0, 2, 1, 2], [0, 1, 0, 1, 0], [1.0, 2.0, 3.0, 4.0, 5.0])
    assert pairs(g) == {(0, 1): 2.0, (1, 1): 4.0, (2, 0): 9.0}
    assert folded == 2
    assert g.validate() == 0


def test_merged_csr_unites_relations():
    # graph_csr.rs:495-510: relations a = 0, b = 1.
    g, folded = O.csr_from_edge_list(2, [0, 0, 1], [1, 1, 0], [1.0, 2.0, 3.0], relations=[1, 0, 0],
                                     relation_count=2)
    assert pairs(g) == {(0, 1): 3.0, (1, 0): 3.0}
    assert pairs(g.transpose()) == {(1, 0): 3.0, (0, 1): 3.0}
    assert folded == 0


def test_transpose_known_answer():
    # transpose.rs:124-140: 0 -> {1, 2}, 1 -> {2}, 2 -> {0, 2}.
    f = O.Csr(3, [0, 2, 3, 5], [1, 2, 2, 0, 2], [1.0, 2.0, 3.0, 4.0, 5.0])
    r = f.transpose()
    assert list(r.offsets) == [0, 1, 2, 5]
    assert list(r.targets) == [2, 0, 0, 1, 2]
    assert list(r.weights) == [4.0, 1.0, 2.0, 3.0, 5.0]


def test_first_bad_edge_is_the_error():
    # edge_list.rs:396-410 rejects 0 -> 2 against an id space of 2.
    with pytest.raises(O.OracleError) as e:
        O.csr_from_edge_list(2, [0], [2])
    assert e.value.code == 21 and e.value.edge == 0
    # The lowest failing edge wins; an edge failing both checks is an id
    # space error (counted(), build/mod.rs:329-352).
    with pytest.raises(O.OracleError) as e:
        O.csr_from_edge_list(4, [0, 1, 9, 0], [1, 2, 0, 1], relations=[0, 0, 5, 7], relation_count=2)
    assert e.value.code == 21 and e.value.edge == 2
    with pytest.raises(O.OracleError) as e:
        O.csr_from_edge_list(4, [0, 1, 2, 9], [1, 2, 0, 1], relations=[0, 3, 0, 0], relation_count=2)
    assert e.value.code == 22 and e.value.edge == 1


def test_weights_sum_in_input_order_then_relation_order():
    big = 1e16
    # One relation: ((1e16 + 1) - 1e16) in input order is 0; sorted by value
    # (or summed pairwise) it would not be.
    g, _ = O.csr_from_edge_list(1, [0, 0, 0], [0, 0, 0], [big, 1.0, -big])
    assert g.weights[0] == 0.0
    g, _ = O.csr_from_edge_list(1, [0, 0, 0], [0, 0, 0], [big, -big, 1.0])
    assert g.weights[0] == 1.0
    # Two relations: relation 0's sum (1) plus relation 1's sum (1e16 - 1e16
    # = 0) is 1, where the input-order sum of all three would be 0.
    g, folded = O.csr_from_edge_list(1, [0, 0, 0], [0, 0, 0], [big, 1.0, -big], relations=[1, 0, 1],
                                     relation_count=2)
    assert g.weights[0] == 1.0 and folded == 1
    # Relation order, not input order: relation 1 first in the input.
    g, _ = O.csr_from_edge_list(1, [0, 0, 0], [0, 0, 0], [-big, big, 1.0], relations=[1, 0, 0],
                                relation_count=2)
    # relation 0: 1e16 + 1 = 1e16; relation 1: -1e16; total 0.
    assert g.weights[0] == 0.0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_brute_force_build_on_integer_weights(seed):
    rng = np.random.default_rng(seed)
    n, m, nrel = 40, 600, 3
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    rel = rng.integers(0, nrel, m)
    w = rng.integers(-8, 9, m).astype(np.float64)  # exact sums in any order
    g, folded = O.csr_from_edge_list(n, src, dst, w, relations=rel, relation_count=nrel)
    naive = {}
    for s, t, r, x in zip(src, dst, rel, w):
        naive.setdefault((int(s), int(t)), {}).setdefault(int(r), []).append(x)
    assert pairs(g) == {k: float(sum(sum(v) for v in per.values())) for k, per in naive.items()}
    assert folded == m - sum(len(per) for per in naive.values())
    assert g.validate() == 0
    # The transpose holds every edge reversed, rows ascending.
    r = g.transpose()
    assert pairs(r) == {(t, s): x for (s, t), x in pairs(g).items()}
    assert r.validate() == 0
    # Unweighted: the same pattern.
    u, _ = O.csr_from_edge_list(n, src, dst)
    assert set(pairs(u)) == set(naive)


def test_raw_rmat_edges_build_to_the_rmat_csr():
    """The raw generator output, self loops dropped, built by the oracle, is the
    CSR the R-MAT generator returns (host code of the library; no GPU)."""
    import pg_warren_amd as W
    src, dst = W.rmat_edges(12, 16, seed=4)
    off, tgt = W.rmat(12, 16, seed=4)
    keep = src != dst
    g, folded = O.csr_from_edge_list(1 << 12, src[keep], dst[keep])
    assert np.array_equal(g.offsets, off) and np.array_equal(g.targets, tgt)
    assert folded == int(keep.sum()) - len(tgt)


# ---------------------------------------------------------------- relations
# compute/product/tests.rs:32-38: relation 0: 0 -> 1 (2), 0 -> 2 (3),
# 1 -> 2 (4), 2 -> 0 (5); relation 1: 0 -> 1 (10). Node 3 is isolated.
EX_EDGES = [(0, 0, 1, 2.0), (0, 0, 2, 3.0), (0, 1, 2, 4.0), (0, 2, 0, 5.0), (1, 0, 1, 10.0)]
EX_X = np.array([1.0, 2.0, 3.0, 4.0])


def example_relations():
    rel, src, dst, w = (np.array(c) for c in zip(*EX_EDGES))
    return O.relation_csrs(4, src, dst, w, rel, 2)


def test_relation_known_answers():
    csrs = example_relations()
    # compute/product/tests.rs:60-69: relation 1 alone; every relation
    # unweighted.
    assert list(O.spmv(O.select(csrs, [1]), EX_X)) == [20.0, 0.0, 0.0, 0.0]
    merged = O.select(csrs)
    assert list(O.spmv(merged, EX_X)) == [33.0, 12.0, 5.0, 0.0]
    assert list(O.spmv(merged, EX_X, weighted=False)) == [5.0, 3.0, 1.0, 0.0]
    assert list(O.spmv(merged, EX_X, rev=merged.transpose(), transpose=True)) == [15.0, 12.0, 11.0, 0.0]
    # Per relation the 0 -> 1 edge counts in both relations: unweighted row 0
    # is 2 + 2 + 3, and min_plus sees the lighter edge's own path.
    per = O.select(csrs, per_relation=True)
    assert list(O.spmv(per, EX_X, weighted=False)) == [7.0, 3.0, 1.0, 0.0]
    assert O.spmv(per, EX_X, "min_plus")[0] == 4.0
    assert O.spmv(merged, EX_X, "min_plus")[0] == 6.0


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_per_relation_is_each_relations_product_added(seed):
    """compute/product/tests.rs:611-650 and :651-662, on integer weights and
    values (every semiring exact)."""
    rng = np.random.default_rng(seed)
    n, m, R = 30, 200, 3
    rel = rng.integers(0, R, m)
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.integers(1, 5, m).astype(np.float64)
    x = rng.integers(-3, 4, n).astype(np.float64)
    csrs = O.relation_csrs(n, src, dst, w, rel, R)
    identity = {"plus_times": 0.0, "min_plus": np.inf, "max_times": -np.inf}
    for sr in ("plus_times", "min_plus", "max_times"):
        found = O.spmv(O.select(csrs, per_relation=True), x, sr)
        parts = [O.spmv(c, x, sr) for c in csrs]
        has = [np.diff(c.offsets.astype(np.int64)) > 0 for c in csrs]
        expected = []
        for row in range(n):
            vals = [p[row] for p, h in zip(parts, has) if h[row]]
            if not vals:
                expected.append(identity[sr])
                continue
            acc = vals[0]
            for v in vals[1:]:
                acc = acc + v if sr == "plus_times" else (min(acc, v) if sr == "min_plus" else max(acc, v))
            expected.append(acc)
        assert list(found) == expected, sr
        # By default: the product of the graph with every edge in one relation.
        dedup, _ = O.csr_from_edge_list(n, src, dst, w)
        assert list(O.spmv(O.select(csrs), x, sr)) == list(O.spmv(dedup, x, sr)), sr


def test_merged_csr_and_merged_rows_agree_bit_for_bit():
    """compute/product/tests.rs:507-530: with fractional weights (-0.0 among
    them) the stored merged CSR and rows merged on the fly give the same
    bits, since f64 sums start from their first term."""
    rng = np.random.default_rng(5)
    n, m, R = 25, 400, 4
    rel = rng.integers(0, R, m)
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.choice([0.1, 0.7, 1e-3, 3.0, -0.0], m)
    stored, _ = O.csr_from_edge_list(n, src, dst, w, relations=rel, relation_count=R)
    on_the_fly = O.select(O.relation_csrs(n, src, dst, w, rel, R))
    assert np.array_equal(stored.offsets, on_the_fly.offsets)
    assert np.array_equal(stored.targets, on_the_fly.targets)
    assert np.array_equal(stored.weights.view(np.uint64), on_the_fly.weights.view(np.uint64))
