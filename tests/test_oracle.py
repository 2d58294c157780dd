"""Pins the CPU oracle against the reference's own known answers.

Each expected value below is transcribed from the reference test it cites
(paths relative to crates/warren-core/src/), or follows from mathematics
(closed forms, brute force on tiny graphs, NetworkX's published algorithm).
"""
import math

import numpy as np
import pytest

I64MAX = 2**63 - 1
I64MIN = -(2**63)


# ---------------------------------------------------------------- fixtures
def example(o, relations=(0, 1)):
    """compute/product/tests.rs:32-38: relation 0: 0->1 (2), 0->2 (3),
    1->2 (4), 2->0 (5); relation 1: 0->1 (10); node 3 isolated. The merged
    CSR sums an edge's weights over the relations."""
    edges = [(0, 0, 1, 2.0), (0, 0, 2, 3.0), (0, 1, 2, 4.0), (0, 2, 0, 5.0), (1, 0, 1, 10.0)]
    edges = [e for e in edges if e[0] in relations]
    src = [e[1] for e in edges]
    dst = [e[2] for e in edges]
    w = [e[3] for e in edges]
    return o.csr_from_edges(4, src, dst, w)


def small_edges():
    """compute/pagerank/tests.rs:56-66."""
    return [(0, 0, 1, 1.0), (0, 1, 2, 1.0), (1, 2, 3, 1.0), (1, 0, 3, 5.0), (0, 3, 0, 1.0),
            (1, 4, 4, 2.0), (1, 0, 1, 0.5)]


def merged(o, n, edges):
    # Merged CSR: relation order then input order within a relation
    # (build/deduplication.rs:1-7).
    edges = sorted(edges, key=lambda e: e[0])
    return o.csr_from_edges(n, [e[1] for e in edges], [e[2] for e in edges],
                            [e[3] for e in edges])


def bfs_example(o):
    """traversal/bfs.rs:151-154: 0->1->2->3, 0->4->3, 2->2, 5 isolated."""
    e = [(0, 1), (1, 2), (2, 3), (0, 4), (4, 3), (2, 2)]
    return o.csr_from_edges(6, [a for a, _ in e], [b for _, b in e])


# ---------------------------------------------------------------- semirings
def test_binary_op_tables(oracle):
    """compute/semiring.rs:451-461."""
    ops = oracle.BINARY_OPS
    assert [oracle.apply(op, 2.0, -3.0, "f64") for op in ops] == \
        [2.0, -3.0, 1.0, -1.0, -6.0, -3.0, 2.0, 1.0, 1.0, 2.0]
    assert [oracle.apply(op, 0.0, 4.0, "f32") for op in ops] == \
        [0.0, 4.0, 1.0, 4.0, 0.0, 0.0, 4.0, 1.0, 0.0, 0.0]
    assert [oracle.apply(op, 5, 0, "i64") for op in ops] == [5, 0, 1, 5, 0, 0, 5, 1, 0, 5]
    assert [bool(oracle.apply(op, 1, 0, "bool")) for op in ops] == \
        [True, False, True, True, False, False, True, True, False, True]


def test_min_max_sql_order(oracle):
    """compute/semiring.rs:463-472."""
    nan, inf = float("nan"), float("inf")
    assert oracle.apply("min", nan, 1.0, "f64") == 1.0
    assert math.isnan(oracle.apply("max", 1.0, nan, "f64"))
    assert oracle.apply("min", inf, nan, "f64") == inf
    assert oracle.apply("lor", nan, 0.0, "f64") == 1.0
    assert math.copysign(1.0, oracle.apply("min", -0.0, 0.0, "f64")) < 0


def test_i64_identities_absorb(oracle):
    """compute/semiring.rs:474-508."""
    m = lambda add, mul, w, v: oracle.apply(mul, w, v, "i64", monoid=add)
    assert m("min", "plus", 3, I64MAX) == I64MAX
    assert m("min", "plus", -3, I64MAX) == I64MAX
    assert m("max", "plus", 3, I64MIN) == I64MIN
    assert m("min", "plus", 3, 4) == 7
    assert m("min", "plus", 1, I64MAX - 1) == I64MAX
    for args in [("min", "plus", 2, I64MAX - 1), ("max", "plus", 1, I64MAX),
                 ("min", "times", -2, I64MAX), ("max", "times", -1, I64MIN),
                 ("min", "times", 2, I64MAX // 2 + 1), ("max", "times", 2, I64MAX)]:
        with pytest.raises(oracle.OracleError) as e:
            m(*args)
        assert e.value.name == "bigint_out_of_range"
    assert m("min", "times", 3, I64MAX) == I64MAX
    assert m("max", "times", 3, I64MIN) == I64MIN
    assert m("min", "times", 1, I64MAX) == I64MAX
    assert m("min", "times", 0, I64MAX) == 0
    assert m("min", "times", -1, I64MAX) == -I64MAX
    assert oracle.apply("plus", 1.0, float("inf"), "f64", monoid="min") == float("inf")
    with pytest.raises(oracle.OracleError):
        oracle.apply("plus", I64MAX, 1, "i64")
    with pytest.raises(oracle.OracleError):
        oracle.apply("times", I64MIN, -1, "i64")


def test_weight_casts(oracle):
    """compute/semiring.rs:519-533 and compute/cast.rs:118-150."""
    fw = oracle.from_weight
    assert fw(0.1, "f64") == 0.1
    assert fw(0.1, "f32") == np.float32(0.1)
    for w, err in [(1e300, "float_overflow"), (1e-300, "float_underflow"),
                   (1e39, "float_overflow"), (1e-50, "float_underflow")]:
        with pytest.raises(oracle.OracleError) as e:
            fw(w, "f32")
        assert e.value.name == err
    assert fw(1.5e-45, "f32") == np.float32(1e-45)
    assert math.isnan(fw(float("nan"), "f32"))
    assert fw(float("inf"), "f32") == np.float32("inf")
    assert fw(2.5, "i64") == 2 and fw(3.5, "i64") == 4 and fw(-2.5, "i64") == -2
    for w in [9.3e18, float("nan"), float(I64MAX)]:
        with pytest.raises(oracle.OracleError) as e:
            fw(w, "i64")
        assert e.value.name == "bigint_out_of_range"
    assert fw(float(I64MIN), "i64") == I64MIN
    assert fw(0.0, "bool") == 0 and fw(float("nan"), "bool") == 1


def test_u32_is_the_image_of_i64(oracle):
    """u32 maps to i64 with 2^32-1 as i64::MAX: every op on the image
    agrees with i64, or is out of range."""
    inf = 2**32 - 1
    rng = np.random.default_rng(3)
    vals = [0, 1, 2, 7, 255, 2**31, 2**32 - 2, inf] + list(rng.integers(0, 2**32 - 1, 20))
    to_i = lambda v: I64MAX if v == inf else int(v)
    for add in ["min", "plus", "lor", "any"]:
        for op in oracle.BINARY_OPS:
            for a in vals[:10]:
                for b in vals[:10]:
                    try:
                        ref = oracle.apply(op, to_i(a), to_i(b), "i64", monoid=add)
                        ok = ref == I64MAX or 0 <= ref <= 2**32 - 2
                    except oracle.OracleError:
                        ok = False
                    if ok:
                        got = oracle.apply(op, a, b, "u32", monoid=add)
                        assert to_i(got) == ref, (add, op, a, b)
                    else:
                        with pytest.raises(oracle.OracleError):
                            oracle.apply(op, a, b, "u32", monoid=add)


# ---------------------------------------------------------------- CSR
def test_csr_build_and_transpose(oracle):
    g = example(oracle)
    assert list(g.offsets) == [0, 2, 3, 4, 4]
    assert list(g.targets) == [1, 2, 2, 0]
    assert list(g.weights) == [12.0, 3.0, 4.0, 5.0]
    r = g.transpose()
    assert list(r.offsets) == [0, 1, 2, 4, 4]
    assert list(r.targets) == [2, 0, 0, 1]
    assert list(r.weights) == [5.0, 12.0, 3.0, 4.0]
    assert g.validate() == 0 and r.validate() == 0
    bad = oracle.Csr(2, [0, 2, 2], [1, 1])  # row not strictly ascending
    assert bad.validate() == 14  # invalid_csr


def test_dedup_sums_in_input_order_and_keeps_self_loops(oracle):
    """graph/csr/graph_csr.rs:478-492: parallel edges fold, weights summed;
    self loops stay."""
    g = oracle.csr_from_edges(3, [0, 0, 1, 0], [1, 1, 1, 1], [0.1, 0.2, 5.0, 0.3])
    assert list(g.targets) == [1, 1]
    assert g.weights[0] == (0.1 + 0.2) + 0.3
    assert g.weights[1] == 5.0


# ---------------------------------------------------------------- products
X4 = np.array([1.0, 2.0, 3.0, 4.0])


def test_plus_times_known_answers(oracle):
    """compute/product/tests.rs:51-58, 60-69."""
    g = example(oracle)
    r = g.transpose()
    assert list(oracle.spmv(g, X4, rev=r)) == [33.0, 12.0, 5.0, 0.0]
    assert list(oracle.spmv(g, X4, rev=r, transpose=True)) == [15.0, 12.0, 11.0, 0.0]
    assert list(oracle.spmv(g, X4, transpose=True)) == [15.0, 12.0, 11.0, 0.0]  # push
    only1 = example(oracle, relations=(1,))
    assert list(oracle.spmv(only1, X4)) == [20.0, 0.0, 0.0, 0.0]
    assert list(oracle.spmv(g, X4, weighted=False)) == [5.0, 3.0, 1.0, 0.0]


def test_min_plus_empty_row_is_identity(oracle):
    """compute/product/tests.rs:71-75."""
    g = example(oracle)
    assert list(oracle.spmv(g, X4, "min_plus")) == [6.0, 7.0, 6.0, float("inf")]


def test_masks_replace_accumulate(oracle):
    """compute/product/tests.rs:77-96."""
    g = example(oracle)
    mask = np.array([1, 0, -7, 0], np.int64)

    def w(complement, replace, acc):
        y = np.full(4, 100.0)
        return list(oracle.spmv(g, X4, mask=mask, mask_complement=complement, replace=replace,
                                accumulate=acc, y=y))
    assert w(False, False, None) == [33.0, 100.0, 5.0, 100.0]
    assert w(False, True, None) == [33.0, 0.0, 5.0, 0.0]
    assert w(False, False, "plus") == [133.0, 100.0, 105.0, 100.0]
    assert w(True, False, None) == [100.0, 12.0, 100.0, 0.0]
    assert w(True, True, "plus") == [0.0, 112.0, 0.0, 100.0]


def test_spmm_known_answer(oracle):
    """compute/product/tests.rs:98-106."""
    g = example(oracle)
    x = np.array([1.0, 10.0, 2.0, 20.0, 3.0, 30.0, 4.0, 40.0])
    y = oracle.spmv(g, x, rev=g.transpose(), transpose=True, width=2)
    assert list(y) == [15.0, 150.0, 12.0, 120.0, 11.0, 110.0, 0.0, 0.0]


def test_i64_min_plus_supersteps(oracle):
    """compute/product/tests.rs:160-185."""
    g = oracle.csr_from_edges(3, [0, 2], [1, 1], [1.0, 1.0])
    r = g.transpose()
    d = np.array([0, I64MAX, I64MAX], np.int64)
    for _ in range(2):
        d = oracle.spmv(g, d.copy(), "min_plus", rev=r, transpose=True, accumulate="min", y=d)
        assert list(d) == [0, 1, I64MAX]
    y = np.full(3, I64MIN, np.int64)
    y = oracle.spmv(g, np.array([5, I64MIN, I64MIN], np.int64), "max_plus", rev=r,
                    transpose=True, accumulate="max", y=y)
    assert list(y) == [I64MIN, 6, I64MIN]


def test_i64_overflow_is_an_error(oracle):
    """compute/product/tests.rs:187-197."""
    g = example(oracle)
    with pytest.raises(oracle.OracleError) as e:
        oracle.spmv(g, np.full(4, I64MAX, np.int64), y=np.zeros(4, np.int64))
    assert e.value.name == "bigint_out_of_range"


def random_graph(o, n, m, seed, weights=True, lo=1, hi=4):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(lo, hi, m).astype(np.float64) if weights else None
    return o.csr_from_edges(n, src, dst, w)


def dense_product(g, x, add, mul, ident):
    """Naive dense reference (compute/reference.rs:114-157): fold each row's
    entries in ascending column order, first entry initialising."""
    n = g.n
    y = []
    for i in range(n):
        acc = None
        for p in range(int(g.offsets[i]), int(g.offsets[i + 1])):
            j = int(g.targets[p])
            t = mul(float(g.weights[p]), x[j])
            acc = t if acc is None else add(acc, t)
        y.append(ident if acc is None else acc)
    return y


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pull_matches_dense_reference_and_push(oracle, seed):
    g = random_graph(oracle, 37, 200, seed)
    r = g.transpose()
    x = np.random.default_rng(seed + 10).random(37)
    got = oracle.spmv(g, x)
    want = dense_product(g, list(x), lambda a, b: a + b, lambda w, v: w * v, 0.0)
    assert list(got) == want
    got = oracle.spmv(g, x, "min_plus")
    want = dense_product(g, list(x), min, lambda w, v: w + v, float("inf"))
    assert list(got) == want
    # The push over the forward CSR equals the pull over the reverse, bit
    # for bit (compute/product/tests.rs:447).
    for sr in oracle.SEMIRINGS:
        for dt in ["f64", "f32", "bool"]:
            xv = (x * 3).astype(oracle.NP_DTYPES[dt])
            a = oracle.spmv(g, xv, sr, rev=r, transpose=True)
            b = oracle.spmv(g, xv, sr, transpose=True)
            assert a.tobytes() == b.tobytes(), (sr, dt)


# ---------------------------------------------------------------- PageRank
def test_pagerank_golden_ranks(oracle):
    """compute/pagerank/tests.rs:78-123, to the last bit."""
    g = merged(oracle, 6, small_edges())
    vals, it = oracle.pagerank(g, g.transpose())
    assert it == 14
    assert list(vals) == [0.28666944658969795, 0.08535717971974721, 0.10168006797333676,
                          0.3029923348432875, 0.19417475728151729, 0.029126213592413387]
    edges = [e for e in small_edges() if e[1] != 1 and e[2] != 1]
    g = merged(oracle, 6, edges)
    tomb = np.array([0, 1, 0, 0, 0, 0], np.uint8)
    vals, it = oracle.pagerank(g, g.transpose(), tombstones=tomb)
    assert it == 69
    assert [vals[i] for i in [0, 2, 3, 4, 5]] == [
        0.33506897700306837, 0.03614457831325302, 0.3516780109487391, 0.24096385542168675,
        0.03614457831325302]
    assert vals[1] == 0.0
    seed = np.arange(1, 7, dtype=np.float64)
    vals, it = oracle.pagerank(g, g.transpose(), tombstones=tomb, seed=seed)
    assert it == 71
    assert [vals[i] for i in [0, 2, 3, 4, 5]] == [
        0.2553952337234808, 0.03237410071942446, 0.28777128327185025, 0.3597111808463956,
        0.06474820143884892]


def test_pagerank_ring_closed_form(oracle):
    """compute/pagerank/tests.rs:125-137: a ring ranks every node 1/4."""
    g = oracle.csr_from_edges(4, [0, 1, 2, 3], [1, 2, 3, 0])
    vals, _ = oracle.pagerank(g, tol=1e-12)
    assert np.all(np.abs(vals - 0.25) < 1e-12)
    vals, _ = oracle.pagerank(g, tol=1e-12, seed=np.array([1.0, 0, 0, 0]))
    assert abs(vals.sum() - 1.0) < 1e-12
    assert vals[0] > vals[1] > vals[2] > vals[3]


def test_pagerank_errors_and_no_iterations(oracle):
    """compute/pagerank/tests.rs:139-169: max_iter 0 keeps the uniform start;
    a NaN tol stops after one iteration; no live node divides by zero."""
    g = merged(oracle, 6, small_edges())
    vals, it = oracle.pagerank(g, max_iter=0)
    assert it == 0 and np.all(vals == 1.0 / 6)
    _, it = oracle.pagerank(g, tol=float("nan"))
    assert it == 1
    _, it = oracle.pagerank(g, tol=None)
    assert it == 100
    with pytest.raises(oracle.OracleError) as e:
        oracle.pagerank(g, tombstones=np.ones(6, np.uint8))
    assert e.value.name == "division_by_zero"
    with pytest.raises(oracle.OracleError) as e:
        oracle.pagerank(g, seed=np.zeros(6))
    assert e.value.name == "division_by_zero"


def test_pagerank_matches_networkx(oracle):
    """NetworkX's power iteration (the reference's SQL oracle,
    test/sql/oracle_pagerank.sql:1-17, within 1e-9)."""
    nx = pytest.importorskip("networkx")
    g = random_graph(oracle, 300, 2000, 7)
    G = nx.DiGraph()
    G.add_nodes_from(range(300))
    for i in range(300):
        for p in range(int(g.offsets[i]), int(g.offsets[i + 1])):
            G.add_edge(i, int(g.targets[p]), weight=float(g.weights[p]))
    want = nx.pagerank(G, alpha=0.85, tol=1e-13, max_iter=1000, weight="weight")
    got, _ = oracle.pagerank(g, tol=1e-13, max_iter=1000)
    assert np.abs(got - np.array([want[i] for i in range(300)])).sum() < 1e-9
    seed = np.random.default_rng(1).random(300)
    want = nx.pagerank(G, alpha=0.85, tol=1e-13, max_iter=1000, weight="weight",
                       personalization={i: seed[i] for i in range(300)})
    got, _ = oracle.pagerank(g, tol=1e-13, max_iter=1000, seed=seed)
    assert np.abs(got - np.array([want[i] for i in range(300)])).sum() < 1e-9


# ---------------------------------------------------------------- BFS
def v(node, depth, via=None):
    return (node, depth, via)


def test_bfs_known_answers(oracle):
    """traversal/bfs.rs:160-235."""
    g = bfs_example(oracle)
    got, st = oracle.bfs(g, [0])
    assert got == [v(0, 0), v(1, 1, 0), v(4, 1, 0), v(2, 2, 1), v(3, 2, 4)]
    assert st["nodes_visited"] == 5 and st["edges_traversed"] == 6 and not st["truncated"]
    assert oracle.bfs(g, [2, 4, 2])[0] == [v(2, 0), v(4, 0), v(3, 1, 2)]
    assert oracle.bfs(g, [4, 1])[0] == [v(4, 0), v(1, 0), v(2, 1, 1), v(3, 1, 4)]
    got, _ = oracle.bfs(g.transpose(), [3])
    assert [(a, b) for a, b, _ in got] == [(3, 0), (2, 1), (4, 1), (1, 2), (0, 2)]
    got, st = oracle.bfs(g, [0], max_depth=1)
    assert len(got) == 3 and not st["truncated"]
    assert oracle.bfs(g, [0], max_depth=0)[0] == [v(0, 0)]
    got, st = oracle.bfs(g, [0], max_results=2)
    assert got == [v(0, 0), v(1, 1, 0)] and st["truncated"]
    got, st = oracle.bfs(g, [0], max_results=5)
    assert len(got) == 5 and not st["truncated"]
    flt = np.zeros(6, np.uint8)
    flt[[1, 2, 3]] = 1
    got, _ = oracle.bfs(g, [0], node_filter=flt)
    assert [(a, b) for a, b, _ in got] == [(0, 0), (1, 1), (2, 2), (3, 3)]
    with pytest.raises(oracle.OracleError) as e:
        oracle.bfs(g, [6])
    assert e.value.name == "node_outside_id_space"


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_pulled_levels_find_the_same_depths_and_vias(oracle, seed):
    """Direction-optimising search (reachable.rs:79-90): same set, depths and
    min-predecessor vias as the push-only search; pulled levels ascend."""
    g = random_graph(oracle, 400, 4000, seed, weights=False)
    r = g.transpose()
    push, _ = oracle.bfs(g, [0, 5])
    pull, st = oracle.bfs(g, [0, 5], pull_rev=r)
    assert st["levels_pulled"] > 0
    assert sorted(push) == sorted(pull)
    depths = oracle.bfs_depths(g, [0, 5]).min(axis=0)
    for node, depth, via in push:
        assert depths[node] == depth
        if via is not None:  # via is the smallest predecessor one level up
            preds = [u for u in range(400) if depths[u] == depth - 1 and
                     node in g.targets[int(g.offsets[u]):int(g.offsets[u + 1])]]
            assert via == min(preds)


def test_bfs_depths_brute_force(oracle):
    g = random_graph(oracle, 30, 60, 4, weights=False)
    d = oracle.bfs_depths(g, [0, 3])
    # Floyd-Warshall on the tiny graph.
    INF = 10**9
    D = [[0 if i == j else INF for j in range(30)] for i in range(30)]
    for i in range(30):
        for p in range(int(g.offsets[i]), int(g.offsets[i + 1])):
            D[i][int(g.targets[p])] = min(D[i][int(g.targets[p])], 1)
    for k in range(30):
        for i in range(30):
            for j in range(30):
                if D[i][k] + D[k][j] < D[i][j]:
                    D[i][j] = D[i][k] + D[k][j]
    for row, s in enumerate([0, 3]):
        for j in range(30):
            assert d[row, j] == (2**32 - 1 if D[s][j] >= INF else D[s][j])


# ---------------------------------------------------------------- RNG
def test_uniform_splitmix(oracle):
    """compute/random.rs:14-26. mix(0) = 0, so uniform(0, 0, 0) = -1; and
    uniform(0, 0, 1) = mix(GOLDEN), SplitMix64's first output from state 0,
    published as 0xe220a8397b1dcdaf (Vigna's splitmix64.c)."""
    assert oracle.uniform(0, 0, 0) == -1.0
    assert oracle.uniform(0, 0, 1) == (0xE220A8397B1DCDAF >> 11) / 2.0**52 - 1.0
    vals = np.array([oracle.uniform(1, 0, i) for i in range(10000)])
    assert np.all((vals >= -1) & (vals < 1)) and abs(vals.mean()) < 0.05
    assert oracle.uniform(7, 0, 3) != oracle.uniform(7, 1, 3)
