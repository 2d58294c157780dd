"""Parity of the device CSR build (wg_graph_build) and the device transpose
(wg_graph_upload's reverse CSR) with the oracle: offsets, targets and
weights bit-exact, forward and reverse, on the reference's known answers,
on random edge lists with parallel edges, self loops and relations, on rows
and columns longer than a row block, and on the error paths."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def same(a, b):
    """Bit-exact equality of two float arrays (NaN payloads and -0.0 included)."""
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def check(ctx, oracle, n, src, dst, w=None, rel=None, nrel=1):
    print(f"build n={n} m={len(src)} weighted={w is not None} relations={nrel}", flush=True)
    ref, folded = oracle.csr_from_edge_list(n, src, dst, w, relations=rel, relation_count=nrel)
    g = W.Graph.build(ctx, n, src, dst, w, relations=rel, relation_count=nrel)
    off, tgt, gw = g.forward()
    assert g.nnz == len(ref.targets)
    assert np.array_equal(off, ref.offsets)
    assert np.array_equal(tgt, ref.targets)
    if w is not None:
        assert same(gw, ref.weights)
    assert g.deduplicated == folded
    rref = ref.transpose()
    roff, rtgt, rw = g.reverse()
    assert np.array_equal(roff, rref.offsets)
    assert np.array_equal(rtgt, rref.targets)
    if w is not None:
        assert same(rw, rref.weights)
    return g


def test_reference_known_answers(ctx, oracle):
    # graph/csr/graph_csr.rs:479-492: parallel edges folded, self loop kept.
    g = check(ctx, oracle, 3, [2, 0, 2, 1, 2], [0, 1, 0, 1, 0], [1.0, 2.0, 3.0, 4.0, 5.0])
    off, tgt, w = g.forward()
    assert list(off) == [0, 1, 2, 3] and list(tgt) == [1, 1, 0] and list(w) == [2.0, 4.0, 9.0]
    assert g.deduplicated == 2
    # graph_csr.rs:495-510: the merged CSR of two relations.
    g = check(ctx, oracle, 2, [0, 0, 1], [1, 1, 0], [1.0, 2.0, 3.0], rel=[1, 0, 0], nrel=2)
    off, tgt, w = g.forward()
    assert list(tgt) == [1, 0] and list(w) == [3.0, 3.0] and g.deduplicated == 0
    # graph/csr/transpose.rs:124-140.
    g = check(ctx, oracle, 3, [0, 0, 1, 2, 2], [1, 2, 2, 0, 2], [1.0, 2.0, 3.0, 4.0, 5.0])
    roff, rtgt, rw = g.reverse()
    assert list(roff) == [0, 1, 2, 5] and list(rtgt) == [2, 0, 0, 1, 2]
    assert list(rw) == [4.0, 1.0, 2.0, 3.0, 5.0]


def test_summation_order(ctx, oracle):
    big = 1e16
    for w, rel, nrel, want in [([big, 1.0, -big], None, 1, 0.0), ([big, -big, 1.0], None, 1, 1.0),
                               ([big, 1.0, -big], [1, 0, 1], 2, 1.0),
                               ([-big, big, 1.0], [1, 0, 0], 2, 0.0)]:
        g = check(ctx, oracle, 1, [0, 0, 0], [0, 0, 0], w, rel=rel, nrel=nrel)
        assert g.forward()[2][0] == want


@pytest.mark.parametrize("seed,n,m,nrel,weighted", [
    (0, 1000, 20000, 1, True), (1, 1000, 20000, 3, True), (2, 1 << 16, 1 << 20, 1, True),
    (3, 1 << 16, 1 << 20, 5, True), (4, 5000, 100000, 1, False), (5, 300, 50000, 2, False)])
def test_random_edge_lists(ctx, oracle, seed, n, m, nrel, weighted):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    # Parallel edges and self loops: repeat a fifth of the edges, loop a few.
    rep = rng.integers(0, m, m // 5)
    src[rng.integers(0, m, m // 5)] = src[rep]
    dst[rng.integers(0, m, m // 5)] = dst[rep]
    loops = rng.integers(0, m, m // 50)
    dst[loops] = src[loops]
    w = rng.standard_normal(m) * 10.0 ** rng.integers(-8, 9, m) if weighted else None
    rel = rng.integers(0, nrel, m) if nrel > 1 else None
    check(ctx, oracle, n, src, dst, w, rel=rel, nrel=nrel)


def test_long_rows_and_columns(ctx, oracle):
    """A hub row and a hub column far beyond a 2048-entry row block, and one
    (source, target) repeated 100,000 times (a long run summed in order)."""
    rng = np.random.default_rng(7)
    n = 4096
    hub_out = np.stack([np.full(30000, 5), rng.integers(0, n, 30000)])
    hub_in = np.stack([rng.integers(0, n, 30000), np.full(30000, 9)])
    run = np.stack([np.full(100000, 17), np.full(100000, 23)])
    rest = rng.integers(0, n, (2, 50000))
    e = np.concatenate([hub_out, hub_in, run, rest], axis=1)
    perm = rng.permutation(e.shape[1])
    src, dst = e[0][perm], e[1][perm]
    w = rng.standard_normal(len(src)) * 1e3
    check(ctx, oracle, n, src, dst, w)
    check(ctx, oracle, n, src, dst)


def test_empty_and_tiny(ctx, oracle):
    g = check(ctx, oracle, 5, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    assert g.nnz == 0 and list(g.forward()[0]) == [0] * 6
    check(ctx, oracle, 1, [0], [0], [2.5])
    check(ctx, oracle, 7, [6, 6, 0], [6, 0, 6])


def test_errors_name_the_first_bad_edge(ctx, oracle):
    cases = [
        (4, [0, 1, 9, 0], [1, 2, 0, 1], [0, 0, 5, 7], 2, "edge_outside_id_space", "edge 2 "),
        (4, [0, 1, 2, 9], [1, 2, 0, 1], [0, 3, 0, 0], 2, "relation_outside_dictionary", "edge 1 "),
        (2, [0, 1], [1, 2], None, 1, "edge_outside_id_space", "edge 1 "),
    ]
    for n, src, dst, rel, nrel, name, where in cases:
        with pytest.raises(W.WarrenError) as e:
            W.Graph.build(ctx, n, src, dst, relations=rel, relation_count=nrel)
        assert e.value.name == name and where in str(e.value)
        with pytest.raises(oracle.OracleError) as o:
            oracle.csr_from_edge_list(n, src, dst, relations=rel, relation_count=nrel)
        assert o.value.code == e.value.code
    # The failed builds leave the context usable.
    check(ctx, oracle, 3, [0, 1], [1, 2])


def test_upload_transposes_on_the_device(ctx, oracle):
    rng = np.random.default_rng(11)
    n = 20000
    ref, _ = oracle.csr_from_edge_list(n, rng.integers(0, n, 300000), rng.integers(0, n, 300000),
                                       rng.standard_normal(300000))
    g = W.Graph(ctx, n, ref.offsets, ref.targets, ref.weights)
    rref = ref.transpose()
    roff, rtgt, rw = g.reverse()
    assert np.array_equal(roff, rref.offsets) and np.array_equal(rtgt, rref.targets)
    assert same(rw, rref.weights)
    # The R-MAT generator's graph (skewed degrees, long rows both ways).
    off, tgt = W.rmat(16, 16, seed=3)
    csr = oracle.Csr(len(off) - 1, off, tgt)
    g = W.Graph(ctx, csr.n, off, tgt)
    rref = csr.transpose()
    roff, rtgt, _ = g.reverse()
    assert np.array_equal(roff, rref.offsets) and np.array_equal(rtgt, rref.targets)


def test_weight_updates_land_in_both_orientations(ctx, oracle):
    rng = np.random.default_rng(12)
    off, tgt = W.rmat(14, 16, seed=5)
    n, m = len(off) - 1, int(off[-1])
    w = rng.random(m)
    g = W.Graph(ctx, n, off, tgt, w)
    edges = rng.choice(m, 500, replace=False)
    vals = rng.random(500) + 2.0
    g.update_weights(edges, vals)
    w2 = w.copy()
    w2[edges] = vals
    rref = oracle.Csr(n, off, tgt, w2).transpose()
    _, _, fw = g.forward()
    _, _, rw = g.reverse()
    assert same(fw, w2) and same(rw, rref.weights)
