"""Relations on the device: per-relation CSR pairs from wg_graph_build,
relation selections (wg_graph_select) merged or per relation, and products
over them, against the oracle's relation CSRs and selections, bit for bit."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu

# compute/product/tests.rs:32-38: relation 0: 0 -> 1 (2), 0 -> 2 (3),
# 1 -> 2 (4), 2 -> 0 (5); relation 1: 0 -> 1 (10). Node 3 is isolated.
EX = [(0, 0, 1, 2.0), (0, 0, 2, 3.0), (0, 1, 2, 4.0), (0, 2, 0, 5.0), (1, 0, 1, 10.0)]
X4 = np.array([1.0, 2.0, 3.0, 4.0])


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def same(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def spmv(ctx, g, x, semiring="plus_times", **kw):
    xv = W.Vector(ctx, x)
    yv = W.Vector(ctx, np.zeros_like(x))
    W.spmv(ctx, g, W.Product(semiring, **kw), xv, yv)
    return yv.numpy()


def example(ctx):
    rel, src, dst, w = (np.array(c) for c in zip(*EX))
    return W.Graph.build(ctx, 4, src, dst, w, relations=rel, relation_count=2)


def test_relation_known_answers(ctx):
    g = example(ctx)
    assert g.relation_count == 2
    # compute/product/tests.rs:60-69.
    assert list(spmv(ctx, g, X4, relations=[1])) == [20.0, 0.0, 0.0, 0.0]
    assert list(spmv(ctx, g, X4)) == [33.0, 12.0, 5.0, 0.0]
    assert list(spmv(ctx, g, X4, weighted=False)) == [5.0, 3.0, 1.0, 0.0]
    assert list(spmv(ctx, g, X4, transpose=True)) == [15.0, 12.0, 11.0, 0.0]
    # Per relation: the 0 -> 1 edge is an entry in each relation.
    assert list(spmv(ctx, g, X4, weighted=False, per_relation=True)) == [7.0, 3.0, 1.0, 0.0]
    assert spmv(ctx, g, X4, "min_plus", per_relation=True)[0] == 4.0
    assert spmv(ctx, g, X4, "min_plus")[0] == 6.0
    # The same selection named twice, in another order, is one selection.
    assert list(spmv(ctx, g, X4, relations=[1, 0, 1], per_relation=True, weighted=False)) == \
        [7.0, 3.0, 1.0, 0.0]
    # No relation: every row empty.
    assert list(spmv(ctx, g, X4, relations=[])) == [0.0, 0.0, 0.0, 0.0]


def test_unknown_relation(ctx):
    g = example(ctx)
    with pytest.raises(W.WarrenError) as e:
        spmv(ctx, g, X4, relations=[3, 2, 0])
    assert e.value.name == "unknown_relation" and "relation 2 " in str(e.value)
    off, tgt = W.rmat(6, 4, seed=1)
    single = W.Graph(ctx, 64, off, tgt)
    assert single.relation_count == 1 and single.select([0]) is single
    assert single.select(per_relation=True) is single
    with pytest.raises(W.WarrenError) as e:
        single.select([1])
    assert e.value.name == "unknown_relation"


def check_view(view, ref):
    off, tgt, w = view.forward()
    assert np.array_equal(off, ref.offsets) and np.array_equal(tgt, ref.targets)
    if ref.weights is not None:
        assert same(w, ref.weights)
    rref = ref.transpose()
    roff, rtgt, rw = view.reverse()
    assert np.array_equal(roff, rref.offsets) and np.array_equal(rtgt, rref.targets)
    if ref.weights is not None:
        assert same(rw, rref.weights)


@pytest.mark.parametrize("seed,n,m,R,weighted", [(0, 200, 3000, 3, True), (1, 5000, 200000, 4, True),
                                                 (2, 3000, 60000, 5, False)])
def test_selections_match_the_oracle(ctx, oracle, seed, n, m, R, weighted):
    rng = np.random.default_rng(seed)
    rel = rng.integers(0, R, m)
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    # Edges repeated across relations, so merging has work to do.
    dup = rng.integers(0, m, m // 4)
    src[rng.integers(0, m, m // 4)] = src[dup]
    w = rng.choice([0.1, 0.7, 1e-3, 3.0, -0.0, 2.5e7], m) if weighted else None
    print(f"relations n={n} m={m} R={R} weighted={weighted}", flush=True)
    g = W.Graph.build(ctx, n, src, dst, w, relations=rel, relation_count=R)
    csrs = oracle.relation_csrs(n, src, dst, w, rel, R)
    merged, folded = oracle.csr_from_edge_list(n, src, dst, w, relations=rel, relation_count=R)
    check_view(g, merged)
    assert g.deduplicated == folded
    for r in range(R):
        check_view(g.select([r]), csrs[r])
    for sel in ([0, 1], [R - 1, 0, 2], list(range(R)), None):
        for per in (False, True):
            check_view(g.select(sel, per_relation=per), oracle.select(csrs, sel, per_relation=per))
    # Products read the selections: min_plus over integer-valued f64 is
    # exact whatever the grouping.
    x = rng.integers(0, 50, n).astype(np.float64)
    for sel, per in (([0, 2], False), ([0, 2], True), (None, True)):
        ref = oracle.select(csrs, sel, per_relation=per)
        for transpose in (False, True):
            got = spmv(ctx, g, x, "min_plus", relations=sel, per_relation=per, transpose=transpose)
            want = oracle.spmv(ref, x, "min_plus", rev=ref.transpose() if transpose else None,
                               transpose=transpose)
            assert np.array_equal(got, want)


def test_selection_feeds_other_entry_points(ctx, oracle):
    """A selection is a graph: PageRank over two of three relations equals the
    oracle's PageRank over that selection."""
    rng = np.random.default_rng(9)
    n, m, R = 2000, 30000, 3
    rel = rng.integers(0, R, m)
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    g = W.Graph.build(ctx, n, src, dst, relations=rel, relation_count=R)
    view = g.select([0, 2])
    ranks, it = W.pagerank(ctx, view, tol=1e-10)
    ranks = ranks.numpy()
    csrs = oracle.relation_csrs(n, src, dst, None, rel, R)
    ref = oracle.select(csrs, [0, 2])
    want, it_ref = oracle.pagerank(ref, ref.transpose(), tol=1e-10)
    assert it == it_ref
    assert np.abs(ranks - want).sum() < 1e-12
