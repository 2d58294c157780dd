"""Batched personalised PageRank: every column against the oracle's FP64
PageRank with that seed vector (same iteration count, <= 1e-9 L1) and
against the plain GPU run."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu

L1_TOL = 1e-9


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def graph(oracle, n, m, seed, weighted):
    """Random edges plus a column longer than one piece (4096 entries)."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([rng.integers(0, n, m), np.arange(n)])
    dst = np.concatenate([rng.integers(0, n, m), np.full(n, 3)])
    w = rng.lognormal(0.0, 1.0, len(src)) if weighted else None
    return oracle.csr_from_edges(n, src, dst, w)


def seed_sets(rng, n, R):
    """Sizes from P(s) ~ s^-2 on 1..64, members uniform without replacement."""
    sizes = np.arange(1, 65)
    p = sizes ** -2.0
    p /= p.sum()
    return [rng.choice(n, rng.choice(sizes, p=p), replace=False) for _ in range(R)]


def dense(n, nodes, weights=None):
    v = np.zeros(n)
    np.add.at(v, nodes, 1.0 if weights is None else weights)
    return v


@pytest.mark.parametrize("R,weighted", [(1, False), (7, True), (64, False), (256, True)])
def test_columns_match_plain_pagerank(ctx, oracle, R, weighted):
    ref = graph(oracle, 6000, 30000, R, weighted)
    g = W.Graph(ctx, ref.n, ref.offsets, ref.targets, ref.weights)
    rng = np.random.default_rng(R + 1)
    sets = seed_sets(rng, ref.n, R)
    ranks, its = W.pagerank_batch(ctx, g, sets, tol=1e-9, max_iter=100)
    got = ranks.numpy().reshape(ref.n, R)
    for c in range(0, R, max(1, R // 12)):
        want, wit = oracle.pagerank(ref, tol=1e-9, max_iter=100, seed=dense(ref.n, sets[c]))
        assert its[c] == wit, (c, its[c], wit)
        assert np.abs(got[:, c] - want).sum() < L1_TOL, c
    # The plain GPU run of one column, to the same bar.
    v, it = W.pagerank(ctx, g, tol=1e-9, max_iter=100, seed=dense(ref.n, sets[0]))
    assert it == its[0]
    assert np.abs(v.numpy() - got[:, 0]).sum() < L1_TOL


def test_weights_repeats_tombstones_and_errors(ctx, oracle):
    ref = graph(oracle, 5000, 20000, 3, False)
    g = W.Graph(ctx, ref.n, ref.offsets, ref.targets)
    rng = np.random.default_rng(4)
    sets = [np.array([1, 2, 2, 9]), np.array([7]), rng.choice(ref.n, 30, replace=False)]
    wts = [np.array([0.5, 1.0, 2.0, 3.0]), np.array([4.0]), rng.random(30)]
    tomb = (rng.random(ref.n) < 0.02).astype(np.uint8)
    tomb[[1, 2, 7]] = 0
    ranks, its = W.pagerank_batch(ctx, g, sets, seed_weights=wts, tol=1e-10, max_iter=60,
                                  tombstones=tomb)
    got = ranks.numpy().reshape(ref.n, 3)
    for c in range(3):
        want, wit = oracle.pagerank(ref, tol=1e-10, max_iter=60, tombstones=tomb,
                                    seed=dense(ref.n, sets[c], wts[c]))
        assert its[c] == wit
        assert np.abs(got[:, c] - want).sum() < L1_TOL
        assert np.all(got[tomb == 1, c] == 0.0)
    with pytest.raises(W.WarrenError) as e:
        W.pagerank_batch(ctx, g, [np.array([5])], seed_weights=[np.array([0.0])])
    assert e.value.name == "division_by_zero"
    with pytest.raises(W.WarrenError) as e:
        W.pagerank_batch(ctx, g, [np.array([ref.n])])
    assert e.value.name == "node_outside_id_space"
    ranks, its = W.pagerank_batch(ctx, g, sets[:2], max_iter=0)
    assert list(its) == [0, 0]
    assert np.all(ranks.numpy() == 1.0 / ref.n)


def test_rmat_scale16_256_seed_sets(ctx, oracle):
    off, tgt = W.rmat(16, 16, seed=3)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    g = W.Graph(ctx, ref.n, off, tgt)
    rng = np.random.default_rng(16)
    deg = np.diff(off.astype(np.int64))
    live = np.flatnonzero(deg > 0)
    sets = [live[s] for s in seed_sets(rng, len(live), 256)]
    ranks, its = W.pagerank_batch(ctx, g, sets, tol=1e-10, max_iter=100)
    got = ranks.numpy().reshape(ref.n, 256)
    for c in [0, 1, 77, 200, 255]:
        want, wit = oracle.pagerank(ref, tol=1e-10, max_iter=100, seed=dense(ref.n, sets[c]))
        assert its[c] == wit
        assert np.abs(got[:, c] - want).sum() < L1_TOL


@pytest.mark.parametrize("weighted", [False, True])
def test_graph_replays_equal_eager_launches(ctx, weighted):
    """Each iteration replays a captured HIP graph (one per message parity
    and active-column set, re-captured after a repack), with the host's
    own copies, repacks and synchronisations between replays; every
    column equals the one-by-one launches' bit for bit, iteration counts
    too, over runs whose columns converge at different iterations."""
    off, tgt = W.rmat(14, 16, seed=3)
    n = len(off) - 1
    w = W.edge_weights(off, tgt, "lognormal", seed=2, p0=0.0, p1=1.0) if weighted else None
    g = W.Graph(ctx, n, off, tgt, w)
    rng = np.random.default_rng(8)
    sets = seed_sets(rng, n, 40)
    try:
        ctx.set_graph_launch(False)
        want, wit = W.pagerank_batch(ctx, g, sets, tol=1e-10, max_iter=80)
        want = want.numpy()
        ctx.set_graph_launch(True)
        for _ in range(2):
            got, it = W.pagerank_batch(ctx, g, sets, tol=1e-10, max_iter=80)
            assert list(it) == list(wit)
            assert got.numpy().tobytes() == want.tobytes()
        assert len(set(wit)) > 1  # columns left at different iterations: repacks ran
    finally:
        ctx.set_graph_launch(True)
    g.close()


@pytest.mark.parametrize("max_iter", [1, 2, 60])
def test_first_iterations_tombstones_and_a_partial_tile(ctx, oracle, max_iter):
    """40 seed sets (three tiles of 16, the last partial) over a weighted
    graph with tombstones, stopped after the first iteration (every column
    from tile 0's gathered sums), the second (zero rows' ranks left alone) or
    run to convergence: every column against the oracle's PageRank with
    that seed vector."""
    ref = graph(oracle, 4000, 20000, 40, True)
    g = W.Graph(ctx, ref.n, ref.offsets, ref.targets, ref.weights)
    rng = np.random.default_rng(41)
    sets = seed_sets(rng, ref.n, 40)
    tomb = (rng.random(ref.n) < 0.03).astype(np.uint8)
    for s in sets:
        tomb[s] = 0
    ranks, its = W.pagerank_batch(ctx, g, sets, tol=1e-10, max_iter=max_iter, tombstones=tomb)
    got = ranks.numpy().reshape(ref.n, 40)
    for c in range(0, 40, 3):
        want, wit = oracle.pagerank(ref, tol=1e-10, max_iter=max_iter, tombstones=tomb,
                                    seed=dense(ref.n, sets[c]))
        assert its[c] == wit, (c, its[c], wit)
        assert np.abs(got[:, c] - want).sum() < L1_TOL, c
        assert np.all(got[tomb == 1, c] == 0.0)
