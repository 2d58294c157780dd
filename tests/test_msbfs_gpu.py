"""Multi-source BFS (up to 512 sources as bitsets) against the oracle's
per-source BFS depths: every (source, node) distance bit-exact, and the
reach counts / farness sums that SourceBatch reports
(compute/centrality/multi_source.rs:92-193) follow from them."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu

UNREACHED = 2**32 - 1


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def hub_csr(oracle, n, m, seed):
    """Random edges plus rows and columns longer than one piece (1024)."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([rng.integers(0, n, m), np.zeros(n - 1, np.int64), np.arange(1, n)])
    dst = np.concatenate([rng.integers(0, n, m), np.arange(1, n), np.full(n - 1, 3)])
    keep = rng.random(len(src)) < 0.9
    return oracle.csr_from_edges(n, src[keep], dst[keep])


def check(ctx, oracle, ref, sources, direction="out", max_depth=None):
    g = W.Graph(ctx, ref.n, ref.offsets, ref.targets)
    adj = ref if direction == "out" else ref.transpose()
    want = oracle.bfs_depths(adj, sources)  # ns x n
    if max_depth is not None:
        want = np.where(want <= max_depth, want, UNREACHED).astype(np.uint32)
    dv, rv, fv, levels = W.msbfs(ctx, g, sources, direction=direction, max_depth=max_depth)
    got = dv.numpy().reshape(ref.n, len(sources)).T
    assert np.array_equal(got, want)
    reached = want != UNREACHED
    assert np.array_equal(rv.numpy(), reached.sum(axis=0))
    assert np.array_equal(fv.numpy(), np.where(reached, want, 0).astype(np.int64).sum(axis=0))
    return levels


@pytest.mark.parametrize("ns", [1, 64, 100, 512])
def test_random_graph_matches_per_source_bfs(ctx, oracle, ns):
    ref = hub_csr(oracle, 3000, 9000, ns)
    sources = np.random.default_rng(ns).integers(0, ref.n, ns)
    check(ctx, oracle, ref, sources)


def test_direction_in_repeats_and_max_depth(ctx, oracle):
    ref = hub_csr(oracle, 2500, 6000, 9)
    sources = np.random.default_rng(3).integers(0, ref.n, 70)
    sources[5] = sources[6]  # a repeated source is its own column
    check(ctx, oracle, ref, sources, direction="in")
    check(ctx, oracle, ref, sources, max_depth=2)
    assert check(ctx, oracle, ref, sources[:3], max_depth=0) == 0


def test_errors(ctx, oracle):
    ref = hub_csr(oracle, 100, 300, 1)
    g = W.Graph(ctx, ref.n, ref.offsets, ref.targets)
    with pytest.raises(W.WarrenError) as e:
        W.msbfs(ctx, g, [100])
    assert e.value.name == "node_outside_id_space"
    with pytest.raises(W.WarrenError) as e:
        W.msbfs(ctx, g, list(range(513)))
    assert e.value.name == "invalid_argument"


def test_rmat_scale16_512_sources(ctx, oracle):
    off, tgt = W.rmat(16, 16, seed=7)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    deg = np.diff(off.astype(np.int64))
    sources = np.random.default_rng(5).choice(np.flatnonzero(deg > 0), 512, replace=False)
    check(ctx, oracle, ref, sources)
