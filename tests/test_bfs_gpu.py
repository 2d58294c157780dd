"""Parity of the HIP BFS with the oracle's level-synchronous search
(traversal/level.rs, bfs.rs, reachable.rs): visit lists equal element for
element (node, depth, via and order), bit-exact."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def up(ctx, csr):
    return W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)


def example(oracle):
    e = [(0, 1), (1, 2), (2, 3), (0, 4), (4, 3), (2, 2)]
    return oracle.csr_from_edges(6, [a for a, _ in e], [b for _, b in e])


def v(node, depth, via=None):
    return (node, depth, via)


def test_known_answers(ctx, oracle):
    """traversal/bfs.rs:160-235 through the device."""
    g = up(ctx, example(oracle))
    got, st = W.bfs(ctx, g, [0])
    assert got == [v(0, 0), v(1, 1, 0), v(4, 1, 0), v(2, 2, 1), v(3, 2, 4)]
    assert st["nodes_visited"] == 5 and not st["truncated"]
    assert W.bfs(ctx, g, [2, 4, 2])[0] == [v(2, 0), v(4, 0), v(3, 1, 2)]
    assert W.bfs(ctx, g, [4, 1])[0] == [v(4, 0), v(1, 0), v(2, 1, 1), v(3, 1, 4)]
    got, _ = W.bfs(ctx, g, [3], direction="in")
    assert [(a, b) for a, b, _ in got] == [(3, 0), (2, 1), (4, 1), (1, 2), (0, 2)]
    got, st = W.bfs(ctx, g, [0], max_depth=1)
    assert len(got) == 3 and not st["truncated"]
    assert W.bfs(ctx, g, [0], max_depth=0)[0] == [v(0, 0)]
    got, st = W.bfs(ctx, g, [0], max_results=2)
    assert got == [v(0, 0), v(1, 1, 0)] and st["truncated"]
    got, st = W.bfs(ctx, g, [0], max_results=5)
    assert len(got) == 5 and not st["truncated"]
    flt = np.zeros(6, np.uint8)
    flt[[1, 2, 3]] = 1
    got, _ = W.bfs(ctx, g, [0], node_filter=flt)
    assert [(a, b) for a, b, _ in got] == [(0, 0), (1, 1), (2, 2), (3, 3)]
    with pytest.raises(W.WarrenError) as e:
        W.bfs(ctx, g, [6])
    assert e.value.name == "node_outside_id_space"


def random_csr(oracle, n, m, seed):
    rng = np.random.default_rng(seed)
    return oracle.csr_from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_graphs_match_the_oracle(ctx, oracle, seed):
    ref = random_csr(oracle, 3000, 12000 + 4000 * seed, seed)
    g = up(ctx, ref)
    rev = ref.transpose()
    rng = np.random.default_rng(seed + 100)
    starts = list(rng.integers(0, ref.n, 3)) + [int(rng.integers(0, ref.n))] * 2
    want, _ = oracle.bfs(ref, starts)
    got, _ = W.bfs(ctx, g, starts)
    assert got == want
    # reachable_set: same decisions as the reference's Beamer switch, so
    # the order and the traversal counters agree too.
    want, wst = oracle.bfs(ref, starts, pull_rev=rev)
    got, gst = W.reachable_set(ctx, g, starts)
    assert list(got) == [a for a, _, _ in want]
    assert gst["edges_traversed"] == wst["edges_traversed"]
    assert gst["levels_pulled"] == wst["levels_pulled"]
    # direction in, filter, limits.
    flt = (rng.random(ref.n) < 0.8).astype(np.uint8)
    want, _ = oracle.bfs(rev, starts, node_filter=flt, max_depth=4)
    got, _ = W.bfs(ctx, g, starts, direction="in", node_filter=flt, max_depth=4)
    assert got == want
    for k in (1, 7, 100, len(want)):
        w2, ws = oracle.bfs(rev, starts, node_filter=flt, max_depth=4, max_results=k)
        g2, gs = W.bfs(ctx, g, starts, direction="in", node_filter=flt, max_depth=4, max_results=k)
        assert g2 == w2 and gs["truncated"] == ws["truncated"]


def test_rmat_scale16_depth3_from_16_seeds(ctx, oracle):
    """The CPU-runnable workload's BFS leg: R-MAT scale 16, depth 3, 16 seeds."""
    off, tgt = W.rmat(16, 16, seed=1)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    g = up(ctx, ref)
    seeds = np.random.default_rng(16).integers(0, ref.n, 16)
    for s in seeds:
        want, wst = oracle.bfs(ref, [s], max_depth=3)
        got, gst = W.bfs(ctx, g, [s], max_depth=3)
        assert got == want
        assert gst["nodes_visited"] == wst["nodes_visited"]


def test_rmat_scale18_depths_and_vias(ctx, oracle):
    off, tgt = W.rmat(18, 16, seed=4)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    g = up(ctx, ref)
    deg = np.diff(off.astype(np.int64))
    sources = np.flatnonzero(deg > 0)[:: max(1, len(np.flatnonzero(deg > 0)) // 5)][:5]
    for s in sources:
        d, via, st = W.bfs_levels(ctx, g, [s])
        d = d.numpy()
        want = oracle.bfs_depths(ref, [s])[0]
        assert np.array_equal(d, want)
        assert st["levels_pulled"] > 0
        # via: the smallest predecessor one level up.
        via = via.numpy()
        reached = np.flatnonzero((d != 2**32 - 1) & (d > 0))
        rev_off, rev_tgt, _ = g.reverse()
        for node in reached[:: max(1, len(reached) // 2000)]:
            preds = rev_tgt[rev_off[node]:rev_off[node + 1]]
            preds = preds[d[preds] == d[node] - 1]
            assert via[node] == preds.min()


def test_hub_rows_pushed_in_pieces(ctx, oracle):
    """Rows longer than a 2048-entry block are pushed a piece per workgroup:
    from the largest hub, a node beside it and a far node, the visits (vias
    and order included) equal the oracle's, pushed only or pulled too."""
    off, tgt = W.rmat(17, 16, seed=9)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    g = up(ctx, ref)
    deg = np.diff(off.astype(np.int64))
    assert deg.max() > 4 * 2048
    hub = int(np.argmax(deg))
    beside = int(tgt[off[hub]])
    far = int(np.flatnonzero(deg == 1)[0])
    for s in (hub, beside, far, [hub, far]):
        print(f"hub rows: starts {s}", flush=True)
        want, wst = oracle.bfs(ref, np.atleast_1d(s))
        got, gst = W.bfs(ctx, g, np.atleast_1d(s))
        assert got == want
        assert gst["nodes_visited"] == wst["nodes_visited"]
        d1, v1, _ = W.bfs_levels(ctx, g, np.atleast_1d(s), pulling=False)
        d1, v1 = d1.numpy(), v1.numpy()
        d2, v2, _ = W.bfs_levels(ctx, g, np.atleast_1d(s))
        assert np.array_equal(d1, d2.numpy()) and np.array_equal(v1, v2.numpy())


def test_small_levels_with_filters_and_depth_limits(ctx, oracle):
    """Searches whose push levels fit one workgroup's level loop (R-MAT scale
    16: frontier rows under n / 4 and 16,384 entries, run several levels to
    a launch), with node filters, depth limits and several starts: depths
    and vias equal the oracle's push, pulled or not, and a pushed-only
    search counts its visits and scanned entries as the oracle's push."""
    off, tgt = W.rmat(16, 16, seed=7)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    n = ref.n
    g = up(ctx, ref)
    rng = np.random.default_rng(21)
    deg = np.diff(off.astype(np.int64))
    live = np.flatnonzero(deg > 0)
    flt = (rng.random(n) < 0.85).astype(np.uint8)
    cases = [[int(s)] for s in rng.choice(live, 3, replace=False)]
    cases.append([int(s) for s in rng.choice(live, 5, replace=False)])
    for starts in cases:
        flt[starts] = 1
        for f in (None, flt):
            for md in (None, 1, 3):
                print(f"small levels: starts {starts} filter={f is not None} max_depth={md}", flush=True)
                push, wst = oracle.bfs(ref, starts, node_filter=f, max_depth=md)
                wd = np.full(n, 2**32 - 1, np.uint32)
                wv = np.full(n, 2**32 - 1, np.uint32)
                for v, dep, u in push:
                    wd[v] = dep
                    if u is not None:
                        wv[v] = u
                for pulling in (False, True):
                    d, via, st = W.bfs_levels(ctx, g, starts, node_filter=f, max_depth=md, pulling=pulling)
                    assert np.array_equal(d.numpy(), wd) and np.array_equal(via.numpy(), wv), pulling
                    if not pulling:
                        assert st["nodes_visited"] == wst["nodes_visited"]
                        assert st["edges_traversed"] == wst["edges_traversed"]
                    d.close()
                    via.close()
