"""The device BFS cursor (wg_bfs_cursor_*, traversal/bfs_cursor.rs): levels
handed out one call at a time, each sorted and written on the device; their
concatenation is the search's visit list, checked against wg_bfs and the
oracle's level-synchronous search, with the reference's cursor answers."""
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
    # bfs_cursor.rs:149-152: 0 -> 1 -> 2 -> 3, 0 -> 4 -> 3, 2 -> 2, 5 isolated.
    e = [(0, 1), (1, 2), (2, 3), (0, 4), (4, 3), (2, 2)]
    return oracle.csr_from_edges(6, [a for a, _ in e], [b for _, b in e])


def test_a_cursor_hands_out_the_bfs_one_level_at_a_time(ctx, oracle):
    # bfs_cursor.rs:184-195
    g = up(ctx, example(oracle))
    cur = W.BfsCursor(ctx, g, [0])
    levels = cur.levels()
    assert [len(lv) for lv in levels] == [1, 2, 2]
    want, wst = W.bfs(ctx, g, [0])
    assert sum(levels, []) == want
    assert cur.stats() == wst
    # Nothing more once finished.
    assert cur.next_level() == ([], False)


def test_filter_and_limits_as_bfs(ctx, oracle):
    # bfs_cursor.rs:197-206: filter {1, 2, 3}, max_depth 2, max_results 2,
    # starts [0, 0].
    g = up(ctx, example(oracle))
    flt = np.zeros(6, np.uint8)
    flt[[1, 2, 3]] = 1
    cur = W.BfsCursor(ctx, g, [0, 0], node_filter=flt, max_depth=2, max_results=2)
    found = sum(cur.levels(), [])
    want, _ = W.bfs(ctx, g, [0, 0], node_filter=flt, max_depth=2, max_results=2)
    assert found == want
    assert cur.stats()["truncated"]


def test_arguments_are_checked_up_front(ctx, oracle):
    # bfs_cursor.rs:218-243
    g = up(ctx, example(oracle))
    with pytest.raises(W.WarrenError) as e:
        W.BfsCursor(ctx, g, [6])
    assert e.value.name == "node_outside_id_space"
    with pytest.raises(W.WarrenError) as e:
        W.BfsCursor(ctx, g, [0], reachable=True, max_results=3)
    assert e.value.name == "invalid_argument"


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5])
def test_a_cursor_yields_exactly_the_bfs(ctx, oracle, seed):
    """bfs_cursor.rs:263-300: over random graphs, starts, filters, depths and
    result limits, in both directions."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 3000))
    m = int(rng.integers(n, 8 * n))
    ref = oracle.csr_from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m))
    rev = ref.transpose()
    g = up(ctx, ref)
    for trial in range(6):
        starts = list(rng.integers(0, n, int(rng.integers(0, 4))))
        flt = (rng.random(n) < 0.85).astype(np.uint8) if trial % 2 else None
        max_depth = None if trial % 3 == 0 else int(rng.integers(0, 5))
        max_results = None if trial % 2 == 0 else int(rng.integers(0, 40))
        direction = "in" if trial % 4 >= 2 else "out"
        print(f"cursor seed={seed} trial={trial} n={n} m={m} {direction}", flush=True)
        cur = W.BfsCursor(ctx, g, starts, direction=direction, node_filter=flt, max_depth=max_depth,
                          max_results=max_results)
        found = sum(cur.levels(), [])
        want, wst = W.bfs(ctx, g, starts, direction=direction, node_filter=flt, max_depth=max_depth,
                          max_results=max_results)
        assert found == want
        assert cur.stats() == wst
        ow, ost = oracle.bfs(rev if direction == "in" else ref, starts, node_filter=flt,
                             max_depth=max_depth, max_results=max_results)
        assert found == ow and cur.stats()["truncated"] == ost["truncated"]
        # Every level's visits share one depth, the levels in order.
        depths = [v[1] for v in found]
        assert depths == sorted(depths)


def test_a_reachable_set_cursor_yields_the_reachable_set(ctx, oracle):
    off, tgt = W.rmat(16, 16, seed=2)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    g = up(ctx, ref)
    starts = [3, 77, 1000]
    cur = W.BfsCursor(ctx, g, starts, reachable=True)
    found = [v[0] for v in sum(cur.levels(), [])]
    want, wst = W.reachable_set(ctx, g, starts)
    assert found == list(want)
    assert cur.stats() == wst
    assert wst["levels_pulled"] > 0


def test_cursors_interleave_on_one_graph(ctx, oracle):
    """Each cursor keeps its own workspace: two open cursors and a whole
    search between their levels do not disturb one another."""
    off, tgt = W.rmat(15, 16, seed=6)
    ref = oracle.Csr(len(off) - 1, off, tgt)
    g = up(ctx, ref)
    a = W.BfsCursor(ctx, g, [1])
    b = W.BfsCursor(ctx, g, [2], direction="in")
    la, lb = [], []
    more_a = more_b = True
    while more_a or more_b:
        if more_a:
            lv, more_a = a.next_level()
            la += lv
        W.bfs(ctx, g, [5], max_depth=2)
        if more_b:
            lv, more_b = b.next_level()
            lb += lv
    assert la == oracle.bfs(ref, [1])[0]
    assert lb == oracle.bfs(ref.transpose(), [2])[0]
