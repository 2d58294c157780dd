"""The device BFS over an Adjacency's CSRs (traversal/adjacency.rs:20-221):
directions Out / In / Both and relation subsets through wg_bfs,
wg_reachable_set, the cursor and wg_bfs_levels, against the oracle's
search over the same CSR lists: visits (order and vias) and stats equal."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def built(ctx, oracle, n, edges, R):
    rel = np.array([e[0] for e in edges], np.int64)
    src = np.array([e[1] for e in edges], np.int64)
    dst = np.array([e[2] for e in edges], np.int64)
    g = W.Graph.build(ctx, n, src, dst, relations=rel if R > 1 else None, relation_count=R)
    merged, _ = oracle.csr_from_edge_list(n, src, dst, relations=rel if R > 1 else None, relation_count=R)
    per = oracle.relation_csrs(n, src, dst, None, rel, R)
    return g, (merged, merged.transpose()), [(c, c.transpose()) for c in per]


def check(ctx, oracle, g, merged, per, n, starts, direction, relations, flt=None, max_depth=None,
          max_results=None):
    print(f"adjacency bfs: n={n} starts={list(starts)} {direction} relations={relations} "
          f"filter={flt is not None} max_depth={max_depth} max_results={max_results}", flush=True)
    rows, rrows = oracle.adjacency(merged, per, direction, relations)
    want, wst = oracle.bfs(rows, starts, node_filter=flt, max_depth=max_depth, max_results=max_results, n=n)
    got, gst = W.bfs(ctx, g, starts, direction=direction, relations=relations, node_filter=flt,
                     max_depth=max_depth, max_results=max_results)
    assert got == want
    assert gst == wst
    cur = W.BfsCursor(ctx, g, starts, direction=direction, relations=relations, node_filter=flt,
                      max_depth=max_depth, max_results=max_results)
    assert [v for level in cur.levels() for v in level] == want
    assert cur.stats() == wst
    cur.close()
    if flt is None and max_results is None:
        want, wst = oracle.bfs(rows, starts, pull_rev=rrows, max_depth=max_depth, n=n)
        got, gst = W.reachable_set(ctx, g, starts, direction=direction, relations=relations, max_depth=max_depth)
        assert list(got) == [v for v, _, _ in want]
        assert gst == wst
        cur = W.BfsCursor(ctx, g, starts, direction=direction, relations=relations, max_depth=max_depth,
                          reachable=True)
        assert [v for level in cur.levels() for v in level] == want
        cur.close()
    # The device-resident search: the push's depths and vias.
    d, via, _ = W.bfs_levels(ctx, g, starts, direction=direction, relations=relations, node_filter=flt,
                             max_depth=max_depth)
    push, _ = oracle.bfs(rows, starts, node_filter=flt, max_depth=max_depth, n=n)
    dd, vv = d.numpy(), via.numpy()
    wd = np.full(n, 2**32 - 1, np.uint32)
    wv = np.full(n, 2**32 - 1, np.uint32)
    for v, dep, u in push:
        wd[v] = dep
        if u is not None:
            wv[v] = u
    assert np.array_equal(dd, wd) and np.array_equal(vv, wv)


def test_both_unites_forward_and_reverse_rows(ctx, oracle):
    g, merged, per = built(ctx, oracle, 3, [(0, 0, 1), (0, 1, 2)], 1)
    got, st = W.bfs(ctx, g, [1], direction="both")
    assert got == [(1, 0, None), (2, 1, 1), (0, 1, 1)]
    assert st["edges_traversed"] == 4 and st["levels_pulled"] == 0
    check(ctx, oracle, g, merged, per, 3, [1], "both", None)


@pytest.mark.parametrize("seed", range(8))
def test_random_relation_graphs(ctx, oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 300))
    R = int(rng.integers(1, 5))
    m = int(rng.integers(n, 6 * n))
    edges = sorted({(int(rng.integers(0, R)), int(rng.integers(0, n)), int(rng.integers(0, n)))
                    for _ in range(m)})
    g, merged, per = built(ctx, oracle, n, edges, R)
    for trial in range(6):
        direction = ["out", "in", "both"][trial % 3]
        relations = [None, sorted(set(rng.integers(0, R, 2).tolist())), list(range(R)), [R - 1]][trial % 4]
        starts = [int(x) for x in rng.integers(0, n, int(rng.integers(1, 4)))]
        flt = (rng.random(n) < 0.85).astype(np.uint8) if trial == 5 else None
        max_depth = 2 if trial == 4 else None
        check(ctx, oracle, g, merged, per, n, starts, direction, relations, flt=flt, max_depth=max_depth)
        # Truncated searches count the entries a push scans up to the cut.
        for k in (1, 3, 17):
            check(ctx, oracle, g, merged, per, n, starts, direction, relations, max_results=k)


def test_rmat_relations_both_ways(ctx, oracle):
    """Wide levels pulled over several CSRs (first-hit and smallest-via
    passes), split rows included."""
    off, tgt = W.rmat(14, 16, seed=11)
    n = len(off) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off.astype(np.int64)))
    rel = (np.random.default_rng(3).random(len(tgt)) < 0.3).astype(np.int64)
    edges = list(zip(rel.tolist(), src.tolist(), tgt.astype(np.int64).tolist()))
    g, merged, per = built(ctx, oracle, n, edges, 2)
    deg = np.diff(off.astype(np.int64))
    hub = int(np.argmax(deg))
    for direction in ("out", "in", "both"):
        for relations in (None, [0, 1], [1]):
            check(ctx, oracle, g, merged, per, n, [hub, 5], direction, relations)
    check(ctx, oracle, g, merged, per, n, [hub], "both", [0, 1], max_results=5000)


def test_errors(ctx, oracle):
    g, _, _ = built(ctx, oracle, 4, [(0, 0, 1), (1, 1, 2)], 2)
    with pytest.raises(W.WarrenError) as e:
        W.bfs(ctx, g, [0], relations=[1, 5, 3])
    assert e.value.name == "unknown_relation" and "relation 3" in str(e.value)
    assert W.bfs(ctx, g, [0, 3], relations=[])[0] == [(0, 0, None), (3, 0, None)]
