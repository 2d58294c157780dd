"""The oracle's BFS over an Adjacency's CSR lists (traversal/adjacency.rs:
20-221, level.rs:88-337): directions Out / In / Both and relation subsets,
pinned by a naive BFS over the edge sets (depths), by the definitions of a
push's order and via and of a pull's via, and by the reference's Both known
answer (adjacency.rs:400-406)."""
import numpy as np
import pytest


def relation_graph(oracle, n, edges, relation_count):
    """Edges (relation, source, target): the merged pair and each
    relation's pair, as a built graph holds them."""
    rel = np.array([e[0] for e in edges], np.int64)
    src = np.array([e[1] for e in edges], np.int64)
    dst = np.array([e[2] for e in edges], np.int64)
    merged, _ = oracle.csr_from_edge_list(n, src, dst, relations=rel if relation_count > 1 else None,
                                          relation_count=relation_count)
    per = oracle.relation_csrs(n, src, dst, None, rel, relation_count)
    return (merged, merged.transpose()), [(c, c.transpose()) for c in per]


def row_lists(edges, direction, relations, relation_count):
    """Per CSR read (in order), a dict node -> sorted row, built from the
    edge set: the definition the oracle's CSR lists must follow."""
    pairs = [None] if relations is None else sorted(set(relations))
    out = []
    for r in pairs:
        sel = [(s, t) for (q, s, t) in edges if r is None or q == r]
        fwd, rev = {}, {}
        for s, t in set(sel):
            fwd.setdefault(s, set()).add(t)
            rev.setdefault(t, set()).add(s)
        for d in {"out": ["f"], "in": ["r"], "both": ["f", "r"]}[direction]:
            m = fwd if d == "f" else rev
            out.append({k: sorted(v) for k, v in m.items()})
    return out


def naive_depths(n, rows, starts, flt):
    depth = {}
    frontier = []
    for s in starts:
        if s not in depth:
            depth[s] = 0
            frontier.append(s)
    d = 0
    while frontier:
        d += 1
        nxt = []
        for u in frontier:
            for row in rows:
                for v in row.get(u, []):
                    if v not in depth and (flt is None or flt[v]):
                        depth[v] = d
                        nxt.append(v)
        frontier = nxt
    return depth


def random_case(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(6, 40))
    R = int(rng.integers(1, 4))
    m = int(rng.integers(0, 3 * n))
    edges = sorted({(int(rng.integers(0, R)), int(rng.integers(0, n)), int(rng.integers(0, n)))
                    for _ in range(m)})
    return rng, n, R, edges


def test_both_unites_forward_and_reverse_rows(oracle):
    """adjacency.rs:400-406: Both reads forward then reverse rows, so from 1
    the push reaches 2 (forward row) before 0 (reverse row)."""
    merged, per = relation_graph(oracle, 3, [(0, 0, 1), (0, 1, 2)], 1)
    rows, rrows = oracle.adjacency(merged, per, "both")
    got, st = oracle.bfs(rows, [1])
    assert got == [(1, 0, None), (2, 1, 1), (0, 1, 1)]
    assert st["edges_traversed"] == 2 + 1 + 1  # rows of 1, then of 2 and 0
    assert oracle.bfs(rows, [1], pull_rev=rrows)[0] in ([(1, 0, None), (2, 1, 1), (0, 1, 1)],
                                                         [(1, 0, None), (0, 1, 1), (2, 1, 1)])


@pytest.mark.parametrize("seed", range(40))
def test_push_depths_vias_and_order_follow_the_definitions(oracle, seed):
    rng, n, R, edges = random_case(oracle, seed)
    if not edges:
        return
    merged, per = relation_graph(oracle, n, edges, R)
    direction = ["out", "in", "both"][seed % 3]
    relations = None if seed % 2 == 0 else sorted(set(rng.integers(0, R, 2).tolist()))
    rows, rrows = oracle.adjacency(merged, per, direction, relations)
    defs = row_lists(edges, direction, relations, R)
    starts = [int(x) for x in rng.integers(0, n, int(rng.integers(1, 3)))]
    flt = (rng.random(n) < 0.8).astype(np.uint8) if seed % 4 == 3 else None
    got, st = oracle.bfs(rows, starts, node_filter=flt, n=n)
    depth = naive_depths(n, defs, starts, flt)
    assert {v: d for v, d, _ in got} == depth
    assert st["levels_pulled"] == 0
    # Each level by (via, first row set of the via holding the node, node);
    # via the smallest node one level up with an edge to it.
    by_level = {}
    for v, d, via in got:
        by_level.setdefault(d, []).append((v, via))
    for d, level in by_level.items():
        if d == 0:
            continue
        keys = []
        for v, via in level:
            preds = [u for u in depth if depth[u] == d - 1 and any(v in row.get(u, []) for row in defs)]
            assert via == min(preds)
            j = min(k for k, row in enumerate(defs) if v in row.get(via, []))
            keys.append((via, j, v))
        assert keys == sorted(keys)
    # Every frontier row entry is scanned once per level.
    scanned = sum(len(row.get(u, [])) for u in depth for row in defs)
    assert st["edges_traversed"] == scanned


@pytest.mark.parametrize("seed", range(40))
def test_pulled_levels_take_the_first_row_set_with_a_frontier_node(oracle, seed):
    rng, n, R, edges = random_case(oracle, 1000 + seed)
    if not edges:
        return
    merged, per = relation_graph(oracle, n, edges, R)
    direction = ["out", "in", "both"][seed % 3]
    relations = None if seed % 2 else sorted(set(rng.integers(0, R, 3).tolist()))
    rows, rrows = oracle.adjacency(merged, per, direction, relations)
    defs = row_lists(edges, direction, relations, R)
    back = {"out": "in", "in": "out", "both": "both"}[direction]
    rdefs = row_lists(edges, back, relations, R)
    starts = [int(x) for x in rng.integers(0, n, 2)]
    got, st = oracle.bfs(rows, starts, pull_rev=rrows, n=n)
    push, _ = oracle.bfs(rows, starts, n=n)
    depth = naive_depths(n, defs, starts, None)
    assert {v: d for v, d, _ in got} == depth
    assert sorted(v for v, _, _ in got) == sorted(v for v, _, _ in push)
    # A level is the push's, or pulled: ascending, each node's via the
    # smallest frontier node in the first reversed row set naming one.
    by_level = {}
    for v, d, via in got:
        by_level.setdefault(d, []).append((v, via))
    pulled = 0
    for d, level in by_level.items():
        if d == 0:
            continue
        pushed = [(v, via) for v, dd, via in push if dd == d]
        if level == pushed:
            continue
        pulled += 1
        assert [v for v, _ in level] == sorted(v for v, _ in level)
        for v, via in level:
            for row in rdefs:
                hits = [u for u in row.get(v, []) if depth.get(u) == d - 1]
                if hits:
                    assert via == hits[0]
                    break
    assert pulled <= st["levels_pulled"]


def test_an_empty_relation_subset_reaches_only_the_starts(oracle):
    merged, per = relation_graph(oracle, 4, [(0, 0, 1), (1, 1, 2)], 2)
    rows, rrows = oracle.adjacency(merged, per, "out", [])
    assert rows == [] and rrows == []
    got, st = oracle.bfs(rows, [0, 3], n=4)
    assert got == [(0, 0, None), (3, 0, None)] and st["edges_traversed"] == 0
