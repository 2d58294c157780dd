"""Dense restatement of the spectral operators for small graphs (test
infrastructure only, never the product path): the matrices of
spectral/operator.rs:33-45 with the active nodes and degrees of
operator.rs:336-364, and the connected components of the active nodes with
their two-colourings (spectral/components.rs:16-140). The eigenvalues a test
compares against come from numpy's eigvalsh of these matrices, as the
reference's own spectrum oracle does (test/sql/oracle_spectrum.sql)."""
from __future__ import annotations

import numpy as np

MATRICES = ("adjacency", "laplacian", "normalized_laplacian", "random_walk")


def adjacency(csr, weighted=True):
    """A as a dense n x n array (weights, or 1 where unweighted)."""
    n = csr.n
    a = np.zeros((n, n))
    off = csr.offsets.astype(np.int64)
    rows = np.repeat(np.arange(n), np.diff(off))
    w = csr.weights if (weighted and csr.weights is not None) else np.ones(csr.nnz)
    np.add.at(a, (rows, csr.targets.astype(np.int64)), w)
    return a


def operator(csr, matrix, weighted=True, symmetrize=True):
    """(M over the active nodes, the active node ids, d^-1/2 at them):
    S = (A + A^T)/2 (or A), active = rows of S with an edge, D = S 1."""
    a = adjacency(csr, weighted)
    au = adjacency(csr, False)
    s = (a + a.T) / 2 if symmetrize else a
    su = (au + au.T) / 2 if symmetrize else au
    active = np.flatnonzero(su.sum(1) > 0)
    s = s[np.ix_(active, active)]
    d = s.sum(1)
    scale = 1.0 / np.sqrt(d) if matrix in ("normalized_laplacian", "random_walk") else None
    if matrix == "adjacency":
        m = s
    elif matrix == "laplacian":
        m = np.diag(d) - s
    elif matrix == "normalized_laplacian":
        m = np.eye(len(active)) - scale[:, None] * s * scale[None, :]
    elif matrix == "random_walk":
        m = scale[:, None] * s * scale[None, :]
    else:
        raise ValueError(matrix)
    return m, active, scale


def spectrum(csr, matrix, k, which, weighted=True, symmetrize=True):
    """The k eigenvalues at the `which` end, in the reference's order
    (largest: by descending magnitude, positive first; smallest: ascending)."""
    m, _, _ = operator(csr, matrix, weighted, symmetrize)
    ev = np.linalg.eigvalsh(m)
    if which == "smallest":
        return np.sort(ev)[:k]
    # Descending magnitude; within a run of magnitudes equal to rounding
    # (1e-9 of the largest), positive values first (ritz.rs:56-80).
    order = sorted(range(len(ev)), key=lambda i: (-abs(ev[i]), -ev[i]))
    width = 1e-9 * np.abs(ev).max()
    out, start = [], 0
    while start < len(order):
        first = abs(ev[order[start]])
        end = start
        while end < len(order) and first - abs(ev[order[end]]) <= width:
            end += 1
        run = order[start:end]
        out += [i for i in run if ev[i] >= 0] + [i for i in run if ev[i] < 0]
        start = end
    return ev[out][:k]


def components(csr, active):
    """(component of each node, colour, bipartite per component): union-find
    over the forward CSR's edges between active nodes, components numbered
    by their lowest node, colours relative to it (components.rs:39-140)."""
    n = csr.n
    act = np.zeros(n, bool)
    act[active] = True
    parent = list(range(n))
    parity = [False] * n
    odd = [False] * n

    def find(v):
        colour = False
        while parent[v] != v:
            colour ^= parity[v]
            v = parent[v]
        return v, colour

    off = csr.offsets.astype(np.int64)
    for s in range(n):
        if not act[s]:
            continue
        for p in range(off[s], off[s + 1]):
            t = int(csr.targets[p])
            if not act[t]:
                continue
            (ra, ca), (rb, cb) = find(s), find(t)
            if ra == rb:
                odd[ra] |= ca == cb
                continue
            lo, hi = min(ra, rb), max(ra, rb)
            parent[hi] = lo
            parity[hi] = ca == cb
            odd[lo] |= odd[hi]
    comp = np.zeros(n, np.int64)
    colour = np.zeros(n, bool)
    number, bipartite = {}, []
    for v in range(n):
        if not act[v]:
            continue
        r, c = find(v)
        if r not in number:
            number[r] = len(bipartite)
            bipartite.append(not odd[r])
        comp[v] = number[r]
        colour[v] = c
    return comp, colour, np.array(bipartite, bool)
