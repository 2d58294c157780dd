"""CPU oracle for warren-core's whole-graph compute path.

TEST INFRASTRUCTURE ONLY: the correctness checker and the timed CPU baseline.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package; the product (``pg_warren_amd``) never does.

A ctypes veneer over ``liboracle.so`` (built from ``warren_oracle.cpp`` by the
Makefile beside it). Each function cites the reference code it restates in
the C++ source.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# Status codes (shared numbering with the product's WG_* codes).
OK = 0
STATUS_NAMES = {
    0: "ok", 1: "cancelled", 2: "dtype_mismatch", 3: "id_space_mismatch",
    4: "width_mismatch", 5: "memory_limit", 6: "out_of_memory",
    7: "division_by_zero", 8: "float_overflow", 9: "float_underflow",
    10: "bigint_out_of_range", 11: "invalid_argument",
    12: "node_outside_id_space", 13: "node_filter_id_space", 14: "invalid_csr",
    15: "unsupported",
}

DTYPES = {"f32": 0, "f64": 1, "i64": 2, "bool": 3, "u32": 4}
NP_DTYPES = {"f32": np.float32, "f64": np.float64, "i64": np.int64,
             "bool": np.uint8, "u32": np.uint32}
SEMIRINGS = ["plus_times", "min_plus", "max_plus", "max_times", "min_times",
             "lor_land", "any_pair", "plus_pair", "plus_first", "plus_second",
             "min_first", "min_second", "max_second"]
BINARY_OPS = ["first", "second", "pair", "plus", "times", "min", "max", "lor",
              "land", "any"]


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(STATUS_NAMES.get(code, str(code)))
        self.code = code
        self.name = STATUS_NAMES.get(code, str(code))


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Csr(C.Structure):
    _fields_ = [("n", C.c_uint64), ("offsets", C.c_void_p), ("targets", C.c_void_p),
                ("weights", C.c_void_p)]


class _Visit(C.Structure):
    _fields_ = [("node", C.c_uint64), ("via", C.c_uint64), ("depth", C.c_uint32),
                ("pad", C.c_uint32)]


class _Stats(C.Structure):
    _fields_ = [("nodes_visited", C.c_uint64), ("edges_traversed", C.c_uint64),
                ("truncated", C.c_uint32), ("levels_pulled", C.c_uint32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.or_csr_validate.argtypes = [P]
        L.or_csr_build.argtypes = [C.c_uint64, C.c_uint64, P, P, P, P, P, P, P]
        L.or_csr_transpose.argtypes = [P, P, P, P]
        L.or_csr_build_merged.argtypes = [C.c_uint64, C.c_uint64, P, P, P, P, C.c_uint32, C.c_int,
                                          P, P, P, P, P, P]
        L.or_spmv.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_int, C.c_int,
                              C.c_int, C.c_int, C.c_uint64, P, P, C.c_int]
        L.or_pagerank.argtypes = [P, P, P, C.c_double, C.c_double, C.c_int, C.c_uint32, P, P, P,
                                  C.c_int]
        L.or_pagerank_backward.argtypes = [P, P, C.c_double, C.c_double, C.c_int, C.c_uint32, P,
                                           P, P, P, P, P, C.c_int]
        L.or_bfs.argtypes = [P, P, P, C.c_uint64, P, C.c_int64, C.c_int64, P, P, P]
        L.or_bfs_rows.argtypes = [C.c_uint64, P, C.c_uint32, P, C.c_uint32, P, C.c_uint64, P, C.c_int64,
                                  C.c_int64, P, P, P]
        L.or_bfs_depths.argtypes = [P, P, C.c_uint64, P, C.c_int]
        L.or_apply.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P]
        L.or_from_weight.argtypes = [C.c_int, C.c_double, P]
        L.or_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.or_uniform.restype = C.c_double
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


def _check(code):
    if code != OK:
        raise OracleError(code)


class Csr:
    """A CSR held in numpy arrays (u64 offsets, u32 targets, f64 weights)."""

    def __init__(self, n, offsets, targets, weights=None):
        self.n = int(n)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.targets = np.ascontiguousarray(targets, dtype=np.uint32)
        self.weights = None if weights is None else np.ascontiguousarray(weights, np.float64)
        self._c = _Csr(self.n, _p(self.offsets), _p(self.targets), _p(self.weights))

    @property
    def nnz(self):
        return int(self.offsets[-1])

    def ref(self):
        return C.byref(self._c)

    def transpose(self):
        m = self.nnz
        off = np.zeros(self.n + 1, np.uint64)
        tgt = np.zeros(m, np.uint32)
        w = np.zeros(m, np.float64) if self.weights is not None else None
        _check(lib().or_csr_transpose(self.ref(), _p(off), _p(tgt), _p(w)))
        return Csr(self.n, off, tgt, w)

    def validate(self):
        return lib().or_csr_validate(self.ref())


def csr_from_edges(n, src, dst, weights=None):
    """GraphCsr::build_as for one relation: dedup summing weights in input
    order, self loops kept."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    m = len(src)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    off = np.zeros(n + 1, np.uint64)
    tgt = np.zeros(max(m, 1), np.uint32)
    wo = np.zeros(max(m, 1), np.float64) if w is not None else None
    nnz = np.zeros(1, np.uint64)
    _check(lib().or_csr_build(n, m, _p(src), _p(dst), _p(w), _p(off), _p(tgt), _p(wo), _p(nnz)))
    k = int(nnz[0])
    return Csr(n, off, tgt[:k].copy(), None if wo is None else wo[:k].copy())


def csr_from_edge_list(n, src, dst, weights=None, relations=None, relation_count=1,
                       keep_parallel=False):
    """GraphCsr::build's merged all-relations CSR (graph/csr/build/mod.rs:248-288):
    returns (Csr, folded). Raises OracleError 21 / 22 for the first edge outside
    the id space / the dictionary, with ``.edge`` set to its index."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    m = len(src)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    rel = None if relations is None else np.ascontiguousarray(relations, np.uint32)
    off = np.zeros(n + 1, np.uint64)
    tgt = np.zeros(max(m, 1), np.uint32)
    wo = np.zeros(max(m, 1), np.float64) if w is not None else None
    out = np.zeros(3, np.uint64)
    code = lib().or_csr_build_merged(n, m, _p(src), _p(dst), _p(w), _p(rel), relation_count,
                                     int(keep_parallel), _p(off), _p(tgt), _p(wo), _p(out[0:]),
                                     _p(out[1:]), _p(out[2:]))
    if code != OK:
        err = OracleError(code)
        err.edge = int(out[2])
        raise err
    k = int(out[0])
    return Csr(n, off, tgt[:k].copy(), None if wo is None else wo[:k].copy()), int(out[1])


def relation_csrs(n, src, dst, weights, relations, relation_count):
    """Each relation's own CSR (GraphCsr's relation CSRs): the edges of
    relation r built on their own."""
    src, dst = np.asarray(src), np.asarray(dst)
    relations = np.asarray(relations)
    out = []
    for r in range(relation_count):
        sel = relations == r
        out.append(csr_from_edge_list(n, src[sel], dst[sel],
                                      None if weights is None else np.asarray(weights)[sel])[0])
    return out


def select(csrs, relations=None, per_relation=False):
    """The matrix a product over `relations` reads (compute/product/matrix.rs:
    247-294), from the relation CSRs `csrs` in relation order: one relation's
    CSR as stored; several, each row's entries merged by column, ties in
    relation order, an edge in several relations one entry (weights summed in
    relation order, 1 for an unweighted relation's entry when any selected
    relation has weights) or with per_relation one entry each."""
    sel = sorted(set(range(len(csrs)) if relations is None else relations))
    if len(sel) == 1:
        return csrs[sel[0]]
    n = csrs[0].n
    weighted = any(csrs[r].weights is not None for r in sel)
    src, dst, w, rel = [], [], [], []
    for i, r in enumerate(sel):
        c = csrs[r]
        rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(c.offsets.astype(np.int64)))
        src.append(rows)
        dst.append(c.targets)
        rel.append(np.full(len(c.targets), i, np.uint32))
        if weighted:
            w.append(c.weights if c.weights is not None else np.ones(len(c.targets)))
    cat = lambda parts, dt: np.concatenate(parts).astype(dt) if parts else np.zeros(0, dt)
    g, _ = csr_from_edge_list(n, cat(src, np.uint32), cat(dst, np.uint32),
                              cat(w, np.float64) if weighted else None, cat(rel, np.uint32),
                              max(len(sel), 1), keep_parallel=per_relation)
    return g


def spmv(fwd, x, semiring="plus_times", dtype=None, *, rev=None, transpose=False,
         weighted=True, mask=None, mask_complement=False, replace=False, accumulate=None,
         y=None, width=1, threads=0):
    """y = A (x) x (or A^T (x) x); returns y (a new array unless given)."""
    if dtype is None:
        dtype = {np.dtype(v): k for k, v in NP_DTYPES.items()}[np.asarray(x).dtype]
    x = np.ascontiguousarray(x, NP_DTYPES[dtype])
    if y is None:
        y = np.zeros_like(x)
    mdt = 0
    if mask is not None:
        mask = np.ascontiguousarray(mask)
        mdt = {np.dtype(v): DTYPES[k] for k, v in NP_DTYPES.items()}[mask.dtype]
    sr = SEMIRINGS.index("lor_land" if semiring == "or_and" else semiring)
    acc = -1 if accumulate is None else BINARY_OPS.index(accumulate)
    _check(lib().or_spmv(fwd.ref(), None if rev is None else rev.ref(), sr, DTYPES[dtype],
                         int(transpose), int(weighted), _p(mask), mdt, int(mask_complement),
                         int(replace), acc, width, _p(x), _p(y), threads))
    return y


def pagerank(fwd, rev=None, *, damping=0.85, tol=1e-6, max_iter=100, seed=None,
             tombstones=None, threads=0):
    n = fwd.n
    out = np.zeros(n, np.float64)
    it = np.zeros(1, np.uint32)
    tomb = None if tombstones is None else np.ascontiguousarray(tombstones, np.uint8)
    sd = None if seed is None else np.ascontiguousarray(seed, np.float64)
    _check(lib().or_pagerank(fwd.ref(), None if rev is None else rev.ref(), _p(tomb),
                             damping, 0.0 if tol is None else tol, int(tol is not None),
                             max_iter, _p(sd), _p(out), _p(it), threads))
    return out, int(it[0])


def pagerank_backward(fwd, x, g, *, damping=0.85, tol=1e-6, max_iter=100, seed=None,
                      tombstones=None, threads=0):
    """Adjoint of pagerank over the weighted fwd CSR: (dL/dw per forward
    entry, the adjoint vector, iterations) from the ranks x and g = dL/dx."""
    n = fwd.n
    lam = np.zeros(n, np.float64)
    dw = np.zeros(max(fwd.nnz, 1), np.float64)
    it = np.zeros(1, np.uint32)
    tomb = None if tombstones is None else np.ascontiguousarray(tombstones, np.uint8)
    sd = None if seed is None else np.ascontiguousarray(seed, np.float64)
    xs = np.ascontiguousarray(x, np.float64)
    gs = np.ascontiguousarray(g, np.float64)
    _check(lib().or_pagerank_backward(fwd.ref(), _p(tomb), damping, 0.0 if tol is None else tol,
                                      int(tol is not None), max_iter, _p(sd), _p(xs), _p(gs),
                                      _p(lam), _p(dw), _p(it), threads))
    return dw[:fwd.nnz], lam, int(it[0])


def adjacency(merged, relation_csrs, direction="out", relations=None):
    """The CSR lists a traversal reads (Adjacency::new / reversed,
    traversal/adjacency.rs:68-152, 288-320): (rows, reversed rows). merged:
    the all-relations (forward, reverse) pair; relation_csrs: each
    relation's (forward, reverse) pair in relation order. Every relation:
    the merged pair, Out [fwd], In [rev], Both [fwd, rev]; a subset (sorted,
    deduplicated): each relation's pair in turn, read the same way. The
    reversed adjacency reads the opposite direction (Both stays Both)."""
    def read(pairs, d):
        out = []
        for f, r in pairs:
            out += {"out": [f], "in": [r], "both": [f, r]}[d]
        return out
    pairs = [merged] if relations is None else [relation_csrs[i] for i in sorted(set(relations))]
    back = {"out": "in", "in": "out", "both": "both"}[direction]
    return read(pairs, direction), read(pairs, back)


def bfs(adj, starts, *, pull_rev=None, node_filter=None, max_depth=None, max_results=None, n=None):
    """Returns (visits as list of (node, depth, via|None), stats dict). adj
    and pull_rev are a Csr or a list of them (an Adjacency's rows, read in
    order); n the id space when the lists are empty."""
    rows = list(adj) if isinstance(adj, (list, tuple)) else [adj]
    rrows = None if pull_rev is None else (list(pull_rev) if isinstance(pull_rev, (list, tuple)) else [pull_rev])
    n = rows[0].n if n is None else n
    st = np.ascontiguousarray(starts, np.uint64)
    vis = (_Visit * max(n, 1))()
    cnt = np.zeros(1, np.uint64)
    stats = _Stats()
    flt = None if node_filter is None else np.ascontiguousarray(node_filter, np.uint8)
    rp = (C.POINTER(_Csr) * max(len(rows), 1))(*[C.pointer(c._c) for c in rows])
    rr = None if rrows is None else (C.POINTER(_Csr) * max(len(rrows), 1))(*[C.pointer(c._c) for c in rrows])
    _check(lib().or_bfs_rows(n, C.cast(rp, C.c_void_p), len(rows), None if rr is None else C.cast(rr, C.c_void_p),
                             0 if rrows is None else len(rrows), _p(st), len(st),
                             _p(flt), -1 if max_depth is None else max_depth,
                             -1 if max_results is None else max_results, C.addressof(vis), _p(cnt),
                             C.byref(stats)))
    out = []
    for i in range(int(cnt[0])):
        v = vis[i]
        out.append((v.node, v.depth, None if v.via == 2**64 - 1 else v.via))
    return out, {"nodes_visited": stats.nodes_visited, "edges_traversed": stats.edges_traversed,
                 "truncated": bool(stats.truncated), "levels_pulled": stats.levels_pulled}


def bfs_depths(adj, sources, threads=0):
    src = np.ascontiguousarray(sources, np.uint64)
    out = np.zeros((len(src), adj.n), np.uint32)
    _check(lib().or_bfs_depths(adj.ref(), _p(src), len(src), _p(out), threads))
    return out


def apply(op, a, b, dtype, monoid=None):
    """BinaryOp `op` on a, b (SemiringElement::apply), or with `monoid` the
    semiring multiply SemiringElement::multiply(monoid, op, a, b)."""
    t = NP_DTYPES[dtype]
    av, bv, out = np.array([a], t), np.array([b], t), np.zeros(1, t)
    mo = -1 if monoid is None else BINARY_OPS.index(monoid)
    _check(lib().or_apply(BINARY_OPS.index(op), mo, DTYPES[dtype], _p(av), _p(bv), _p(out)))
    return out[0]


def from_weight(weight, dtype):
    out = np.zeros(1, NP_DTYPES[dtype])
    _check(lib().or_from_weight(DTYPES[dtype], weight, _p(out)))
    return out[0]


def uniform(seed, stream, node):
    return lib().or_uniform(seed, stream, node)
