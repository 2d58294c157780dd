"""pg_warren_amd: warren-core's whole-graph compute path on one MI355X.

The Python host layer over libwarren_gpu.so, mirroring warren-core's API
(crates/warren-core/src/): ``Graph`` is the device ``GraphCsr`` (forward +
reverse CSR, graph/csr/graph_csr.rs:15-153), ``Vector`` an ``AnyVector``
(compute/vector.rs:9-50), ``Product`` the product descriptor
(compute/product/mod.rs:33-56) and ``spmv`` / ``spmm`` the products
(compute/product/mod.rs:77-110). Every call runs the HIP kernels; errors are
raised as :class:`WarrenError` carrying warren-core's error kind.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib

STATUS_NAMES = {
    0: "ok", 1: "cancelled", 2: "dtype_mismatch", 3: "id_space_mismatch",
    4: "width_mismatch", 5: "memory_limit", 6: "out_of_memory",
    7: "division_by_zero", 8: "float_overflow", 9: "float_underflow",
    10: "bigint_out_of_range", 11: "invalid_argument",
    12: "node_outside_id_space", 13: "node_filter_id_space", 14: "invalid_csr",
    15: "unsupported", 16: "device_error", 17: "not_symmetric", 18: "non_positive_degree",
    19: "eigenpair_count", 20: "not_finite", 21: "edge_outside_id_space",
    22: "relation_outside_dictionary", 23: "unknown_relation", 24: "invalid_segment",
    25: "combiner_type", 26: "sender_outside_id_space",
}
DTYPES = {"f32": 0, "f64": 1, "i64": 2, "bool": 3, "u32": 4}
NP_DTYPES = {"f32": np.float32, "f64": np.float64, "i64": np.int64, "bool": np.uint8,
             "u32": np.uint32}
# Semiring::ALL order (compute/semiring.rs:149-163).
SEMIRINGS = ["plus_times", "min_plus", "max_plus", "max_times", "min_times", "lor_land",
             "any_pair", "plus_pair", "plus_first", "plus_second", "min_first", "min_second",
             "max_second"]
# BinaryOp::ALL order (compute/semiring.rs:42-53).
BINARY_OPS = ["first", "second", "pair", "plus", "times", "min", "max", "lor", "land", "any"]


class WarrenError(RuntimeError):
    """A failed call: ``name`` is warren-core's error kind."""

    def __init__(self, code, message=""):
        self.code = code
        self.name = STATUS_NAMES.get(code, str(code))
        super().__init__(f"{self.name}: {message}" if message else self.name)


def _p(a):
    return None if a is None else a.ctypes.data


def semiring_index(name):
    """Semiring::from_str (compute/semiring.rs:219-233): canonical names and
    the ``or_and`` alias."""
    canonical = "lor_land" if name == "or_and" else name
    if canonical not in SEMIRINGS:
        raise WarrenError(11, f"unknown semiring {name!r}")
    return SEMIRINGS.index(canonical)


def _dtype_of(arr):
    for k, v in NP_DTYPES.items():
        if np.dtype(v) == arr.dtype:
            return k
    if arr.dtype == np.bool_:
        return "bool"
    raise WarrenError(11, f"unsupported element type {arr.dtype}")


class Context:
    """Device, stream, memory budget and cancellation (Workers +
    ComputeMemory + Interrupts)."""

    def __init__(self, device=0, memory_limit=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        code = self.lib.wg_context_create(device, memory_limit, C.byref(h))
        if code:
            raise WarrenError(code, "context creation")
        self.h = h
        # Vectors and graphs hold device memory of this context: they are
        # released before it is (the C ABI requires children to go first).
        self._children = weakref.WeakSet()
        self._pinned = None  # (pointer, bytes): page-locked room for results

    def check(self, code):
        if code:
            buf = C.create_string_buffer(512)
            self.lib.wg_last_error(self.h, buf, 512)
            raise WarrenError(code, buf.value.decode(errors="replace"))

    def cancel(self):
        self.lib.wg_cancel(self.h)

    def cancel_reset(self):
        self.lib.wg_cancel_reset(self.h)

    def set_progress(self, port):
        """Reports of long work go to port.phase(name, total or None),
        port.done(units) and port.detail(name, value) (progress.rs:9-16);
        None reports nowhere."""
        if port is None:
            self._progress = None
            self.lib.wg_set_progress(self.h, None)
            return
        c = _lib.ProgressPort(
            _lib.PHASE_FN(lambda _u, name, total: port.phase(name.decode(), None if total < 0 else total)),
            _lib.DONE_FN(lambda _u, done: port.done(done)),
            _lib.DETAIL_FN(lambda _u, name, value: port.detail(name.decode(), value)), None)
        self._progress = c  # the callbacks live as long as the port is set
        self.lib.wg_set_progress(self.h, C.byref(c))

    def memory_in_use(self):
        return self.lib.wg_memory_in_use(self.h)

    def set_memory_limit(self, limit):
        """A new budget for later calls (0: none), as ComputeMemory::new."""
        self.check(self.lib.wg_set_memory_limit(self.h, int(limit)))

    def synchronize(self):
        self.check(self.lib.wg_synchronize(self.h))

    def set_graph_launch(self, on):
        """Replay iterative launch sequences as captured HIP graphs (default)
        or enqueue them one by one."""
        self.lib.wg_set_graph_launch(self.h, int(bool(on)))

    def set_blocking(self, blocking):
        self.lib.wg_set_blocking(self.h, int(blocking))

    @property
    def stream(self):
        return self.lib.wg_stream(self.h)

    def host_array(self, ctype, count):
        """A ctypes array of count ctype over the context's page-locked
        buffer (wg_host_alloc, grown as needed and reused: its contents last
        until the next call), which results are copied into by DMA."""
        need = max(C.sizeof(ctype) * max(int(count), 1), 1)
        if self._pinned is None or self._pinned[1] < need:
            self._free_pinned()
            ptr = C.c_void_p()
            self.check(self.lib.wg_host_alloc(self.h, need, C.byref(ptr)))
            self._pinned = (ptr.value, need)
        return (ctype * max(int(count), 1)).from_address(self._pinned[0])

    def _free_pinned(self):
        if self._pinned is not None:
            self.lib.wg_host_free(self._pinned[0])
            self._pinned = None

    def close(self):
        if getattr(self, "h", None):
            for child in list(self._children):
                child.close()
            self._free_pinned()
            self.lib.wg_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Vector:
    """A device AnyVector: row-major id_space x width of one dtype."""

    def __init__(self, ctx, values=None, *, dtype=None, id_space=None, width=1):
        self.ctx = ctx
        if values is not None:
            values = np.asarray(values)
            if dtype is None:
                dtype = _dtype_of(values)
            values = np.ascontiguousarray(values, NP_DTYPES[dtype])
            id_space = values.shape[0]
            width = 1 if values.ndim == 1 else values.shape[1]
        self.dtype, self.id_space, self.width = dtype, int(id_space), int(width)
        h = C.c_void_p()
        ctx.check(ctx.lib.wg_vector_create(ctx.h, DTYPES[dtype], self.id_space, self.width,
                                           _p(values), C.byref(h)))
        self.h = h
        ctx._children.add(self)

    def numpy(self):
        shape = (self.id_space,) if self.width == 1 else (self.id_space, self.width)
        out = np.empty(shape, NP_DTYPES[self.dtype])
        self.ctx.check(self.ctx.lib.wg_vector_read(self.ctx.h, self.h, _p(out)))
        return out

    def write(self, values):
        values = np.ascontiguousarray(values, NP_DTYPES[self.dtype])
        self.ctx.check(self.ctx.lib.wg_vector_write(self.ctx.h, self.h, _p(values)))

    @property
    def device_ptr(self):
        return self.ctx.lib.wg_vector_device_ptr(self.h)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h is not None:
            self.ctx.lib.wg_vector_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A device GraphCsr: the merged all-relations CSR and its reverse."""

    def __init__(self, ctx, n, offsets, targets, weights=None, target_width=None):
        """target_width: 4 or 8 bytes; by default u64 targets exactly when the
        id space is 2^32 or more (TargetWidth::for_id_space), else u32. A
        graph with u64 targets serves products only."""
        self.ctx = ctx
        self.n = int(n)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        if target_width is None:
            target_width = 8 if self.n >= 1 << 32 else 4
        self.wide = target_width == 8
        targets = np.ascontiguousarray(targets, np.uint64 if self.wide else np.uint32)
        if weights is not None:
            weights = np.ascontiguousarray(weights, np.float64)
        h = C.c_void_p()
        upload = ctx.lib.wg_graph_upload_u64 if self.wide else ctx.lib.wg_graph_upload
        ctx.check(upload(ctx.h, self.n, _p(offsets), _p(targets), _p(weights), C.byref(h)))
        self.h = h
        ctx._children.add(self)
        self.nnz = int(offsets[-1])
        self.weighted = weights is not None
        self.deduplicated = 0

    def read_u64(self, reverse=False):
        """A u64-target graph's forward or reverse CSR: (offsets, targets, weights)."""
        off = np.zeros(self.n + 1, np.uint64)
        tgt = np.zeros(max(self.nnz, 1), np.uint64)
        w = np.zeros(max(self.nnz, 1), np.float64) if self.weighted else None
        self.ctx.check(self.ctx.lib.wg_graph_read_u64(self.ctx.h, self.h, int(reverse), _p(off), _p(tgt), _p(w)))
        return off, tgt[:self.nnz], None if w is None else w[:self.nnz]

    @classmethod
    def build(cls, ctx, n, sources, targets, weights=None, relations=None, relation_count=1):
        """GraphCsr::build_as (graph/csr/graph_csr.rs:210-221) on the device from
        an edge list: rows sorted, parallel edges folded (weights summed in input
        order, then across relations in relation order), self loops kept. Returns
        the graph; ``graph.deduplicated`` counts the edges folded into another of
        the same relation. Arrays are numpy (host) arrays."""
        sources, targets = np.asarray(sources), np.asarray(targets)
        for col in (sources, targets):
            if col.dtype != np.uint32 and col.size and (col.min() < 0 or col.max() >= 2**32):
                bad = int(np.flatnonzero((sources < 0) | (sources >= 2**32) | (targets < 0)
                                         | (targets >= 2**32))[0])
                raise WarrenError(21, f"edge {bad} from {sources[bad]} to {targets[bad]} lies "
                                      "outside the id space")
        src = np.ascontiguousarray(sources, np.uint32)
        dst = np.ascontiguousarray(targets, np.uint32)
        if src.shape != dst.shape:
            raise ValueError("sources and targets differ in length")
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        rel = None if relations is None else np.ascontiguousarray(relations, np.uint32)
        for a in (w, rel):
            if a is not None and a.shape != src.shape:
                raise ValueError("edge columns differ in length")
        h = C.c_void_p()
        folded = C.c_uint64()
        ctx.check(ctx.lib.wg_graph_build(ctx.h, int(n), len(src), _p(src), _p(dst), _p(w), _p(rel),
                                         int(relation_count), 0, C.byref(h), C.byref(folded)))
        return cls._adopt(ctx, h, int(n), w is not None, int(folded.value))

    @classmethod
    def load_segments(cls, ctx, id_space, forward, reverse=None):
        """AnyCsr::decode (graph/segment/csr.rs:84-127) straight to the device:
        forward (and optionally reverse) as dicts of segment bytes with keys
        'offsets', 'targets' and optionally 'weights' and 'rows'."""
        keep = []

        def seg(d):
            c = _lib.CsrSegments()
            for k in ("rows", "offsets", "targets", "weights"):
                b = d.get(k)
                if b is not None:
                    arr = np.frombuffer(bytes(b), np.uint8) if len(b) else np.zeros(1, np.uint8)
                    keep.append(arr)
                    setattr(c, k, arr.ctypes.data)
                    setattr(c, k + "_len", len(b))
            return c

        fw = seg(forward)
        rv = None if reverse is None else seg(reverse)
        h = C.c_void_p()
        ctx.check(ctx.lib.wg_graph_load_segments(ctx.h, int(id_space), C.byref(fw),
                                                 None if rv is None else C.byref(rv), C.byref(h)))
        return cls._adopt(ctx, h, int(id_space), forward.get("weights") is not None)

    def encode_segments(self, reverse=False):
        """AnyCsr::encode with dense offsets: {'offsets', 'targets', 'weights'}."""
        lo, lt, lw = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.lib.wg_csr_segment_lengths(self.n, self.nnz, C.byref(lo), C.byref(lt), C.byref(lw))
        o = np.zeros(lo.value, np.uint8)
        t = np.zeros(lt.value, np.uint8)
        w = np.zeros(lw.value, np.uint8) if self.weighted else None
        self.ctx.check(self.ctx.lib.wg_graph_encode_segments(self.ctx.h, self.h, int(reverse), _p(o),
                                                             _p(t), _p(w)))
        return {"offsets": o.tobytes(), "targets": t.tobytes(),
                "weights": None if w is None else w.tobytes()}

    def encode_relation_segments(self, relation, reverse=False):
        """A relation's CSR as its segments in the reference's layout (sparse
        offsets where they pay, graph/csr/sparse.rs): {'rows' (sparse only),
        'offsets', 'targets', 'weights'}."""
        ln = [C.c_uint64() for _ in range(4)]
        self.ctx.check(self.ctx.lib.wg_relation_segment_lengths(self.ctx.h, self.h, int(relation), int(reverse),
                                                                *[C.byref(x) for x in ln]))
        bufs = [np.zeros(x.value, np.uint8) if x.value else None for x in ln]
        self.ctx.check(self.ctx.lib.wg_relation_encode_segments(self.ctx.h, self.h, int(relation), int(reverse),
                                                                *[_p(b) for b in bufs]))
        out = dict(zip(("rows", "offsets", "targets", "weights"),
                       (None if b is None else b.tobytes() for b in bufs)))
        if out["rows"] is None:
            del out["rows"]
        return out

    @classmethod
    def _adopt(cls, ctx, h, n, weighted, folded=0):
        g = cls.__new__(cls)
        g.ctx = ctx
        g.n = n
        g.h = h
        ctx._children.add(g)
        nnz = C.c_uint64()
        ctx.check(ctx.lib.wg_graph_info(h, None, C.byref(nnz), None))
        g.nnz = int(nnz.value)
        g.weighted = weighted
        g.deduplicated = folded
        g.wide = False
        return g

    def select(self, relations=None, per_relation=False):
        """The graph of a relation selection (GraphCsr::selected and the matrix a
        product reads, compute/product/matrix.rs:126-157): kept with this graph,
        usable wherever a graph is."""
        rels, nrels = None, 0
        if relations is not None:
            arr = np.ascontiguousarray(relations, np.uint32).reshape(-1)
            nrels = arr.size
            arr = arr if nrels else np.zeros(1, np.uint32)
            rels = arr.ctypes.data
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.wg_graph_select(self.ctx.h, self.h, rels, nrels, int(per_relation),
                                                    C.byref(h)))
        if h.value == self.h.value:
            return self
        weighted = C.c_int()
        self.ctx.check(self.ctx.lib.wg_graph_info(h, None, None, C.byref(weighted)))
        view = Graph._adopt(self.ctx, h, self.n, bool(weighted.value))
        view._owner = self  # the view lives as long as this graph
        return view

    @property
    def relation_count(self):
        c = C.c_uint32()
        self.ctx.check(self.ctx.lib.wg_graph_relation_count(self.h, C.byref(c)))
        return int(c.value)

    def forward(self):
        off = np.zeros(self.n + 1, np.uint64)
        tgt = np.zeros(max(self.nnz, 1), np.uint32)
        w = np.zeros(max(self.nnz, 1), np.float64) if self.weighted else None
        self.ctx.check(self.ctx.lib.wg_graph_read_forward(self.ctx.h, self.h, _p(off), _p(tgt),
                                                          _p(w)))
        return off, tgt[:self.nnz], None if w is None else w[:self.nnz]

    def reverse(self):
        off = np.zeros(self.n + 1, np.uint64)
        tgt = np.zeros(max(self.nnz, 1), np.uint32)
        w = np.zeros(max(self.nnz, 1), np.float64) if self.weighted else None
        self.ctx.check(self.ctx.lib.wg_graph_read_reverse(self.ctx.h, self.h, _p(off), _p(tgt),
                                                          _p(w)))
        return off, tgt[:self.nnz], None if w is None else w[:self.nnz]

    def update_weights(self, edges, values):
        """Sets the weights of the given edges (forward-CSR indices) in place."""
        e = np.ascontiguousarray(edges, np.uint64)
        v = np.ascontiguousarray(values, np.float64)
        if e.shape != v.shape:
            raise ValueError("edges and values differ in length")
        self.ctx.check(self.ctx.lib.wg_graph_update_weights(self.ctx.h, self.h, _p(e), _p(v), len(e)))

    def close(self):
        if getattr(self, "h", None) and self.ctx.h is not None:
            self.ctx.lib.wg_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Product:
    """What to multiply and how to write the result (Product,
    compute/product/mod.rs:33-56). weighted=None: if the graph has weights."""

    def __init__(self, semiring="plus_times", *, transpose=False, weighted=None, mask=None,
                 mask_complement=False, replace=False, accumulate=None, relations=None,
                 per_relation=False):
        self.semiring = semiring
        self.transpose = transpose
        self.weighted = weighted
        self.mask = mask
        self.mask_complement = mask_complement
        self.replace = replace
        self.accumulate = accumulate
        self.relations = relations  # None: every relation (RelationSelection::All)
        self.per_relation = per_relation

    def c(self):
        acc = -1
        if self.accumulate is not None:
            if self.accumulate not in BINARY_OPS:
                raise WarrenError(11, f"unknown binary op {self.accumulate!r}")
            acc = BINARY_OPS.index(self.accumulate)
        rels, nrels = None, 0
        if self.relations is not None:
            # Kept on the instance: the struct holds a raw pointer to it.
            self._rels = np.ascontiguousarray(self.relations, np.uint32).reshape(-1)
            if self._rels.size == 0:
                self._rels = np.zeros(1, np.uint32)  # a valid pointer for an empty list
                nrels = 0
            else:
                nrels = self._rels.size
            rels = self._rels.ctypes.data
        return _lib.Product(semiring_index(self.semiring), int(self.transpose),
                            -1 if self.weighted is None else int(self.weighted),
                            int(self.mask_complement),
                            None if self.mask is None else self.mask.h.value,
                            int(self.replace), acc, rels, nrels, int(self.per_relation))


def spmv(ctx, graph, product, x, y):
    """y = A (x) x, or A^T (x) x (compute/product/mod.rs:77-92)."""
    p = product.c()
    ctx.check(ctx.lib.wg_spmv(ctx.h, graph.h, C.byref(p), x.h, y.h))
    return y


def spmm(ctx, graph, product, x, y):
    """Y = A (x) X for row-major X of any width (compute/product/mod.rs:97-110)."""
    p = product.c()
    ctx.check(ctx.lib.wg_spmm(ctx.h, graph.h, C.byref(p), x.h, y.h))
    return y


def pagerank(ctx, graph, *, damping=0.85, tol=1e-6, max_iter=100, seed=None, tombstones=None,
             weighted=None, out=None):
    """PageRank (compute/pagerank/mod.rs:96-151): NetworkX's power
    iteration in f64, stopping once an iteration's L1 change falls below live
    nodes x tol (tol None: never). seed / tombstones: device Vectors or
    array-likes (tombstones as bool). Returns (ranks Vector, iterations)."""
    if seed is not None and not isinstance(seed, Vector):
        seed = Vector(ctx, np.asarray(seed))
    if tombstones is not None and not isinstance(tombstones, Vector):
        tombstones = Vector(ctx, np.asarray(tombstones, dtype=np.uint8), dtype="bool")
    if out is None:
        out = Vector(ctx, dtype="f64", id_space=graph.n)
    opts = _lib.PagerankOpts(damping, 0.0 if tol is None else tol, int(tol is not None),
                             max_iter, -1 if weighted is None else int(weighted), 0)
    it = C.c_uint32()
    ctx.check(ctx.lib.wg_pagerank(ctx.h, graph.h, None if tombstones is None else tombstones.h,
                                  C.byref(opts), None if seed is None else seed.h, out.h,
                                  C.byref(it)))
    return out, it.value


def pagerank_backward(ctx, graph, ranks, dl_dranks, *, damping=0.85, tol=1e-6, max_iter=100,
                      seed=None, tombstones=None, out=None, adjoint=False):
    """dL/dw for every edge (forward-CSR order) of a weighted pagerank run
    with these options, from its ranks and dL/dranks. Returns (dL/dw Vector
    over nnz, iterations) or, with adjoint=True, (dL/dw, iterations, the
    adjoint Vector)."""
    if seed is not None and not isinstance(seed, Vector):
        seed = Vector(ctx, np.asarray(seed))
    if tombstones is not None and not isinstance(tombstones, Vector):
        tombstones = Vector(ctx, np.asarray(tombstones, dtype=np.uint8), dtype="bool")
    if not isinstance(ranks, Vector):
        ranks = Vector(ctx, np.asarray(ranks, np.float64))
    if not isinstance(dl_dranks, Vector):
        dl_dranks = Vector(ctx, np.asarray(dl_dranks, np.float64))
    if out is None:
        out = Vector(ctx, dtype="f64", id_space=graph.nnz)
    adj = Vector(ctx, dtype="f64", id_space=graph.n) if adjoint else None
    opts = _lib.PagerankOpts(damping, 0.0 if tol is None else tol, int(tol is not None),
                             max_iter, 1, 0)
    it = C.c_uint32()
    ctx.check(ctx.lib.wg_pagerank_backward(ctx.h, graph.h, None if tombstones is None else tombstones.h,
                                           C.byref(opts), None if seed is None else seed.h, ranks.h,
                                           dl_dranks.h, out.h, None if adj is None else adj.h,
                                           C.byref(it)))
    return (out, it.value, adj) if adjoint else (out, it.value)


def pagerank_batch(ctx, graph, seed_sets, *, seed_weights=None, damping=0.85, tol=1e-6,
                   max_iter=100, tombstones=None, weighted=None, out=None):
    """Personalised PageRank for many seed sets at once: column c is
    pagerank(seed = the vector holding seed_weights[c] (or 1) at
    seed_sets[c]). Returns (ranks Vector f64 n x R, iterations per column)."""
    R = len(seed_sets)
    offs = np.zeros(R + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seed_sets])
    nodes = np.ascontiguousarray(np.concatenate([np.asarray(s, np.uint64) for s in seed_sets])
                                 if R else np.zeros(0, np.uint64), dtype=np.uint64)
    wts = None
    if seed_weights is not None:
        wts = np.ascontiguousarray(np.concatenate([np.asarray(w, np.float64) for w in seed_weights]))
    if tombstones is not None and not isinstance(tombstones, Vector):
        tombstones = Vector(ctx, np.asarray(tombstones, dtype=np.uint8), dtype="bool")
    if out is None:
        out = Vector(ctx, dtype="f64", id_space=graph.n, width=R)
    opts = _lib.PagerankOpts(damping, 0.0 if tol is None else tol, int(tol is not None),
                             max_iter, -1 if weighted is None else int(weighted), 0)
    its = np.zeros(R, np.uint32)
    ctx.check(ctx.lib.wg_pagerank_batch(ctx.h, graph.h, None if tombstones is None else tombstones.h,
                                        C.byref(opts), _p(offs), _p(nodes), _p(wts), R, out.h,
                                        _p(its)))
    return out, its


DIRECTIONS = {"out": 0, "in": 1, "both": 2}


def _stats(st):
    return {"nodes_visited": st.nodes_visited, "edges_traversed": st.edges_traversed,
            "truncated": bool(st.truncated), "levels_pulled": st.levels_pulled}


def _starts(starts):
    return np.ascontiguousarray(np.atleast_1d(np.asarray(starts, dtype=np.uint64)))


def _filter(node_filter):
    return None if node_filter is None else np.ascontiguousarray(node_filter, dtype=np.uint8)


def _relations(relations):
    """(array or None, count): a traversal's relation subset, None for every
    relation (Adjacency::new, traversal/adjacency.rs:68-105)."""
    if relations is None:
        return None, 0
    r = np.atleast_1d(np.asarray(relations, dtype=np.int64))
    buf = np.zeros(max(len(r), 1), np.uint32)  # never a null pointer, even for no relation
    buf[:len(r)] = r
    return buf, len(r)


def bfs(ctx, graph, starts, *, direction="out", relations=None, node_filter=None, max_depth=None,
        max_results=None):
    """bfs (traversal/bfs.rs:34-45): [(node, depth, via or None)] in BFS order
    and the traversal stats. direction "out", "in" or "both"; relations a
    subset of the graph's relations (None: all)."""
    st = _starts(starts)
    flt = _filter(node_filter)
    rel, nrel = _relations(relations)
    out = ctx.host_array(_lib.Visit, graph.n)
    cnt = C.c_uint64()
    stats = _lib.TraversalStats()
    ctx.check(ctx.lib.wg_bfs(ctx.h, graph.h, DIRECTIONS[direction], _p(rel), nrel, _p(st), len(st), _p(flt),
                             -1 if max_depth is None else max_depth,
                             -1 if max_results is None else max_results, C.addressof(out),
                             C.byref(cnt), C.byref(stats)))
    arr = np.ctypeslib.as_array(out)[:cnt.value]
    visits = [(int(v["node"]), int(v["depth"]), None if v["via"] == 2**64 - 1 else int(v["via"]))
              for v in arr]
    return visits, _stats(stats)


def reachable_set(ctx, graph, starts, *, direction="out", relations=None, max_depth=None):
    """reachable_set (traversal/reachable.rs:48-90): node ids in BFS order."""
    st = _starts(starts)
    rel, nrel = _relations(relations)
    out = ctx.host_array(C.c_uint64, graph.n)
    cnt = C.c_uint64()
    stats = _lib.TraversalStats()
    ctx.check(ctx.lib.wg_reachable_set(ctx.h, graph.h, DIRECTIONS[direction], _p(rel), nrel, _p(st), len(st),
                                       -1 if max_depth is None else max_depth, C.addressof(out),
                                       C.byref(cnt), C.byref(stats)))
    return np.ctypeslib.as_array(out)[:cnt.value].astype(np.uint64), _stats(stats)


COMBINERS = ["collect", "count", "sum", "min", "max", "first", "concat", "bool_or", "bool_and", "mode"]
GATHER_DIRECTIONS = {"out": 0, "in": 1, "both": 2}


def _gather_opts(ctx, combiner, direction, relations, per_relation, mask, mask_complement, include_self):
    """wg_gather_opts, and what must outlive the call (relation list, mask)."""
    if combiner not in COMBINERS:
        raise WarrenError(11, f"unknown combiner {combiner!r}")
    rels, nrels, keep = None, 0, None
    if relations is not None:
        keep = np.ascontiguousarray(relations, np.uint32).reshape(-1)
        nrels = keep.size
        keep = keep if nrels else np.zeros(1, np.uint32)
        rels = keep.ctypes.data
    if mask is not None and not isinstance(mask, Vector):
        mask = Vector(ctx, np.asarray(mask))
    opts = _lib.GatherOpts(COMBINERS.index(combiner), GATHER_DIRECTIONS[direction], rels, nrels,
                           int(per_relation), None if mask is None else mask.h.value,
                           int(mask_complement), int(include_self))
    return opts, (keep, mask)


def _senders(graph, senders):
    snd = np.asarray(senders)
    if snd.size and (snd.min() < 0 or snd.max() >= 2**32):
        bad = int(np.flatnonzero((snd < 0) | (snd >= 2**32))[0])
        raise WarrenError(26, f"sender {snd[bad]} outside the id space {graph.n}")
    return np.ascontiguousarray(snd, np.uint32)


def _gather_bytes(ctx, graph, senders, messages, combiner, **kw):
    """gather over byte messages (wg_gather_bytes): collect gives a list of
    bytes a receiver, first and concat one bytes a receiver, count i64."""
    snd = _senders(graph, senders)
    if len(snd) != len(messages):
        raise ValueError("senders and messages differ in length")
    lens = np.array([len(m) for m in messages], np.uint64)
    offsets = np.zeros(len(messages) + 1, np.uint64)
    offsets[1:] = np.cumsum(lens)
    data = np.frombuffer(b"".join(bytes(m) for m in messages), np.uint8) if offsets[-1] else np.zeros(1, np.uint8)
    opts, _keep = _gather_opts(ctx, combiner, kw["direction"], kw["relations"], kw["per_relation"], kw["mask"],
                               kw["mask_complement"], kw["include_self"])
    h = C.c_void_p()
    ctx.check(ctx.lib.wg_gather_bytes(ctx.h, graph.h, C.byref(opts), len(snd), _p(snd), _p(data), _p(offsets),
                                      C.byref(h)))
    try:
        nr, nv, nb = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ctx.lib.wg_gathered_info(h, C.byref(nr), C.byref(nv))
        recv = np.zeros(max(nr.value, 1), np.uint64)
        if combiner == "count":
            vals = np.zeros(max(nv.value, 1), np.int64)
            ctx.check(ctx.lib.wg_gathered_read(ctx.h, h, _p(recv), _p(vals), None))
            return recv[:nr.value], vals[:nv.value]
        offs = np.zeros(nr.value + 1, np.uint64) if combiner == "collect" else None
        ctx.check(ctx.lib.wg_gathered_read(ctx.h, h, _p(recv), None, _p(offs)))
        ctx.lib.wg_gathered_bytes_info(h, C.byref(nb))
        buf = np.zeros(max(nb.value, 1), np.uint8)
        soffs = np.zeros(nv.value + 1, np.uint64)
        ctx.check(ctx.lib.wg_gathered_read_bytes(ctx.h, h, _p(buf), _p(soffs)))
    finally:
        ctx.lib.wg_gathered_destroy(h)
    raw = buf.tobytes()
    strings = [raw[int(soffs[i]):int(soffs[i + 1])] for i in range(nv.value)]
    recv = recv[:nr.value]
    if combiner == "collect":
        return recv, [strings[int(offs[i]):int(offs[i + 1])] for i in range(nr.value)]
    return recv, strings


def gather(ctx, graph, senders, messages, combiner, *, direction="in", relations=None,
           per_relation=False, mask=None, mask_complement=False, include_self=False):
    """gather (compute/gather/mod.rs:96-145): messages[i] sent by senders[i]
    (push order), combined at each receiver in collect order. Returns
    (receivers, values): values one per receiver (count: i64), or for
    collect a list of arrays, one per receiver. messages: a numpy array of
    f32, f64, i64 or bool, or a list of bytes objects (byte messages:
    collect, count, first and concat, the strings gathered on the device);
    mask: a Vector or None."""
    if isinstance(messages, (list, tuple)) and messages and all(
            isinstance(m, (bytes, bytearray)) for m in messages):
        return _gather_bytes(ctx, graph, senders, messages, combiner, direction=direction,
                             relations=relations, per_relation=per_relation, mask=mask,
                             mask_complement=mask_complement, include_self=include_self)
    msg = np.asarray(messages)
    dt = _dtype_of(msg)
    msg = np.ascontiguousarray(msg, NP_DTYPES[dt])
    snd = _senders(graph, senders)
    if snd.shape != msg.shape:
        raise ValueError("senders and messages differ in length")
    if combiner not in COMBINERS:
        raise WarrenError(11, f"unknown combiner {combiner!r}")
    opts, _keep = _gather_opts(ctx, combiner, direction, relations, per_relation, mask, mask_complement,
                               include_self)
    h = C.c_void_p()
    ctx.check(ctx.lib.wg_gather(ctx.h, graph.h, C.byref(opts), DTYPES[dt], len(snd), _p(snd), _p(msg),
                                C.byref(h)))
    try:
        nr, nv = C.c_uint64(), C.c_uint64()
        ctx.lib.wg_gathered_info(h, C.byref(nr), C.byref(nv))
        recv = np.zeros(max(nr.value, 1), np.uint64)
        vt = np.int64 if combiner == "count" else NP_DTYPES[dt]
        vals = np.zeros(max(nv.value, 1), vt)
        offs = np.zeros(nr.value + 1, np.uint64) if combiner == "collect" else None
        ctx.check(ctx.lib.wg_gathered_read(ctx.h, h, _p(recv), _p(vals), _p(offs)))
    finally:
        ctx.lib.wg_gathered_destroy(h)
    recv, vals = recv[:nr.value], vals[:nv.value]
    if dt == "bool":
        vals = vals.astype(bool) if combiner != "count" else vals
    if combiner == "collect":
        return recv, [vals[int(offs[i]):int(offs[i + 1])] for i in range(nr.value)]
    return recv, vals


class BfsCursor:
    """BfsCursor (traversal/bfs_cursor.rs:23-135): the search of ``bfs`` (or,
    with reachable=True, of ``reachable_set``) one level per ``next_level``
    call, the starts first. next_level returns (visits, more): visits as
    [(node, depth, via or None)] and more False once no level follows."""

    def __init__(self, ctx, graph, starts, *, direction="out", relations=None, node_filter=None,
                 max_depth=None, max_results=None, reachable=False):
        self.ctx, self.graph = ctx, graph
        st = _starts(starts)
        flt = _filter(node_filter)
        rel, nrel = _relations(relations)
        h = C.c_void_p()
        ctx.check(ctx.lib.wg_bfs_cursor_create(ctx.h, graph.h, DIRECTIONS[direction], _p(rel), nrel, _p(st),
                                               len(st),
                                               _p(flt), -1 if max_depth is None else max_depth,
                                               -1 if max_results is None else max_results,
                                               int(reachable), C.byref(h)))
        self.h = h
        ctx._children.add(self)

    def next_level(self):
        cnt = C.c_uint64()
        more = C.c_int()
        out = self.ctx.host_array(_lib.Visit, self.graph.n)
        self.ctx.check(self.ctx.lib.wg_bfs_cursor_next(self.ctx.h, self.h, C.addressof(out),
                                                       C.byref(cnt), C.byref(more)))
        arr = np.ctypeslib.as_array(out)[:cnt.value]
        visits = [(int(v["node"]), int(v["depth"]), None if v["via"] == 2**64 - 1 else int(v["via"]))
                  for v in arr]
        return visits, bool(more.value)

    def levels(self):
        """Every remaining level, as lists of visits (empty levels left out)."""
        out, more = [], True
        while more:
            level, more = self.next_level()
            if level:
                out.append(level)
        return out

    def stats(self):
        stats = _lib.TraversalStats()
        self.ctx.check(self.ctx.lib.wg_bfs_cursor_stats(self.h, C.byref(stats)))
        return _stats(stats)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h is not None:
            self.ctx.lib.wg_bfs_cursor_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bfs_levels(ctx, graph, starts, *, direction="out", relations=None, node_filter=None, max_depth=None,
               pulling=True, depth=None, via=None):
    """The device-resident search: (depth, via) u32 Vectors (0xffffffff where
    unreached) and the traversal stats."""
    st = _starts(starts)
    flt = _filter(node_filter)
    if depth is None:
        depth = Vector(ctx, dtype="u32", id_space=graph.n)
    if via is None:
        via = Vector(ctx, dtype="u32", id_space=graph.n)
    stats = _lib.TraversalStats()
    rel, nrel = _relations(relations)
    ctx.check(ctx.lib.wg_bfs_levels(ctx.h, graph.h, DIRECTIONS[direction], _p(rel), nrel, _p(st), len(st),
                                    _p(flt), -1 if max_depth is None else max_depth,
                                    int(pulling), depth.h, via.h, C.byref(stats)))
    return depth, via, _stats(stats)


MATRICES = {"adjacency": 0, "laplacian": 1, "normalized_laplacian": 2, "random_walk": 3}
WHICH = {"largest": 0, "smallest": 1}


def eigen(ctx, graph, k, *, matrix="adjacency", which="largest", weighted=None, symmetrize=True,
          tol=1e-6, max_iter=None, precision="f32", seed=0, basis_size=None, vectors=True):
    """eigen (spectral/lanczos.rs:206-224): the k eigenpairs of `matrix` at the
    `which` end by thick-restart Lanczos with full reorthogonalisation.
    Returns (values, residuals, eigenvectors Vector n x k or None, info dict)."""
    if matrix not in MATRICES:
        raise WarrenError(11, f'unknown matrix "{matrix}"')
    if which not in WHICH:
        raise WarrenError(11, f'unknown which "{which}"')
    if precision not in ("f32", "f64"):
        raise WarrenError(11, f'unknown precision "{precision}"')
    opts = _lib.EigenOpts(MATRICES[matrix], WHICH[which], -1 if weighted is None else int(weighted),
                          int(bool(symmetrize)), int(k), int(basis_size or 0), float(tol),
                          int(max_iter or 0), DTYPES[precision], 0, int(seed))
    vals = np.zeros(max(int(k), 1), np.float64)
    res = np.zeros(max(int(k), 1), np.float64)
    vec = Vector(ctx, dtype=precision, id_space=graph.n, width=int(k)) if vectors and k > 0 else None
    info = _lib.EigenInfo()
    code = ctx.lib.wg_eigen(ctx.h, graph.h, C.byref(opts), _p(vals), _p(res),
                            None if vec is None else vec.h, C.byref(info))
    try:
        ctx.check(code)
    except WarrenError as err:
        err.working_set = info.working_set  # set when refused at admission
        raise
    return vals[:k], res[:k], vec, {"iterations": info.iterations, "components": info.components,
                                    "converged": bool(info.converged), "exact": info.exact,
                                    "restarts": info.restarts, "basis_size": info.basis_size,
                                    "working_set": info.working_set}


def sym_eigen(ctx, a):
    """The eigenpairs of the symmetric matrix a (m x m, m <= 256) solved on the
    device, as each Rayleigh-Ritz step of eigen solves its projected problem
    (spectral/ritz.rs:34-52): returns (values, vectors), vectors[:, l] the unit
    eigenvector of values[l], in the solver's order."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    m = a.shape[0]
    if a.shape != (m, m):
        raise WarrenError(11, "a must be square")
    vals = np.zeros(max(m, 1), np.float64)
    vecs = np.zeros((max(m, 1), max(m, 1)), np.float64)
    ctx.check(ctx.lib.wg_sym_eigen(ctx.h, m, _p(a), _p(vals), _p(vecs)))
    return vals[:m], vecs[:m, :m]


def msbfs(ctx, graph, sources, *, direction="out", max_depth=None, depths=True, reach=True,
          farness=True):
    """Multi-source BFS (compute/centrality/multi_source.rs:34-193, up to 512
    sources): returns (depths Vector u32 n x ns or None, reach Vector u32 or
    None, farness Vector i64 or None, levels)."""
    src = _starts(sources)
    ns = len(src)
    dv = Vector(ctx, dtype="u32", id_space=graph.n, width=ns) if depths else None
    rv = Vector(ctx, dtype="u32", id_space=graph.n) if reach else None
    fv = Vector(ctx, dtype="i64", id_space=graph.n) if farness else None
    lv = C.c_uint32()
    ctx.check(ctx.lib.wg_msbfs(ctx.h, graph.h, DIRECTIONS[direction], _p(src), ns,
                               -1 if max_depth is None else max_depth,
                               None if dv is None else dv.h, None if rv is None else rv.h,
                               None if fv is None else fv.h, C.byref(lv)))
    return dv, rv, fv, lv.value


def set_spmv_variant(ctx, variant):
    """Forces an SpMV launch variant (0..31: bits direct, non-temporal, XCD order,
    hot-first gathers, LDS hot-column cache; 32: the tiled kernel), or -1 to
    autotune."""
    ctx.lib.wg_set_spmv_variant(ctx.h, variant)


def set_spmm_shape(ctx, shape):
    """Forces an SpMM launch shape (lanes per row | 256 for 16-byte column
    groups), or -1 to autotune."""
    ctx.lib.wg_set_spmm_shape(ctx.h, shape)


def spmm_shape(graph, semiring, dtype, width, *, transpose=False, weighted=True):
    """The SpMM autotuner's shape for this graph orientation and width, or -1:
    (lanes per row, 16-byte column groups)."""
    s = graph.ctx.lib.wg_spmm_shape(graph.h, int(transpose), DTYPES[dtype], semiring_index(semiring),
                                    int(weighted), int(width))
    return None if s < 0 else (s & 255, bool(s & 256))


def spmv_variant(graph, semiring, dtype, *, transpose=False, weighted=True):
    """The autotuner's choice for this graph orientation, or -1."""
    return graph.ctx.lib.wg_spmv_variant(graph.h, int(transpose), DTYPES[dtype],
                                         semiring_index(semiring), int(weighted))


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, symmetrize=False, threads=0):
    """A synthetic R-MAT graph's CSR (offsets u64, targets u32): self loops and
    duplicate edges removed, ids not permuted."""
    lib = _lib.load()
    n = 1 << scale
    cap = (edge_factor << scale) * (2 if symmetrize else 1)
    off = np.zeros(n + 1, np.uint64)
    tgt = np.zeros(cap, np.uint32)
    nnz = np.zeros(1, np.uint64)
    code = lib.wg_rmat_csr(scale, edge_factor, a, b, c, seed, int(symmetrize), _p(off), _p(tgt),
                           cap, _p(nnz), threads)
    if code:
        raise WarrenError(code, "rmat")
    return off, tgt[:int(nnz[0])].copy()


def rmat_edges(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, threads=0):
    """The raw R-MAT edge list ``rmat`` starts from, in generation order: self
    loops and parallel edges included. Returns (sources, targets), u32."""
    lib = _lib.load()
    m = edge_factor << scale
    src = np.zeros(m, np.uint32)
    dst = np.zeros(m, np.uint32)
    code = lib.wg_rmat_edges(scale, edge_factor, a, b, c, seed, _p(src), _p(dst), threads)
    if code:
        raise WarrenError(code, "rmat_edges")
    return src, dst


WEIGHT_KINDS = {"uniform01": 0, "int": 1, "lognormal": 2}


def edge_weights(offsets, targets, kind="uniform01", seed=7, p0=0.0, p1=1.0, threads=0):
    """Per-edge weights as a function of (seed, source, target)."""
    lib = _lib.load()
    offsets = np.ascontiguousarray(offsets, np.uint64)
    targets = np.ascontiguousarray(targets, np.uint32)
    w = np.zeros(len(targets), np.float64)
    code = lib.wg_edge_weights(len(offsets) - 1, _p(offsets), _p(targets), seed,
                               WEIGHT_KINDS[kind], p0, p1, _p(w), threads)
    if code:
        raise WarrenError(code, "edge_weights")
    return w
