"""ctypes binding of libwarren_gpu.so (the C ABI in include/warren_gpu.h).

The library is built in-tree by pg_warren_amd/build.py. Loading fails loudly
when it is missing: there is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwarren_gpu.so")

P = C.c_void_p
U64 = C.c_uint64
U32 = C.c_uint32
I32 = C.c_int
I64 = C.c_int64
D = C.c_double

_LIB = None


class Product(C.Structure):
    """wg_product: compute/product/mod.rs:33-56."""
    _fields_ = [("semiring", C.c_int32), ("transpose", C.c_int32), ("weighted", C.c_int32),
                ("mask_complement", C.c_int32), ("mask", P), ("replace", C.c_int32),
                ("accumulate", C.c_int32), ("relations", P), ("nrelations", C.c_uint32),
                ("per_relation", C.c_int32)]


class CsrSegments(C.Structure):
    """wg_csr_segments: graph/segment/csr.rs:22-29."""
    _fields_ = [("rows", P), ("rows_len", C.c_uint64), ("offsets", P), ("offsets_len", C.c_uint64),
                ("targets", P), ("targets_len", C.c_uint64), ("weights", P),
                ("weights_len", C.c_uint64)]


class GatherOpts(C.Structure):
    """wg_gather_opts: compute/gather/mod.rs:48-64."""
    _fields_ = [("combiner", C.c_int32), ("direction", C.c_int32), ("relations", P),
                ("nrelations", C.c_uint32), ("per_relation", C.c_int32), ("mask", P),
                ("mask_complement", C.c_int32), ("include_self", C.c_int32)]


class PagerankOpts(C.Structure):
    """wg_pagerank_opts: compute/pagerank/mod.rs:27-47."""
    _fields_ = [("damping", C.c_double), ("tol", C.c_double), ("has_tol", C.c_int32),
                ("max_iter", C.c_uint32), ("weighted", C.c_int32), ("pad", C.c_int32)]


class Visit(C.Structure):
    """wg_visit: traversal/bfs.rs:15-23."""
    _fields_ = [("node", C.c_uint64), ("via", C.c_uint64), ("depth", C.c_uint32),
                ("pad", C.c_uint32)]


class TraversalStats(C.Structure):
    """wg_traversal_stats: traversal/stats.rs:12-20."""
    _fields_ = [("nodes_visited", C.c_uint64), ("edges_traversed", C.c_uint64),
                ("truncated", C.c_uint32), ("levels_pulled", C.c_uint32)]


class EigenOpts(C.Structure):
    """wg_eigen_opts: EigenOptions (spectral/lanczos.rs:140-156) and
    GraphMatrix (spectral/operator.rs:79-93)."""
    _fields_ = [("matrix", C.c_int32), ("which", C.c_int32), ("weighted", C.c_int32),
                ("symmetrize", C.c_int32), ("k", C.c_uint32), ("basis_size", C.c_uint32),
                ("tol", C.c_double), ("max_iter", C.c_uint64), ("precision", C.c_int32),
                ("pad", C.c_int32), ("seed", C.c_uint64)]


class EigenInfo(C.Structure):
    """wg_eigen_info: Eigen (spectral/lanczos.rs:174-198)."""
    _fields_ = [("iterations", C.c_uint64), ("components", C.c_uint64),
                ("converged", C.c_uint32), ("exact", C.c_uint32), ("restarts", C.c_uint32),
                ("basis_size", C.c_uint32), ("working_set", C.c_uint64)]


PHASE_FN = C.CFUNCTYPE(None, P, C.c_char_p, C.c_int64)
DONE_FN = C.CFUNCTYPE(None, P, C.c_uint64)
DETAIL_FN = C.CFUNCTYPE(None, P, C.c_char_p, C.c_double)


class ProgressPort(C.Structure):
    """wg_progress: the Progress port (progress.rs:9-16)."""
    _fields_ = [("phase", PHASE_FN), ("done", DONE_FN), ("detail", DETAIL_FN), ("user", P)]


# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    ("wg_context_create", I32, [I32, U64, P]),
    ("wg_context_destroy", None, [P]),
    ("wg_last_error", I32, [P, C.c_char_p, C.c_size_t]),
    ("wg_cancel", None, [P]),
    ("wg_cancel_reset", None, [P]),
    ("wg_set_progress", None, [P, P]),
    ("wg_memory_in_use", U64, [P]),
    ("wg_set_memory_limit", I32, [P, U64]),
    ("wg_synchronize", I32, [P]),
    ("wg_set_blocking", None, [P, I32]),
    ("wg_set_graph_launch", None, [P, I32]),
    ("wg_stream", P, [P]),
    ("wg_graph_upload", I32, [P, U64, P, P, P, P]),
    ("wg_graph_upload_u64", I32, [P, U64, P, P, P, P]),
    ("wg_graph_read_u64", I32, [P, P, I32, P, P, P]),
    ("wg_graph_destroy", None, [P]),
    ("wg_graph_info", I32, [P, P, P, P]),
    ("wg_graph_read_reverse", I32, [P, P, P, P, P]),
    ("wg_graph_read_forward", I32, [P, P, P, P, P]),
    ("wg_graph_select", I32, [P, P, P, U32, I32, P]),
    ("wg_gather", I32, [P, P, P, I32, U64, P, P, P]),
    ("wg_gathered_info", I32, [P, P, P]),
    ("wg_gathered_read", I32, [P, P, P, P, P]),
    ("wg_gather_bytes", I32, [P, P, P, U64, P, P, P, P]),
    ("wg_gathered_bytes_info", I32, [P, P]),
    ("wg_gathered_read_bytes", I32, [P, P, P, P]),
    ("wg_gathered_destroy", None, [P]),
    ("wg_graph_load_segments", I32, [P, U64, P, P, P]),
    ("wg_csr_segment_lengths", I32, [U64, U64, P, P, P]),
    ("wg_graph_encode_segments", I32, [P, P, I32, P, P, P]),
    ("wg_relation_segment_lengths", I32, [P, P, U32, I32, P, P, P, P]),
    ("wg_relation_encode_segments", I32, [P, P, U32, I32, P, P, P, P]),
    ("wg_host_alloc", I32, [P, U64, P]),
    ("wg_host_free", None, [P]),
    ("wg_bfs_cursor_create", I32, [P, P, I32, P, U32, P, U64, P, C.c_int64, C.c_int64, I32, P]),
    ("wg_bfs_cursor_next", I32, [P, P, P, P, P]),
    ("wg_bfs_cursor_stats", I32, [P, P]),
    ("wg_bfs_cursor_destroy", None, [P]),
    ("wg_graph_relation_count", I32, [P, P]),
    ("wg_graph_build", I32, [P, U64, U64, P, P, P, P, U32, I32, P, P]),
    ("wg_graph_update_weights", I32, [P, P, P, P, U64]),
    ("wg_vector_create", I32, [P, I32, U64, U64, P, P]),
    ("wg_vector_destroy", None, [P]),
    ("wg_vector_write", I32, [P, P, P]),
    ("wg_vector_read", I32, [P, P, P]),
    ("wg_vector_info", I32, [P, P, P, P]),
    ("wg_vector_device_ptr", P, [P]),
    ("wg_spmv", I32, [P, P, P, P, P]),
    ("wg_spmv_noalloc", I32, [P, P, P, P, P]),
    ("wg_spmm", I32, [P, P, P, P, P]),
    ("wg_set_spmv_variant", None, [P, I32]),
    ("wg_spmv_variant", I32, [P, I32, I32, I32, I32]),
    ("wg_set_spmm_shape", None, [P, I32]),
    ("wg_spmm_shape", I32, [P, I32, I32, I32, I32, U64]),
    ("wg_pagerank", I32, [P, P, P, P, P, P, P]),
    ("wg_pagerank_batch", I32, [P, P, P, P, P, P, P, U32, P, P]),
    ("wg_pagerank_backward", I32, [P, P, P, P, P, P, P, P, P, P]),
    ("wg_eigen", I32, [P, P, P, P, P, P, P]),
    ("wg_sym_eigen", I32, [P, U32, P, P, P]),
    ("wg_bfs_levels", I32, [P, P, I32, P, U32, P, U64, P, I64, I32, P, P, P]),
    ("wg_bfs", I32, [P, P, I32, P, U32, P, U64, P, I64, I64, P, P, P]),
    ("wg_reachable_set", I32, [P, P, I32, P, U32, P, U64, I64, P, P, P]),
    ("wg_msbfs", I32, [P, P, I32, P, U32, I64, P, P, P, P]),
    ("wg_device_copy", I32, [P, P, P, U64]),
    ("wg_rmat_csr", I32, [U32, U32, D, D, D, U64, I32, P, P, U64, P, I32]),
    ("wg_rmat_edges", I32, [U32, U32, D, D, D, U64, P, P, I32]),
    ("wg_edge_weights", I32, [U64, P, P, U64, I32, D, D, P, I32]),
]


def load():
    """Loads libwarren_gpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python pg_warren_amd/build.py` "
                "(or __graft_entry__.build()); pg_warren_amd has no CPU fallback")
        lib = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB
