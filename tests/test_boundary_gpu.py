"""The C ABI's ownership and cancellation contract, called through ctypes as a
Rust or C adapter would call it.

* wg_context_destroy with live children: their device memory goes with the
  context; destroying them afterwards frees host records only, and any other
  call given one is refused (the reference's borrow rules, which Rust
  enforces at compile time, as run-time checks).
* wg_cancel from a second thread stops a running call (Cancellation::cancel,
  crates/warren-core/src/interrupts.rs:40-70, raised by the thread that waits
  for the pool).
"""
import ctypes as C
import threading
import time

import numpy as np
import pytest

import pg_warren_amd as W
from pg_warren_amd import _lib

pytestmark = pytest.mark.gpu

WG_OK, WG_CANCELLED, WG_INVALID_ARGUMENT = 0, 1, 11
WG_F64, WG_BOOL = 1, 3


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _ring(n):
    off = np.arange(n + 1, dtype=np.uint64)
    tgt = ((np.arange(n) + 1) % n).astype(np.uint32)
    return off, tgt


def test_context_destroy_orphans_live_children():
    """Regression: destroying a context before its graph, vector, cursor and
    gathered result used to leave destroy_graph reading the freed context
    (a segfault at exit once every test had passed)."""
    lib = _lib.load()
    ctx = C.c_void_p()
    assert lib.wg_context_create(0, 0, C.byref(ctx)) == WG_OK
    off, tgt = _ring(64)
    g = C.c_void_p()
    assert lib.wg_graph_upload(ctx, 64, _p(off), _p(tgt), None, C.byref(g)) == WG_OK
    x = np.arange(64, dtype=np.float64)
    v = C.c_void_p()
    assert lib.wg_vector_create(ctx, WG_F64, 64, 1, _p(x), C.byref(v)) == WG_OK
    starts = np.array([0], np.uint64)
    cur = C.c_void_p()
    assert lib.wg_bfs_cursor_create(ctx, g, 0, None, 0, _p(starts), 1, None, -1, -1, 0, C.byref(cur)) == WG_OK
    senders = np.arange(64, dtype=np.uint32)
    msgs = np.ones(64)
    opts = _lib.GatherOpts(2, 1, None, 0, 0, None, 0, 0)
    res = C.c_void_p()
    assert lib.wg_gather(ctx, g, C.byref(opts), WG_F64, 64, _p(senders), _p(msgs), C.byref(res)) == WG_OK
    used = lib.wg_memory_in_use(ctx)
    assert used > 0
    lib.wg_context_destroy(ctx)
    # A second context refuses the orphans and frees nothing of them.
    ctx2 = C.c_void_p()
    assert lib.wg_context_create(0, 0, C.byref(ctx2)) == WG_OK
    out = np.zeros(64)
    assert lib.wg_vector_read(ctx2, v, _p(out)) == WG_INVALID_ARGUMENT
    prod = _lib.Product(0, 0, -1, 0, None, 0, -1, None, 0, 0)
    assert lib.wg_spmv(ctx2, g, C.byref(prod), v, v) == WG_INVALID_ARGUMENT
    visits = (_lib.Visit * 64)()
    count, more = C.c_uint64(), C.c_int()
    assert lib.wg_bfs_cursor_next(ctx2, cur, visits, C.byref(count), C.byref(more)) == WG_INVALID_ARGUMENT
    rec = np.zeros(64, np.uint64)
    assert lib.wg_gathered_read(ctx2, res, _p(rec), None, None) == WG_INVALID_ARGUMENT
    # Destroying the orphans afterwards, in any order, is safe.
    lib.wg_bfs_cursor_destroy(cur)
    lib.wg_graph_destroy(g)
    lib.wg_gathered_destroy(res)
    lib.wg_vector_destroy(v)
    assert lib.wg_memory_in_use(ctx2) == 0
    lib.wg_context_destroy(ctx2)


def test_graph_destroy_releases_its_cursors():
    lib = _lib.load()
    ctx = C.c_void_p()
    assert lib.wg_context_create(0, 0, C.byref(ctx)) == WG_OK
    off, tgt = _ring(32)
    g = C.c_void_p()
    assert lib.wg_graph_upload(ctx, 32, _p(off), _p(tgt), None, C.byref(g)) == WG_OK
    starts = np.array([3], np.uint64)
    cur = C.c_void_p()
    assert lib.wg_bfs_cursor_create(ctx, g, 0, None, 0, _p(starts), 1, None, -1, -1, 0, C.byref(cur)) == WG_OK
    lib.wg_graph_destroy(g)
    visits = (_lib.Visit * 32)()
    count, more = C.c_uint64(), C.c_int()
    assert lib.wg_bfs_cursor_next(ctx, cur, visits, C.byref(count), C.byref(more)) == WG_INVALID_ARGUMENT
    lib.wg_bfs_cursor_destroy(cur)
    assert lib.wg_memory_in_use(ctx) == 0
    lib.wg_context_destroy(ctx)


def test_cancel_from_another_thread_stops_a_running_pagerank():
    """A 100,000-iteration PageRank (no tolerance) on R-MAT scale 20 runs for
    several seconds; wg_cancel from a second thread ends it with
    WG_CANCELLED within a batch of iterations, and the context is usable
    again after wg_cancel_reset."""
    ctx = W.Context(0)
    off, tgt = W.rmat(20, 16, seed=5)
    n = len(off) - 1
    g = W.Graph(ctx, n, off, tgt)
    _, it = W.pagerank(ctx, g, tol=None, max_iter=20)  # warm: builds the graph's buffers
    assert it == 20
    t0 = time.monotonic()
    timer = threading.Timer(0.2, ctx.cancel)
    timer.start()
    with pytest.raises(W.WarrenError) as e:
        W.pagerank(ctx, g, tol=None, max_iter=100_000)
    elapsed = time.monotonic() - t0
    timer.join()
    assert e.value.name == "cancelled"
    # The full run takes several seconds; the cancel lands within a batch.
    assert elapsed < 2.0, elapsed
    ctx.cancel_reset()
    _, it = W.pagerank(ctx, g, tol=None, max_iter=5)
    assert it == 5
    g.close()
    ctx.close()


def test_cancel_word_stops_a_running_batched_pagerank():
    """Batched personalised PageRank polls the context's host-mapped cancel
    word once a work item inside its persistent pass (a word a running
    kernel sees), besides the host's check between iterations: a cancel
    from a second thread ends a long run with WG_CANCELLED, no cut-short
    pass is taken as an iteration, and after wg_cancel_reset the same call
    gives the ranks of an uncancelled run."""
    ctx = W.Context(0)
    off, tgt = W.rmat(18, 16, seed=6)
    n = len(off) - 1
    g = W.Graph(ctx, n, off, tgt)
    rng = np.random.default_rng(3)
    seeds = [rng.choice(n, 1 + (i % 5), replace=False) for i in range(32)]
    want, _ = W.pagerank_batch(ctx, g, seeds, tol=None, max_iter=3)
    want = want.numpy()
    t0 = time.monotonic()
    timer = threading.Timer(0.2, ctx.cancel)
    timer.start()
    with pytest.raises(W.WarrenError) as e:
        W.pagerank_batch(ctx, g, seeds, tol=None, max_iter=100_000)
    elapsed = time.monotonic() - t0
    timer.join()
    assert e.value.name == "cancelled"
    assert elapsed < 2.0, elapsed
    ctx.cancel_reset()
    got, it = W.pagerank_batch(ctx, g, seeds, tol=None, max_iter=3)
    assert np.all(np.asarray(it) == 3)
    assert np.array_equal(got.numpy(), want)
    g.close()
    ctx.close()
