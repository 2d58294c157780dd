"""wg_spmv_noalloc (the product that allocates nothing) against wg_spmv, through
the C ABI: the same bits on integer-valued input for every semiring and
dtype, on two calls in a row, the second over a smaller graph (the static
partials keep the first call's contents), with no device memory charged to
the context by either call."""
import ctypes as C

import numpy as np
import pytest

import pg_warren_amd as W
from pg_warren_amd import _lib

pytestmark = pytest.mark.gpu

WG_OK, WG_MEMORY_LIMIT, WG_UNSUPPORTED = 0, 5, 15


def hub_graph(n, m, seed):
    """Random edges plus a hub row and column longer than one 2048-entry
    block, so split rows leave partials in the scratch."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([rng.integers(0, n, m), np.zeros(n - 1, np.int64), np.arange(1, n)])
    dst = np.concatenate([rng.integers(0, n, m), np.arange(1, n), np.full(n - 1, 7)])
    key = np.unique(src * n + dst)
    src, dst = key // n, key % n
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    w = rng.integers(1, 4, len(dst)).astype(np.float64)
    return off, dst.astype(np.uint32), w


def values(dtype, n, rng):
    if dtype in ("f32", "f64"):
        return rng.integers(0, 5, n).astype(W.NP_DTYPES[dtype])
    if dtype == "i64":
        return rng.integers(-3, 6, n).astype(np.int64)
    if dtype == "u32":
        return rng.integers(0, 6, n).astype(np.uint32)
    return (rng.random(n) < 0.5).astype(np.uint8)


def noalloc(ctx, g, prod, x, y):
    p = prod.c()
    return ctx.lib.wg_spmv_noalloc(ctx.h, g.h, C.byref(p), x.h, y.h)


@pytest.mark.parametrize("dtype", ["f32", "f64", "i64", "bool", "u32"])
def test_noalloc_matches_spmv_on_two_calls(dtype):
    ctx = W.Context(0)
    rng = np.random.default_rng(3)
    graphs = [hub_graph(6000, 60000, 1), hub_graph(3000, 20000, 2)]  # the second smaller
    for off, tgt, w in graphs:
        n = len(off) - 1
        g = W.Graph(ctx, n, off, tgt, w)
        for sr in W.SEMIRINGS:
            if dtype == "u32" and sr in ("max_plus", "max_times", "max_second"):
                continue
            x = values(dtype, n, rng)
            for transpose in (False, True):
                prod = W.Product(sr, transpose=transpose)
                xv = W.Vector(ctx, x, dtype=dtype)
                want = W.spmv(ctx, g, prod, xv, W.Vector(ctx, dtype=dtype, id_space=n)).numpy()
                yv = W.Vector(ctx, dtype=dtype, id_space=n)
                before = ctx.memory_in_use()
                assert noalloc(ctx, g, prod, xv, yv) == WG_OK, (sr, dtype, transpose)
                assert ctx.memory_in_use() == before
                assert yv.numpy().tobytes() == want.tobytes(), (sr, dtype, transpose, n)
        g.close()
    ctx.close()


def test_noalloc_casts_weights_itself_and_refuses_what_would_allocate():
    """On a fresh graph (no weight cast cached) the f32 product casts the f64
    weights per entry, with wg_spmv's bits and its overflow error; relation
    selections and split rows beyond the scratch are refused."""
    ctx = W.Context(0)
    off, tgt, w = hub_graph(4000, 30000, 5)
    n = len(off) - 1
    g = W.Graph(ctx, n, off, tgt, w)
    x = np.random.default_rng(6).integers(0, 5, n).astype(np.float32)
    xv = W.Vector(ctx, x)
    yv = W.Vector(ctx, dtype="f32", id_space=n)
    before = ctx.memory_in_use()
    assert noalloc(ctx, g, W.Product("plus_times"), xv, yv) == WG_OK
    assert ctx.memory_in_use() == before
    got = yv.numpy()
    want = W.spmv(ctx, g, W.Product("plus_times"), xv, W.Vector(ctx, dtype="f32", id_space=n)).numpy()
    assert got.tobytes() == want.tobytes()
    # An f64 weight beyond f32's range: the same error kind as wg_spmv.
    big = W.Graph(ctx, 2, np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([1e300]))
    x2 = W.Vector(ctx, np.ones(2, np.float32))
    code = noalloc(ctx, big, W.Product("plus_times"), x2, W.Vector(ctx, dtype="f32", id_space=2))
    assert code == 8  # WG_FLOAT_OVERFLOW
    with pytest.raises(W.WarrenError) as e:
        W.spmv(ctx, big, W.Product("plus_times"), x2, W.Vector(ctx, dtype="f32", id_space=2))
    assert e.value.name == "float_overflow"
    # A relation selection would be built (allocated).
    code = noalloc(ctx, g, W.Product("plus_times", relations=[0]), xv, yv)
    assert code == WG_UNSUPPORTED
    # Split rows whose partials exceed the 1 MiB scratch: i64 plus keeps 48
    # bytes a piece; 11,000 rows of 4096 entries split into 22,000 pieces
    # (1.06 MB), while a 4-byte f32 partial needs 88 KB.
    rows, width = 11_000, 4096
    off3 = (np.arange(rows + 1, dtype=np.uint64) * width)
    tgt3 = np.tile(np.arange(width, dtype=np.uint32), rows)
    g3 = W.Graph(ctx, rows, off3, tgt3)
    xi = W.Vector(ctx, np.ones(rows, np.int64))
    code = noalloc(ctx, g3, W.Product("plus_times"), xi, W.Vector(ctx, dtype="i64", id_space=rows))
    assert code == WG_MEMORY_LIMIT
    xf = W.Vector(ctx, np.arange(rows, dtype=np.float32) % 7)
    y3 = W.Vector(ctx, dtype="f32", id_space=rows)
    assert noalloc(ctx, g3, W.Product("min_plus"), xf, y3) == WG_OK
    want3 = W.spmv(ctx, g3, W.Product("min_plus"), xf, W.Vector(ctx, dtype="f32", id_space=rows)).numpy()
    assert y3.numpy().tobytes() == want3.tobytes()
    g3.close()
    g.close()
    ctx.close()
