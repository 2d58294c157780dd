"""The plain-method workload at its own size (R-MAT scale 22: n = 4,194,304,
m = 65,244,064), on the benchmark's inputs, against the oracle:

* plus_times over f32 weights in (0, 1] with f32 accumulation (the tiled
  kernel): within the rounding bound of the oracle's f32 left fold,
  2 * len * eps32 * sum|terms| per row (compute/product/kernel.rs:100-109);
* min_plus over u32 weights 1..255 (10% of x "infinite") and or_and over
  bool: bit for bit;
* BFS from 4 sources with an out-edge: depths equal to the oracle's;
* PageRank, damping 0.85, tol 1e-10, at most 100 iterations, f64: the same
  iteration count and an L1 distance of at most 1e-9 from the oracle's.
"""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

EPS32 = float(np.finfo(np.float32).eps)


def test_scale22_products_bfs_and_pagerank_match_the_oracle(oracle):
    off, tgt = W.rmat(22, 16, seed=1)
    n, m = len(off) - 1, len(tgt)
    assert n == 1 << 22 and m > 60_000_000
    wf = W.edge_weights(off, tgt, "uniform01", seed=1)
    wi = W.edge_weights(off, tgt, "int", seed=2, p0=1, p1=255)
    rng = np.random.default_rng(5)
    xf = ((rng.integers(0, 2**24, n) + 1) / 2.0**24).astype(np.float32)
    xu = rng.integers(0, 2**24, n).astype(np.uint32)
    xu[rng.random(n) < 0.1] = 2**32 - 1
    xb = (rng.random(n) < 0.5).astype(np.uint8)
    threads = 16
    ref_f = oracle.Csr(n, off, tgt, wf)
    ref_i = oracle.Csr(n, off, tgt, wi)

    ctx = W.Context(0)
    gf = W.Graph(ctx, n, off, tgt, wf)
    gi = W.Graph(ctx, n, off, tgt, wi)

    def spmv(g, sr, x, dtype):
        return W.spmv(ctx, g, W.Product(sr), W.Vector(ctx, x, dtype=dtype),
                      W.Vector(ctx, dtype=dtype, id_space=n)).numpy()

    got = spmv(gf, "plus_times", xf, "f32").astype(np.float64)
    assert W.spmv_variant(gf, "plus_times", "f32") == 32  # the tiled kernel
    want = oracle.spmv(ref_f, xf, "plus_times", "f32", threads=threads).astype(np.float64)
    abssum = oracle.spmv(ref_f, xf.astype(np.float64), "plus_times", "f64", threads=threads)
    deg = np.diff(off.astype(np.int64))
    assert np.all(np.abs(got - want) <= 2 * deg * EPS32 * abssum + 1e-30)
    assert spmv(gi, "min_plus", xu, "u32").tobytes() == \
        oracle.spmv(ref_i, xu, "min_plus", "u32", threads=threads).tobytes()
    assert spmv(gf, "lor_land", xb, "bool").tobytes() == \
        oracle.spmv(ref_f, xb, "lor_land", "bool", threads=threads).tobytes()

    sources = np.random.default_rng(64).choice(np.flatnonzero(deg > 0), 4, replace=False)
    depth = W.Vector(ctx, dtype="u32", id_space=n)
    via = W.Vector(ctx, dtype="u32", id_space=n)
    for s in sources:
        W.bfs_levels(ctx, gf, [int(s)], depth=depth, via=via)
        assert np.array_equal(depth.numpy(), oracle.bfs_depths(ref_f, [int(s)])[0]), int(s)

    ranks, it = W.pagerank(ctx, gf, tol=1e-10, max_iter=100, weighted=False)
    pr_ref, it_ref = oracle.pagerank(oracle.Csr(n, off, tgt), tol=1e-10, max_iter=100, threads=threads)
    assert it == it_ref
    assert np.abs(ranks.numpy() - pr_ref).sum() <= 1e-9
    ctx.close()
