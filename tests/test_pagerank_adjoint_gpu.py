"""The HIP PageRank adjoint (wg_pagerank_backward) against the oracle's FP64
adjoint, central finite differences of the HIP forward PageRank (weights
moved in place with wg_graph_update_weights), and the per-row scale
invariance sum_e w_e dL/dw_e = 0. The adjoint's sums run in another order
than the oracle's, so parity is to a relative 1e-9 of the largest entry."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def random_csr(oracle, rng, n, m, dangling=0.1):
    src = rng.integers(0, int(n * (1 - dangling)), m)
    dst = rng.integers(0, n, m)
    w = rng.lognormal(0.0, 1.0, m)
    return oracle.csr_from_edges(n, src, dst, w)


@pytest.mark.parametrize("variant", ["uniform", "seeded", "tombstoned"])
@pytest.mark.parametrize("n,m", [(50, 300), (3000, 40000)])
def test_matches_oracle_adjoint(ctx, oracle, variant, n, m):
    rng = np.random.default_rng(n + len(variant))
    csr = random_csr(oracle, rng, n, m)
    kw = {}
    if variant == "seeded":
        seed = np.zeros(n)
        seed[rng.choice(n, 5, replace=False)] = rng.uniform(0.5, 2.0, 5)
        kw["seed"] = seed
    if variant == "tombstoned":
        tomb = np.zeros(n, np.uint8)
        tomb[rng.choice(n, max(1, n // 20), replace=False)] = 1
        kw["tombstones"] = tomb
    g = W.Graph(ctx, n, csr.offsets, csr.targets, csr.weights)
    x, _ = W.pagerank(ctx, g, tol=None, max_iter=200, **kw)
    xs = x.numpy()
    gx = np.zeros(n)
    t = rng.choice(n, min(n, 40), replace=False)
    gx[t] = 2.0 * xs[t]
    dw, it, lam = W.pagerank_backward(ctx, g, x, gx, tol=None, max_iter=200, adjoint=True, **kw)
    assert it == 200
    odw, olam, oit = oracle.pagerank_backward(csr, xs, gx, tol=None, max_iter=200, **kw)
    assert oit == 200
    got = dw.numpy()
    scale = np.abs(odw).max()
    assert np.abs(got - odw).max() <= 1e-9 * scale
    assert np.abs(lam.numpy() - olam).max() <= 1e-9 * np.abs(olam).max()


def test_tolerance_stops_like_the_oracle(ctx, oracle):
    rng = np.random.default_rng(11)
    csr = random_csr(oracle, rng, 2000, 20000)
    g = W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)
    x, _ = W.pagerank(ctx, g, tol=1e-12, max_iter=300)
    gx = rng.normal(size=csr.n) * 1e-3
    dw, it = W.pagerank_backward(ctx, g, x, gx, tol=1e-12, max_iter=300)
    odw, _, oit = oracle.pagerank_backward(csr, x.numpy(), gx, tol=1e-12, max_iter=300)
    assert abs(it - oit) <= 1
    assert np.abs(dw.numpy() - odw).max() <= 1e-8 * np.abs(odw).max()


def test_central_differences_on_device(ctx, oracle):
    rng = np.random.default_rng(5)
    n = 400
    csr = random_csr(oracle, rng, n, 3000)
    g = W.Graph(ctx, n, csr.offsets, csr.targets, csr.weights)
    targets = rng.choice(n, 30, replace=False)

    def loss():
        x, _ = W.pagerank(ctx, g, tol=None, max_iter=300)
        xs = x.numpy()
        return float(np.sum(xs[targets] ** 2)), x, xs

    _, x, xs = loss()
    gx = np.zeros(n)
    gx[targets] = 2.0 * xs[targets]
    dw, _ = W.pagerank_backward(ctx, g, x, gx, tol=None, max_iter=300)
    dw = dw.numpy()
    h = 1e-6
    edges = np.concatenate([np.argsort(-np.abs(dw))[:20], rng.choice(csr.nnz, 20, replace=False)])
    scale = np.abs(dw).max()
    for e in edges:
        w0 = csr.weights[e]
        g.update_weights([e], [w0 + h])
        lp, _, _ = loss()
        g.update_weights([e], [w0 - h])
        lm, _, _ = loss()
        g.update_weights([e], [w0])
        fd = (lp - lm) / (2 * h)
        assert abs(fd - dw[e]) <= 1e-6 * scale, (e, fd, dw[e])


def test_rows_are_scale_invariant(ctx, oracle):
    rng = np.random.default_rng(9)
    csr = random_csr(oracle, rng, 5000, 60000)
    g = W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)
    x, _ = W.pagerank(ctx, g, tol=None, max_iter=200)
    dw, _ = W.pagerank_backward(ctx, g, x, rng.normal(size=csr.n), tol=None, max_iter=200)
    dw = dw.numpy()
    off = csr.offsets.astype(np.int64)
    rows = np.repeat(np.arange(csr.n), np.diff(off))
    per_row = np.bincount(rows, weights=csr.weights * dw, minlength=csr.n)
    assert np.abs(per_row).max() <= 1e-11 * max(1.0, np.abs(dw).max())


def test_update_weights_reaches_every_orientation(ctx, oracle):
    rng = np.random.default_rng(3)
    csr = random_csr(oracle, rng, 300, 2000)
    g = W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)
    x = W.Vector(ctx, rng.uniform(0, 1, csr.n).astype(np.float32))
    y = W.Vector(ctx, dtype="f32", id_space=csr.n)
    W.spmv(ctx, g, W.Product("plus_times", transpose=True), x, y)  # casts the f32 weights
    edges = rng.choice(csr.nnz, 50, replace=False)
    neww = rng.uniform(0.1, 3.0, 50)
    g.update_weights(edges, neww)
    w2 = csr.weights.copy()
    w2[edges] = neww
    c2 = oracle.Csr(csr.n, csr.offsets, csr.targets, w2)
    got, it = W.pagerank(ctx, g, tol=1e-10)
    want, oit = oracle.pagerank(c2, tol=1e-10)
    assert it == oit
    assert np.abs(got.numpy() - want).sum() <= 1e-9
    W.spmv(ctx, g, W.Product("plus_times", transpose=True), x, y)
    want_y = oracle.spmv(c2, x.numpy(), "plus_times", rev=c2.transpose(), transpose=True)
    np.testing.assert_allclose(y.numpy(), want_y, rtol=1e-5)
    off, tgt, rw = g.reverse()
    rc = c2.transpose()
    assert np.array_equal(rw, rc.weights)


def test_errors(ctx, oracle):
    csr = oracle.csr_from_edges(4, [0, 1, 2], [1, 2, 3])
    g = W.Graph(ctx, 4, csr.offsets, csr.targets)
    x, _ = W.pagerank(ctx, g)
    with pytest.raises(W.WarrenError):
        W.pagerank_backward(ctx, g, x, np.ones(4))
    wg = W.Graph(ctx, 4, csr.offsets, csr.targets, np.ones(3))
    with pytest.raises(W.WarrenError):
        wg.update_weights([3], [1.0])
    with pytest.raises(W.WarrenError):
        W.pagerank_backward(ctx, wg, x, np.ones(5))


def test_config_shape_rmat_with_hub_rows(ctx, oracle):
    """The differentiable-PageRank workload's shape at scale 16: R-MAT with
    log-normal(0, 1) weights, row-normalised; PageRank personalised from 32
    seeds; loss = sum of squared ranks over 1,000 targets; 200 iterations
    each way. Its hub rows exceed 2048 entries in both directions, so the
    forward and adjoint products fold split rows piece by piece. Checked
    against the oracle's FP64 adjoint and against central differences of the
    device's own forward run (h = 1e-6) on 20 edges, a hub row's among them."""
    off, tgt = W.rmat(16, 16, seed=4)
    n, m = len(off) - 1, len(tgt)
    deg = np.diff(off.astype(np.int64))
    indeg = np.bincount(tgt, minlength=n)
    assert deg.max() > 2048 and indeg.max() > 2048
    w = W.edge_weights(off, tgt, "lognormal", seed=4, p0=0.0, p1=1.0)
    rows = np.repeat(np.arange(n), deg)
    w = w / np.bincount(rows, weights=w, minlength=n)[rows]
    csr = oracle.Csr(n, off, tgt, w)
    g = W.Graph(ctx, n, off, tgt, w)
    rng = np.random.default_rng(20)
    seed = np.zeros(n)
    seed[rng.choice(np.flatnonzero(deg > 0), 32, replace=False)] = 1.0
    targets = rng.choice(n, 1000, replace=False)
    sv = W.Vector(ctx, seed)
    out = W.Vector(ctx, dtype="f64", id_space=n)

    def forward():
        W.pagerank(ctx, g, tol=None, max_iter=200, seed=sv, out=out)
        xs = out.numpy()
        return float(np.sum(xs[targets] ** 2)), xs

    _, x = forward()
    ox, oit = oracle.pagerank(csr, tol=None, max_iter=200, seed=seed)
    assert oit == 200
    assert np.abs(x - ox).sum() <= 1e-12
    gx = np.zeros(n)
    gx[targets] = 2.0 * x[targets]
    dwv, it = W.pagerank_backward(ctx, g, x, gx, tol=None, max_iter=200, seed=seed)
    assert it == 200
    dw = dwv.numpy()
    odw, _, oit = oracle.pagerank_backward(csr, x, gx, tol=None, max_iter=200, seed=seed)
    assert oit == 200
    scale = np.abs(odw).max()
    assert scale > 0
    assert np.abs(dw - odw).max() <= 1e-9 * scale
    hub = int(np.argmax(deg))
    hub_edges = np.arange(int(off[hub]), int(off[hub + 1]), dtype=np.int64)
    edges = np.concatenate([hub_edges[np.argsort(-np.abs(dw[hub_edges]))[:5]],
                            np.argsort(-np.abs(dw))[:10], rng.choice(m, 5, replace=False)])
    h = 1e-6
    for e in edges:
        w0 = float(w[e])
        g.update_weights([e], [w0 + h])
        lp, _ = forward()
        g.update_weights([e], [w0 - h])
        lm, _ = forward()
        g.update_weights([e], [w0])
        fd = (lp - lm) / (2 * h)
        assert abs(fd - dw[e]) <= 1e-7 * scale, (int(e), fd, dw[e])
    for v in (sv, out, dwv):
        v.close()
    g.close()
