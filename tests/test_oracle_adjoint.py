"""The oracle's PageRank adjoint against what the mathematics fixes: central
finite differences of the oracle's own forward PageRank (itself pinned to
the reference's golden ranks), the zero gradient of a graph whose ranks do
not depend on its weights, and the scale invariance of each row (the ranks
see a row's weights only through w / s, so sum_e w_e dL/dw_e = 0 per row)."""
import numpy as np
import pytest

import oracle as O


def random_graph(rng, n=24, m=90, dangling=3):
    src = rng.integers(0, n - dangling, m)
    dst = rng.integers(0, n, m)
    w = rng.uniform(0.5, 2.0, m)
    return O.csr_from_edges(n, src, dst, w)


def loss_and_grad(x, targets, coef):
    return float(np.sum(coef * x[targets] ** 2)), targets, 2.0 * coef * x[targets]


def forward(csr, **kw):
    x, _ = O.pagerank(csr, tol=None, max_iter=500, **kw)
    return x


@pytest.mark.parametrize("variant", ["uniform", "seeded", "tombstoned"])
def test_adjoint_matches_central_differences(oracle, variant):
    rng = np.random.default_rng({"uniform": 1, "seeded": 2, "tombstoned": 3}[variant])
    csr = random_graph(rng)
    n = csr.n
    kw = {}
    if variant == "seeded":
        seed = np.zeros(n)
        seed[rng.choice(n, 4, replace=False)] = rng.uniform(0.5, 2.0, 4)
        kw["seed"] = seed
    if variant == "tombstoned":
        tomb = np.zeros(n, np.uint8)
        tomb[rng.choice(n, 3, replace=False)] = 1
        kw["tombstones"] = tomb
    targets = rng.choice(n, 8, replace=False)
    coef = rng.uniform(0.5, 1.5, 8)
    x = forward(csr, **kw)
    g = np.zeros(n)
    g[targets] = 2.0 * coef * x[targets]
    dw, lam, it = O.pagerank_backward(csr, x, g, tol=None, max_iter=500, **kw)
    assert it == 500
    h = 1e-6
    fd = np.zeros(csr.nnz)
    for e in range(csr.nnz):
        wp = csr.weights.copy()
        wp[e] += h
        wm = csr.weights.copy()
        wm[e] -= h
        xp = forward(O.Csr(n, csr.offsets, csr.targets, wp), **kw)
        xm = forward(O.Csr(n, csr.offsets, csr.targets, wm), **kw)
        fd[e] = (np.sum(coef * xp[targets] ** 2) - np.sum(coef * xm[targets] ** 2)) / (2 * h)
    scale = np.abs(dw).max()
    assert scale > 1e-6
    np.testing.assert_allclose(dw, fd, rtol=0, atol=1e-7 * scale)


def test_rows_are_scale_invariant(oracle):
    rng = np.random.default_rng(7)
    csr = random_graph(rng, n=40, m=200)
    x = forward(csr)
    g = rng.normal(size=csr.n)
    dw, _, _ = O.pagerank_backward(csr, x, g, tol=None, max_iter=500)
    off = csr.offsets.astype(np.int64)
    for a in range(csr.n):
        b, e = off[a], off[a + 1]
        if e > b:
            s = np.dot(csr.weights[b:e], dw[b:e])
            assert abs(s) <= 1e-12 * max(1.0, np.abs(dw).max())


def test_two_cycle_has_no_weight_gradient(oracle):
    # 0 <-> 1: both ranks are 1/2 whatever the weights.
    csr = O.csr_from_edges(2, [0, 1], [1, 0], [3.0, 0.25])
    x = forward(csr)
    np.testing.assert_allclose(x, [0.5, 0.5], rtol=0, atol=1e-15)
    dw, lam, _ = O.pagerank_backward(csr, x, np.array([1.0, -2.0]), tol=None, max_iter=200)
    np.testing.assert_allclose(dw, 0.0, atol=1e-15)


def test_adjoint_fixed_point_closed_form(oracle):
    # A single node with a self loop (weight 2): s = 2, P = 1, so
    # lambda = d * lambda + g  =>  lambda = g / (1 - d).
    csr = O.csr_from_edges(1, [0], [0], [2.0])
    x = forward(csr)
    _, lam, _ = O.pagerank_backward(csr, x, np.array([0.3]), tol=None, max_iter=400)
    assert abs(lam[0] - 0.3 / 0.15) < 1e-12


def test_unweighted_graph_is_refused(oracle):
    csr = O.csr_from_edges(3, [0, 1], [1, 2])
    with pytest.raises(O.OracleError):
        O.pagerank_backward(csr, np.ones(3) / 3, np.ones(3))
