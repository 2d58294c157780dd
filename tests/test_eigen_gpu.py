"""Parity of the HIP Lanczos (wg_eigen) with the dense spectral oracle:
eigenvalues against numpy's eigvalsh of the operator matrix (the reference's
spectrum oracle), the path graph's closed form (spectral/lanczos/tests.rs:
75-101), explicit residuals, orthonormal eigenvectors of fixed sign, zero at
inactive nodes, and the exact pairs the components give (deflation.rs)."""
import numpy as np
import pytest

import pg_warren_amd as W
from oracle import spectral as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def upload(ctx, csr):
    return W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)


def random_csr(oracle, rng, n, m, weighted=True, symmetric=False, isolated=0):
    src = rng.integers(0, n - isolated, m)
    dst = rng.integers(0, n - isolated, m)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.uniform(0.5, 2.0, len(src)) if weighted else None
    if symmetric:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
        w = None if w is None else np.concatenate([w, w])
    return oracle.csr_from_edges(n, src, dst, w)


def check_vectors(vec, vals, csr, matrix, weighted=True, symmetrize=True, tol=1e-8):
    m, active, scale = S.operator(csr, matrix, weighted, symmetrize)
    x = vec.numpy().astype(np.float64)
    inactive = np.setdiff1d(np.arange(csr.n), active)
    assert np.all(x[inactive] == 0.0)
    xa = x[active]
    if matrix == "random_walk":
        xa = xa / scale[:, None]  # back to the symmetric matrix's eigenvectors
        xa = xa / np.linalg.norm(xa, axis=0)
    for l, lam in enumerate(vals):
        v = xa[:, l]
        r = np.linalg.norm(m @ v - lam * v)
        assert r <= tol * max(1.0, abs(lam)), (l, lam, r)
        top = np.argmax(np.abs(x[:, l]))
        assert x[top, l] > 0
    # Orthonormality: f32 eigenvectors carry their output rounding (each
    # component within 2^-24 relative, so a Gram entry within 2^-23 of the
    # exact vectors'), hence 1e-6 for them; the reference's own bound is
    # sqrt(accuracy) (spectral/lanczos/tests.rs:200-203), far looser.
    g = xa.T @ xa
    np.testing.assert_allclose(g, np.eye(len(vals)), atol=1e-7 if vec.dtype == "f64" else 1e-6)


def test_path_laplacian_closed_form(ctx, oracle):
    n = 60
    src = list(range(n - 1)) + list(range(1, n))
    dst = list(range(1, n)) + list(range(n - 1))
    csr = oracle.csr_from_edges(n, src, dst)
    g = upload(ctx, csr)
    vals, res, vec, info = W.eigen(ctx, g, 4, matrix="laplacian", which="smallest", precision="f64",
                                   tol=1e-8)
    assert info["converged"] and info["components"] == 1 and info["exact"] == 1
    x = vec.numpy()
    for j in range(4):
        want = 2 - 2 * np.cos(np.pi * j / n)
        assert abs(vals[j] - want) < 1e-9
        assert res[j] < 1e-7
        exact = np.cos(np.pi * j * (np.arange(n) + 0.5) / n)
        exact /= np.linalg.norm(exact)
        assert abs(abs(x[:, j] @ exact) - 1) < 1e-8


@pytest.mark.parametrize("matrix,which", [("adjacency", "largest"), ("adjacency", "smallest"),
                                          ("laplacian", "largest"), ("laplacian", "smallest"),
                                          ("normalized_laplacian", "smallest"),
                                          ("normalized_laplacian", "largest"),
                                          ("random_walk", "largest")])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_matches_dense_spectrum(ctx, oracle, matrix, which, precision):
    rng = np.random.default_rng(hash((matrix, which)) % 1000)
    csr = random_csr(oracle, rng, 300, 1500, isolated=5)
    g = upload(ctx, csr)
    k = 6
    tol = 1e-10 if precision == "f64" else 1e-6
    vals, res, vec, info = W.eigen(ctx, g, k, matrix=matrix, which=which, precision=precision, tol=tol)
    want = S.spectrum(csr, matrix, k, which)
    scale = np.abs(np.linalg.eigvalsh(S.operator(csr, matrix)[0])).max()
    atol = (1e-9 if precision == "f64" else 2e-5) * scale
    np.testing.assert_allclose(vals, want, rtol=0, atol=atol)
    assert info["converged"]
    assert np.all(res <= (1e-7 if precision == "f64" else 1e-4) * scale)
    check_vectors(vec, vals, csr, matrix, tol=1e-7 if precision == "f64" else 1e-4)


def test_thick_restarts(ctx, oracle):
    rng = np.random.default_rng(17)
    csr = random_csr(oracle, rng, 2000, 12000, symmetric=True)
    g = upload(ctx, csr)
    vals, res, vec, info = W.eigen(ctx, g, 8, matrix="normalized_laplacian", which="smallest",
                                   precision="f64", tol=1e-10, basis_size=24)
    assert info["restarts"] > 0 and info["converged"]
    np.testing.assert_allclose(vals, S.spectrum(csr, "normalized_laplacian", 8, "smallest"), atol=1e-9)
    check_vectors(vec, vals, csr, "normalized_laplacian")


def test_components_are_deflated(ctx, oracle):
    # Four components (one bipartite 4-cycle, a triangle, an edge, a random
    # block) and two isolated nodes: the normalised Laplacian's 0 four times
    # from the components, Lanczos for the rest.
    rng = np.random.default_rng(3)
    edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 4), (7, 8)]
    blk = rng.integers(9, 60, (200, 2))
    edges += [(a, b) for a, b in blk if a != b]
    src = [a for a, b in edges] + [b for a, b in edges]
    dst = [b for a, b in edges] + [a for a, b in edges]
    csr = oracle.csr_from_edges(62, src, dst)
    g = upload(ctx, csr)
    vals, res, vec, info = W.eigen(ctx, g, 6, matrix="normalized_laplacian", which="smallest",
                                   precision="f64", tol=1e-10)
    assert info["components"] == 4 and info["exact"] == 4
    np.testing.assert_allclose(vals, S.spectrum(csr, "normalized_laplacian", 6, "smallest"), atol=1e-9)
    assert np.all(vals[:4] == 0.0)
    check_vectors(vec, vals, csr, "normalized_laplacian")
    vals, res, vec, info = W.eigen(ctx, g, 3, matrix="random_walk", which="largest", precision="f64",
                                   tol=1e-10)
    np.testing.assert_allclose(vals, S.spectrum(csr, "random_walk", 3, "largest"), atol=1e-9)


def test_unsymmetrised_and_errors(ctx, oracle):
    rng = np.random.default_rng(8)
    sym = random_csr(oracle, rng, 100, 400, symmetric=True)
    g = upload(ctx, sym)
    vals, _, _, _ = W.eigen(ctx, g, 3, symmetrize=False, precision="f64", tol=1e-10, vectors=False)
    np.testing.assert_allclose(vals, S.spectrum(sym, "adjacency", 3, "largest", symmetrize=False),
                               atol=1e-9)
    asym = random_csr(oracle, rng, 100, 400)
    ga = upload(ctx, asym)
    with pytest.raises(W.WarrenError) as e:
        W.eigen(ctx, ga, 3, symmetrize=False)
    assert e.value.code == 17
    with pytest.raises(W.WarrenError) as e:
        W.eigen(ctx, ga, 0)
    assert e.value.code == 19
    with pytest.raises(W.WarrenError) as e:
        W.eigen(ctx, ga, 101)
    assert e.value.code == 19
    neg = oracle.csr_from_edges(3, [0, 1, 1, 2], [1, 0, 2, 1], [1.0, 1.0, -3.0, -3.0])
    with pytest.raises(W.WarrenError) as e:
        W.eigen(ctx, upload(ctx, neg), 1, matrix="normalized_laplacian")
    assert e.value.code == 18


def test_rmat_normalized_laplacian(ctx, oracle):
    off, tgt = W.rmat(12, 16, seed=5, symmetrize=True)
    csr = oracle.Csr(len(off) - 1, off, tgt)
    g = upload(ctx, csr)
    vals, res, vec, info = W.eigen(ctx, g, 12, matrix="normalized_laplacian", which="smallest",
                                   precision="f64", tol=1e-10, basis_size=64)
    want = S.spectrum(csr, "normalized_laplacian", 12, "smallest")
    np.testing.assert_allclose(vals, want, atol=1e-9)
    assert np.all(res < 1e-8)


def largest_component(oracle, off, tgt):
    """The largest connected component of a symmetric CSR, relabelled 0.. in
    id order (its rows stay ascending)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    n = len(off) - 1
    a = sp.csr_matrix((np.ones(len(tgt)), tgt.astype(np.int64), off.astype(np.int64)), shape=(n, n))
    _, lab = connected_components(a, directed=False)
    keep = np.flatnonzero(lab == np.bincount(lab).argmax())
    sub = a[keep][:, keep].tocsr()
    sub.sort_indices()
    return oracle.Csr(len(keep), sub.indptr.astype(np.uint64), sub.indices.astype(np.uint32))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_config_shape_32_smallest_with_a_basis_of_128(ctx, oracle, precision):
    """The eigensolver workload's shape at scale 12: the 32 smallest
    eigenpairs of the normalised Laplacian of symmetrised R-MAT's largest
    component with a Krylov basis of 128 (one exact pair from the deflation,
    31 from Lanczos, thick restarts), f64 and f32 basis storage, against the
    dense operator's eigvalsh; eigenvectors checked as the reference checks
    them (residual, orthonormality, sign)."""
    off, tgt = W.rmat(12, 16, seed=5, symmetrize=True)
    csr = largest_component(oracle, off, tgt)
    g = upload(ctx, csr)
    tol = 1e-10 if precision == "f64" else 1e-5
    vals, res, vec, info = W.eigen(ctx, g, 32, matrix="normalized_laplacian", which="smallest",
                                   precision=precision, tol=tol, basis_size=128)
    assert info["basis_size"] == 128 and info["converged"]
    assert info["exact"] == 1 and info["restarts"] > 0
    want = S.spectrum(csr, "normalized_laplacian", 32, "smallest")
    scale = 2.0  # the normalised Laplacian's spectrum lies in [0, 2]
    atol = (1e-9 if precision == "f64" else 2e-5) * scale
    np.testing.assert_allclose(vals, want, atol=atol)
    if precision == "f64":
        assert np.all(res <= 1e-8 * np.maximum(1.0, np.abs(vals)))
        check_vectors(vec, vals, csr, "normalized_laplacian", weighted=False)
    g.close()


def _sym(rng, m):
    a = rng.standard_normal((m, m))
    return (a + a.T) / 2


@pytest.mark.parametrize("m", [1, 2, 3, 17, 64, 128, 256])
def test_projected_solver_matches_eigh(ctx, m):
    """The device solve of each Rayleigh-Ritz step (spectral/ritz.rs:34-52)
    against LAPACK's eigh: values within 1e-12 ||A||, orthonormal vectors,
    and A V = V diag(values)."""
    rng = np.random.default_rng(m)
    a = _sym(rng, m)
    vals, vecs = W.sym_eigen(ctx, a)
    scale = max(np.abs(np.linalg.eigvalsh(a)).max(), 1.0)
    assert np.allclose(np.sort(vals), np.linalg.eigvalsh(a), rtol=0, atol=1e-12 * scale)
    assert np.abs(vecs.T @ vecs - np.eye(m)).max() <= 1e-12 * m
    assert np.abs(a @ vecs - vecs * vals).max() <= 1e-11 * scale


def test_projected_solver_special_shapes(ctx):
    """Shapes a thick restart produces, and closed forms."""
    # Diagonal: nothing to reduce; the values are the diagonal itself.
    d = np.array([3.0, -1.0, 2.5, 0.0, 7.0])
    vals, vecs = W.sym_eigen(ctx, np.diag(d))
    assert np.array_equal(vals, d) and np.array_equal(np.abs(vecs), np.eye(5))
    # Arrowhead (kept Ritz values on the diagonal, couplings in the last
    # row and column) followed by a tridiagonal tail.
    rng = np.random.default_rng(7)
    m, kept = 40, 24
    a = np.zeros((m, m))
    a[np.arange(kept), np.arange(kept)] = np.sort(rng.random(kept))
    a[kept, :kept] = a[:kept, kept] = rng.random(kept) * 1e-3
    for i in range(kept, m):
        a[i, i] = rng.random()
        if i + 1 < m:
            a[i, i + 1] = a[i + 1, i] = rng.random()
    vals, vecs = W.sym_eigen(ctx, a)
    assert np.allclose(np.sort(vals), np.linalg.eigvalsh(a), rtol=0, atol=1e-13)
    assert np.abs(vecs.T @ vecs - np.eye(m)).max() <= 1e-12 * m
    # The path Laplacian: 2 - 2 cos(pi j / m), j = 0..m-1.
    m = 50
    lap = 2 * np.eye(m) - np.eye(m, k=1) - np.eye(m, k=-1)
    lap[0, 0] = lap[-1, -1] = 1.0
    vals, _ = W.sym_eigen(ctx, lap)
    want = 2 - 2 * np.cos(np.pi * np.arange(m) / m)
    assert np.allclose(np.sort(vals), want, rtol=0, atol=1e-13)
    # A repeated eigenvalue: I + u u^T has 1 (m - 1 times) and 1 + |u|^2.
    u = rng.random(12)
    vals, vecs = W.sym_eigen(ctx, np.eye(12) + np.outer(u, u))
    assert np.allclose(np.sort(vals), [1.0] * 11 + [1 + u @ u], rtol=0, atol=1e-13)
    assert np.abs(vecs.T @ vecs - np.eye(12)).max() <= 1e-13 * 12


def test_projected_solver_errors(ctx):
    a = np.eye(4)
    a[1, 2] = a[2, 1] = np.nan
    with pytest.raises(W.WarrenError) as e:
        W.sym_eigen(ctx, a)
    assert e.value.name == "not_finite"
    with pytest.raises(W.WarrenError) as e:
        W.sym_eigen(ctx, np.eye(257))
    assert e.value.name == "invalid_argument"


class _Recorded:
    """Keeps every report (progress.rs recorded::Recorded)."""

    def __init__(self):
        self.reports = []

    def phase(self, name, total):
        self.reports.append(("phase", name, total))

    def done(self, done):
        self.reports.append(("done", done))

    def detail(self, name, value):
        self.reports.append(("detail", name, value))

    def phases(self):
        return [(r[1], r[2]) for r in self.reports if r[0] == "phase"]

    def dones(self):
        return [r[1] for r in self.reports if r[0] == "done"]

    def last_detail(self, name):
        vals = [r[2] for r in self.reports if r[0] == "detail" and r[1] == name]
        return vals[-1] if vals else None


@pytest.mark.parametrize("k,max_iter", [(3, None), (4, 30), (2, 5)])
def test_lanczos_reports_its_steps_and_reporting_changes_nothing(ctx, oracle, k, max_iter):
    """spectral/lanczos/tests.rs:766-803: one phase "lanczos" of the step
    limit, steps reported in order up to the call's count, converged pairs
    at most k (k when converged), restarts counted; results unchanged."""
    rng = np.random.default_rng(40 + k)
    n = 90
    # A ring (connected) plus random chords, symmetrised.
    src = np.concatenate([np.arange(n), rng.integers(0, n, 200)])
    dst = np.concatenate([(np.arange(n) + 1) % n, rng.integers(0, n, 200)])
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.uniform(0.5, 2.0, len(src))
    g = upload(ctx, oracle.csr_from_edges(n, np.concatenate([src, dst]), np.concatenate([dst, src]),
                                          np.concatenate([w, w])))
    kw = dict(matrix="normalized_laplacian", which="smallest", precision="f64", seed=5,
              max_iter=max_iter, tol=1e-9)
    plain = W.eigen(ctx, g, k, **kw)
    rec = _Recorded()
    ctx.set_progress(rec)
    try:
        got = W.eigen(ctx, g, k, **kw)
    finally:
        ctx.set_progress(None)
    assert np.array_equal(plain[0], got[0]) and np.array_equal(plain[1], got[1])
    assert np.array_equal(plain[2].numpy(), got[2].numpy())
    info = got[3]
    size = max(2 * k + 1, 20)
    total = max(max_iter if max_iter else 1000 * size, k)
    assert rec.phases() == [("lanczos", total)]
    dones = rec.dones()
    assert dones == sorted(dones) and dones[-1] == info["iterations"] <= total
    conv = rec.last_detail("converged")
    assert conv <= k and (not info["converged"] or conv == k)
    assert rec.last_detail("restarts") == info["restarts"]
