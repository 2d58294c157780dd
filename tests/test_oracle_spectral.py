"""The dense spectral oracle against closed forms: the path graph's Laplacian
spectrum 2 - 2cos(pi j / n) (spectral/lanczos/tests.rs:75-101), the
normalised Laplacian's eigenvalue 0 once per component, the random walk's
-1 once per bipartite component, and the two-colouring of an even cycle."""
import numpy as np

import oracle as O
from oracle import spectral as S


def path(n):
    src = list(range(n - 1)) + list(range(1, n))
    dst = list(range(1, n)) + list(range(n - 1))
    return O.csr_from_edges(n, src, dst)


def test_path_laplacian_closed_form(oracle):
    n = 40
    ev = S.spectrum(path(n), "laplacian", n, "smallest")
    want = 2 - 2 * np.cos(np.pi * np.arange(n) / n)
    np.testing.assert_allclose(ev, want, atol=1e-12)


def test_components_and_their_eigenvalues(oracle):
    # A triangle (odd cycle), a 4-cycle (bipartite), an edge, and 2 isolated nodes.
    edges = [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 6), (6, 3), (7, 8)]
    src = [a for a, b in edges] + [b for a, b in edges]
    dst = [b for a, b in edges] + [a for a, b in edges]
    csr = O.csr_from_edges(11, src, dst)
    m, active, _ = S.operator(csr, "normalized_laplacian")
    assert list(active) == list(range(9))
    ev = np.linalg.eigvalsh(m)
    assert np.sum(np.abs(ev) < 1e-12) == 3
    rw = np.linalg.eigvalsh(S.operator(csr, "random_walk")[0])
    assert np.sum(np.abs(rw + 1) < 1e-12) == 2
    comp, colour, bip = S.components(csr, active)
    assert list(comp[:9]) == [0, 0, 0, 1, 1, 1, 1, 2, 2]
    assert list(bip) == [False, True, True]
    assert list(colour[3:7]) == [False, True, False, True]
    assert list(colour[7:9]) == [False, True]
