"""Parity of the HIP PageRank with the oracle (compute/pagerank/mod.rs).

The bar: within 1e-9 L1 of the oracle's FP64 ranks at the same iteration
count (the measured error is far smaller: the only difference is the order
of f64 additions); on the reference's golden graphs the ranks agree to the
last bit.
"""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu

L1_TOL = 1e-9


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def small_edges():
    return [(0, 0, 1, 1.0), (0, 1, 2, 1.0), (1, 2, 3, 1.0), (1, 0, 3, 5.0), (0, 3, 0, 1.0),
            (1, 4, 4, 2.0), (1, 0, 1, 0.5)]


def merged(oracle, n, edges):
    edges = sorted(edges, key=lambda e: e[0])
    return oracle.csr_from_edges(n, [e[1] for e in edges], [e[2] for e in edges],
                                 [e[3] for e in edges])


def gpu_rank(ctx, csr, **kw):
    g = W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)
    v, it = W.pagerank(ctx, g, **kw)
    return v.numpy(), it


def test_golden_ranks(ctx, oracle):
    """compute/pagerank/tests.rs:78-123, to the last bit."""
    g = merged(oracle, 6, small_edges())
    vals, it = gpu_rank(ctx, g)
    assert it == 14
    assert list(vals) == [0.28666944658969795, 0.08535717971974721, 0.10168006797333676,
                          0.3029923348432875, 0.19417475728151729, 0.029126213592413387]
    edges = [e for e in small_edges() if e[1] != 1 and e[2] != 1]
    g = merged(oracle, 6, edges)
    tomb = np.array([0, 1, 0, 0, 0, 0], np.uint8)
    vals, it = gpu_rank(ctx, g, tombstones=tomb)
    assert it == 69
    assert [vals[i] for i in [0, 2, 3, 4, 5]] == [
        0.33506897700306837, 0.03614457831325302, 0.3516780109487391, 0.24096385542168675,
        0.03614457831325302]
    assert vals[1] == 0.0
    vals, it = gpu_rank(ctx, g, tombstones=tomb, seed=np.arange(1, 7, dtype=np.float64))
    assert it == 71
    assert [vals[i] for i in [0, 2, 3, 4, 5]] == [
        0.2553952337234808, 0.03237410071942446, 0.28777128327185025, 0.3597111808463956,
        0.06474820143884892]


def test_special_cases_and_errors(ctx, oracle):
    """compute/pagerank/tests.rs:139-169."""
    g = merged(oracle, 6, small_edges())
    vals, it = gpu_rank(ctx, g, max_iter=0)
    assert it == 0 and np.all(vals == 1.0 / 6)
    assert gpu_rank(ctx, g, tol=float("nan"))[1] == 1
    assert gpu_rank(ctx, g, tol=None)[1] == 100
    for kw, name in [(dict(tombstones=np.ones(6, np.uint8)), "division_by_zero"),
                     (dict(seed=np.zeros(6)), "division_by_zero"),
                     (dict(seed=np.ones(6, np.uint8).astype(bool)), "dtype_mismatch"),
                     (dict(tol=1e308), "float_overflow")]:
        with pytest.raises(W.WarrenError) as e:
            gpu_rank(ctx, g, **kw)
        assert e.value.name == name, kw
    # A ring ranks every node alike (tests.rs:125-137).
    ring = oracle.csr_from_edges(4, [0, 1, 2, 3], [1, 2, 3, 0])
    vals, _ = gpu_rank(ctx, ring, tol=1e-12)
    assert np.all(np.abs(vals - 0.25) < 1e-12)


@pytest.mark.parametrize("weighted", [True, False])
@pytest.mark.parametrize("scale", [17, 19])
def test_rmat_matches_the_oracle(ctx, oracle, weighted, scale):
    """Scale 17 runs the blocked gather; scale 19 (2^22 entries or more)
    the tiled one, with its hot and warm message copies."""
    off, tgt = W.rmat(scale, 16, seed=21)
    n = len(off) - 1
    assert scale == 17 or len(tgt) >= 1 << 22
    w = W.edge_weights(off, tgt, "lognormal", seed=3, p0=0.0, p1=1.0) if weighted else None
    ref = oracle.Csr(n, off, tgt, w)
    want, wit = oracle.pagerank(ref, tol=1e-10, max_iter=100)
    got, it = gpu_rank(ctx, ref, tol=1e-10, max_iter=100)
    assert it == wit
    err = np.abs(got - want).sum()
    assert err < L1_TOL, err
    seed = np.random.default_rng(1).random(n)
    seed[seed < 0.9] = 0.0
    tomb = (np.random.default_rng(2).random(n) < 0.01).astype(np.uint8)
    want, wit = oracle.pagerank(ref, tol=1e-10, max_iter=60, seed=seed, tombstones=tomb)
    got, it = gpu_rank(ctx, ref, tol=1e-10, max_iter=60, seed=seed, tombstones=tomb)
    assert it == wit
    assert np.abs(got - want).sum() < L1_TOL
    assert np.all(got[tomb == 1] == 0.0)
