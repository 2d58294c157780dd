"""Weight casts on the device against the reference's own known answers
(compute/cast.rs:99-135 and compute/semiring.rs:519-533 tests) and the
oracle: f64 -> i64 rounds half to even, then range-checks; f64 -> f32 fails
on overflow and underflow (a non-zero weight that narrows to zero); f64 ->
bool is "non-zero" (NaN true). Each weight is cast where the product reads
it (compute/product/kernel.rs:40-47), in every kernel variant."""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu

VARIANTS = [0, 1, 2, 3, 8, 16, 18, 32]
I64MIN = -(2**63)


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def star(oracle, weights):
    """Row i holds one entry (i -> 0) of weight weights[i]."""
    n = len(weights)
    return oracle.csr_from_edges(n, np.arange(n), np.zeros(n, np.int64), np.asarray(weights, np.float64))


def product(ctx, csr, dtype, semiring="plus_times"):
    g = W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)
    x = W.Vector(ctx, np.ones(csr.n), dtype=dtype)
    y = W.Vector(ctx, dtype=dtype, id_space=csr.n)
    try:
        W.spmv(ctx, g, W.Product(semiring), x, y)
        return y.numpy()
    finally:
        for h in (x, y, g):
            h.close()


# (weight, its i64 image): cast.rs:110-118, semiring.rs:526-528.
I64_CASES = [(2.5, 2), (3.5, 4), (-2.5, -2), (0.5, 0), (-0.5, 0), (1.5, 2), (-1.5, -2),
             (2.4999999999999996, 2), (4503599627370495.5, 4503599627370496),
             (-9223372036854775808.0, I64MIN), (1e18, 10**18), (7.0, 7)]


@pytest.mark.parametrize("variant", VARIANTS)
def test_f64_to_i64_rounds_half_to_even(ctx, oracle, variant):
    W.set_spmv_variant(ctx, variant)
    try:
        csr = star(oracle, [w for w, _ in I64_CASES])
        got = product(ctx, csr, "i64")
        assert list(got) == [v for _, v in I64_CASES]
        x = np.ones(csr.n, np.int64)
        assert got.tobytes() == oracle.spmv(csr, x, "plus_times", "i64").tobytes()
    finally:
        W.set_spmv_variant(ctx, -1)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("weight", [9.3e18, 9223372036854775808.0, float("nan"), -9.3e18,
                                    float("inf")])
def test_f64_to_i64_out_of_range_is_an_error(ctx, oracle, variant, weight):
    W.set_spmv_variant(ctx, variant)
    try:
        csr = star(oracle, [1.0, weight, 2.0])
        with pytest.raises(W.WarrenError) as e:
            product(ctx, csr, "i64")
        assert e.value.name == "bigint_out_of_range"
        with pytest.raises(oracle.OracleError) as o:
            oracle.spmv(csr, np.ones(3, np.int64), "plus_times", "i64")
        assert o.value.name == "bigint_out_of_range"
    finally:
        W.set_spmv_variant(ctx, -1)


# (weight, its f32 image): cast.rs:101-108, semiring.rs:522-525.
F32_CASES = [(0.1, np.float32(0.1)), (1.5e-45, np.float32(1e-45)), (1e-40, np.float32(1e-40)),
             (float("inf"), np.float32(np.inf)), (-float("inf"), np.float32(-np.inf)),
             (3.4028234663852886e38, np.float32(3.4028234663852886e38)), (-0.0, np.float32(-0.0)),
             (0.0, np.float32(0.0))]


@pytest.mark.parametrize("variant", VARIANTS)
def test_f64_to_f32_narrows_and_keeps_subnormals(ctx, oracle, variant):
    W.set_spmv_variant(ctx, variant)
    try:
        csr = star(oracle, [w for w, _ in F32_CASES] + [float("nan")])
        got = product(ctx, csr, "f32", "plus_times")
        want = np.array([v for _, v in F32_CASES] + [np.float32(np.nan)], np.float32)
        # times by x = 1 keeps every image (and -0.0 * 1 = -0.0).
        assert got[:-1].tobytes() == want[:-1].tobytes()
        assert np.isnan(got[-1])
        ref = oracle.spmv(csr, np.ones(csr.n, np.float32), "plus_times", "f32")
        assert np.array_equal(got, ref, equal_nan=True)
    finally:
        W.set_spmv_variant(ctx, -1)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("weight,name", [(1e-50, "float_underflow"), (-1e-300, "float_underflow"),
                                         (1e39, "float_overflow"), (-1e300, "float_overflow")])
def test_f64_to_f32_overflow_and_underflow_are_errors(ctx, oracle, variant, weight, name):
    W.set_spmv_variant(ctx, variant)
    try:
        csr = star(oracle, [1.0, 2.0, weight, 3.0])
        with pytest.raises(W.WarrenError) as e:
            product(ctx, csr, "f32")
        assert e.value.name == name
        with pytest.raises(oracle.OracleError) as o:
            oracle.spmv(csr, np.ones(4, np.float32), "plus_times", "f32")
        assert o.value.name == name
        # A product that never reads the weight never casts it.
        assert list(product(ctx, csr, "f32", "plus_second")) == [1.0, 1.0, 1.0, 1.0]
    finally:
        W.set_spmv_variant(ctx, -1)


@pytest.mark.parametrize("variant", VARIANTS)
def test_f64_to_bool_is_non_zero(ctx, oracle, variant):
    W.set_spmv_variant(ctx, variant)
    try:
        weights = [0.0, -0.0, float("nan"), 0.5, -3.0, 1e-300, float("inf")]
        csr = star(oracle, weights)
        got = product(ctx, csr, "bool", "lor_land")
        assert list(got) == [0, 0, 1, 1, 1, 1, 1]
        ref = oracle.spmv(csr, np.ones(csr.n, np.uint8), "lor_land", "bool")
        assert got.tobytes() == ref.tobytes()
    finally:
        W.set_spmv_variant(ctx, -1)


def test_casts_on_a_graph_of_many_rows(ctx, oracle):
    """Fractional weights over a graph large enough for the autotuned and
    tiled kernels: every row's i64 sum of half-even images, bit for bit."""
    rng = np.random.default_rng(11)
    n, m = 20000, 300000
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.integers(-40, 40, m) / 2.0  # halves: every tie case
    csr = oracle.csr_from_edges(n, src, dst, w)
    x = rng.integers(-3, 4, n).astype(np.int64)
    want = oracle.spmv(csr, x, "plus_times", "i64")
    g = W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights)
    for variant in (-1, 0, 32):
        W.set_spmv_variant(ctx, variant)
        y = W.Vector(ctx, dtype="i64", id_space=n)
        W.spmv(ctx, g, W.Product("plus_times"), W.Vector(ctx, x, dtype="i64"), y)
        assert y.numpy().tobytes() == want.tobytes(), variant
    W.set_spmv_variant(ctx, -1)
    g.close()
