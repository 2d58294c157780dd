"""Parity of the HIP semiring products with the oracle.

Bit-exact for every monoid except float plus, whose rounding depends on the
grouping of the sum: there the check is exactness on integer-valued floats
(every partial sum representable) and, for random floats, a rounding bound
|gpu - oracle| <= 2 * len * eps * sum|terms| per row.
"""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu

I64MAX = 2**63 - 1
I64MIN = -(2**63)
U32INF = 2**32 - 1
EPS = {"f32": np.finfo(np.float32).eps, "f64": np.finfo(np.float64).eps}


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    _graphs.clear()
    c.close()


_graphs = {}


def gpu_graph(ctx, csr):
    key = id(csr)
    if key not in _graphs:
        _graphs[key] = (csr, W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights))
    return _graphs[key][1]


def run_gpu(ctx, csr, x, semiring, dtype, *, transpose=False, weighted=None, mask=None,
            mask_complement=False, replace=False, accumulate=None, y0=None, spmm=False):
    g = gpu_graph(ctx, csr)
    xv = W.Vector(ctx, x, dtype=dtype)
    yv = W.Vector(ctx, np.zeros_like(x) if y0 is None else y0, dtype=dtype)
    mv = None if mask is None else W.Vector(ctx, mask)
    p = W.Product(semiring, transpose=transpose, weighted=weighted, mask=mv,
                  mask_complement=mask_complement, replace=replace, accumulate=accumulate)
    (W.spmm if spmm else W.spmv)(ctx, g, p, xv, yv)
    return yv.numpy()


def run_oracle(oracle, csr, x, semiring, dtype, *, transpose=False, weighted=None, mask=None,
               mask_complement=False, replace=False, accumulate=None, y0=None, width=1):
    w = csr.weights is not None if weighted is None else weighted
    y = np.zeros_like(x) if y0 is None else y0.copy()
    return oracle.spmv(csr, x.reshape(-1), semiring, dtype, rev=csr.transpose() if transpose else None,
                       transpose=transpose, weighted=w, mask=mask, mask_complement=mask_complement,
                       replace=replace, accumulate=accumulate, y=y.reshape(-1), width=width)


def example(oracle):
    src, dst, w = [0, 0, 1, 2], [1, 2, 2, 0], [12.0, 3.0, 4.0, 5.0]
    return oracle.csr_from_edges(4, src, dst, w)


X4 = np.array([1.0, 2.0, 3.0, 4.0])


def test_known_answers(ctx, oracle):
    """compute/product/tests.rs:51-106 through the device."""
    g = example(oracle)
    assert list(run_gpu(ctx, g, X4, "plus_times", "f64")) == [33.0, 12.0, 5.0, 0.0]
    assert list(run_gpu(ctx, g, X4, "plus_times", "f64", transpose=True)) == [15.0, 12.0, 11.0, 0.0]
    assert list(run_gpu(ctx, g, X4, "plus_times", "f64", weighted=False)) == [5.0, 3.0, 1.0, 0.0]
    assert list(run_gpu(ctx, g, X4, "min_plus", "f64")) == [6.0, 7.0, 6.0, float("inf")]
    mask = np.array([1, 0, -7, 0], np.int64)
    y100 = np.full(4, 100.0)

    def w(c, r, a):
        return list(run_gpu(ctx, g, X4, "plus_times", "f64", mask=mask, mask_complement=c,
                            replace=r, accumulate=a, y0=y100))
    assert w(False, False, None) == [33.0, 100.0, 5.0, 100.0]
    assert w(False, True, None) == [33.0, 0.0, 5.0, 0.0]
    assert w(False, False, "plus") == [133.0, 100.0, 105.0, 100.0]
    assert w(True, False, None) == [100.0, 12.0, 100.0, 0.0]
    assert w(True, True, "plus") == [0.0, 112.0, 0.0, 100.0]
    x2 = np.array([[1.0, 10.0], [2.0, 20.0], [3.0, 30.0], [4.0, 40.0]])
    y = run_gpu(ctx, g, x2, "plus_times", "f64", transpose=True, spmm=True)
    assert list(y.reshape(-1)) == [15.0, 150.0, 12.0, 120.0, 11.0, 110.0, 0.0, 0.0]


def test_i64_min_plus_supersteps(ctx, oracle):
    """compute/product/tests.rs:160-185."""
    g = oracle.csr_from_edges(3, [0, 2], [1, 1], [1.0, 1.0])
    d = np.array([0, I64MAX, I64MAX], np.int64)
    for _ in range(2):
        d = run_gpu(ctx, g, d, "min_plus", "i64", transpose=True, accumulate="min", y0=d)
        assert list(d) == [0, 1, I64MAX]
    y = run_gpu(ctx, g, np.array([5, I64MIN, I64MIN], np.int64), "max_plus", "i64",
                transpose=True, accumulate="max", y0=np.full(3, I64MIN, np.int64))
    assert list(y) == [I64MIN, 6, I64MIN]


def test_errors(ctx, oracle):
    g = example(oracle)
    with pytest.raises(W.WarrenError) as e:
        run_gpu(ctx, g, np.full(4, I64MAX, np.int64), "plus_times", "i64")
    assert e.value.name == "bigint_out_of_range"
    with pytest.raises(W.WarrenError) as e:
        run_gpu(ctx, g, np.full(4, 2**32 - 2, np.uint32), "plus_times", "u32")
    assert e.value.name == "bigint_out_of_range"
    big = oracle.csr_from_edges(2, [0], [1], [1e300])
    with pytest.raises(W.WarrenError) as e:
        run_gpu(ctx, big, np.ones(2, np.float32), "plus_times", "f32")
    assert e.value.name == "float_overflow"
    # (x) second never reads the weight, so the bad weight is not cast.
    assert list(run_gpu(ctx, big, np.ones(2, np.float32), "plus_second", "f32")) == [1.0, 0.0]
    gg = gpu_graph(ctx, g)
    x = W.Vector(ctx, np.ones(4))
    for y, name in [(W.Vector(ctx, np.zeros(4, np.int64)), "dtype_mismatch"),
                    (W.Vector(ctx, np.zeros(3)), "id_space_mismatch"),
                    (W.Vector(ctx, np.zeros((4, 2))), "width_mismatch")]:
        with pytest.raises(W.WarrenError) as e:
            W.spmv(ctx, gg, W.Product(), x, y)
        assert e.value.name == name
    with pytest.raises(W.WarrenError) as e:
        W.spmv(ctx, gg, W.Product("plus_min"), x, W.Vector(ctx, np.zeros(4)))
    assert e.value.name == "invalid_argument"
    with pytest.raises(W.WarrenError) as e:
        run_gpu(ctx, g, np.ones(4, np.uint32), "max_plus", "u32")
    assert e.value.name == "unsupported"
    ctx.cancel()
    with pytest.raises(W.WarrenError) as e:
        W.spmv(ctx, gg, W.Product(), x, W.Vector(ctx, np.zeros(4)))
    assert e.value.name == "cancelled"
    ctx.cancel_reset()


def hub_graph(oracle, seed, n=5000, m=40000):
    """Random edges plus a hub whose row and column exceed one block (2048
    entries), so long rows split into pieces both ways."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([rng.integers(0, n, m), np.zeros(n - 1, np.int64),
                          np.arange(1, n)])
    dst = np.concatenate([rng.integers(0, n, m), np.arange(1, n), np.full(n - 1, 7)])
    w = rng.integers(1, 4, len(src)).astype(np.float64)
    return oracle.csr_from_edges(n, src, dst, w)


def values(dtype, n, rng, add):
    if dtype in ("f32", "f64"):
        v = rng.integers(0, 5, n).astype(np.float64)
        if add in ("min", "max"):
            v[rng.random(n) < 0.05] = np.nan
            v[rng.random(n) < 0.05] = -0.0
        return v.astype(W.NP_DTYPES[dtype])
    if dtype == "i64":
        v = rng.integers(-3, 6, n).astype(np.int64)
        if add == "min":
            v[rng.random(n) < 0.1] = I64MAX
        if add == "max":
            v[rng.random(n) < 0.1] = I64MIN
        return v
    if dtype == "u32":
        v = rng.integers(0, 6, n).astype(np.uint32)
        if add == "min":
            v[rng.random(n) < 0.1] = U32INF
        return v
    return (rng.random(n) < 0.5).astype(np.uint8)


ADD_OF = {"plus_times": "plus", "min_plus": "min", "max_plus": "max", "max_times": "max",
          "min_times": "min", "lor_land": "lor", "any_pair": "any", "plus_pair": "plus",
          "plus_first": "plus", "plus_second": "plus", "min_first": "min", "min_second": "min",
          "max_second": "max"}


@pytest.mark.parametrize("variant", list(range(16)) + [16, 18, 32])
@pytest.mark.parametrize("dtype", ["f32", "f64", "i64", "bool", "u32"])
def test_every_semiring_matches_the_oracle_bit_for_bit(ctx, oracle, dtype, variant):
    """Integer-valued inputs: every partial sum is exact, so even float plus
    must agree to the bit; NaN and -0 exercise the SQL min/max order. Every
    launch variant the autotuner may pick is checked."""
    W.set_spmv_variant(ctx, variant)
    g = hub_graph(oracle, 1)
    rng = np.random.default_rng(2)
    for sr in W.SEMIRINGS:
        add = ADD_OF[sr]
        if dtype == "u32" and add == "max":
            continue
        x = values(dtype, g.n, rng, add)
        for transpose in (False, True):
            for weighted in (None, False):
                try:
                    want = run_oracle(oracle, g, x, sr, dtype, transpose=transpose,
                                      weighted=weighted)
                except oracle.OracleError as e:
                    with pytest.raises(W.WarrenError) as ge:
                        run_gpu(ctx, g, x, sr, dtype, transpose=transpose, weighted=weighted)
                    assert ge.value.name == e.name, (sr, dtype, transpose)
                    continue
                got = run_gpu(ctx, g, x, sr, dtype, transpose=transpose, weighted=weighted)
                assert got.tobytes() == want.tobytes(), (sr, dtype, transpose, weighted)
    W.set_spmv_variant(ctx, -1)


@pytest.mark.parametrize("variant", [0, 3, 7, 8, 15, 16, 18, 32])
@pytest.mark.parametrize("dtype", ["f32", "f64", "i64", "bool"])
def test_masks_replace_accumulate_match_the_oracle(ctx, oracle, dtype, variant):
    W.set_spmv_variant(ctx, variant)
    g = hub_graph(oracle, 3)
    rng = np.random.default_rng(4)
    for sr in ["plus_times", "min_plus", "lor_land", "max_second"]:
        x = values(dtype, g.n, rng, ADD_OF[sr])
        y0 = values(dtype, g.n, rng, ADD_OF[sr])
        mask = rng.integers(0, 2, g.n).astype(np.float32)
        for complement in (False, True):
            for replace in (False, True):
                for acc in (None, "plus", "min", "second", "lor"):
                    kw = dict(mask=mask, mask_complement=complement, replace=replace,
                              accumulate=acc, y0=y0, transpose=bool(replace))
                    try:
                        want = run_oracle(oracle, g, x, sr, dtype, **kw)
                    except oracle.OracleError as e:
                        with pytest.raises(W.WarrenError) as ge:
                            run_gpu(ctx, g, x, sr, dtype, **kw)
                        assert ge.value.name == e.name
                        continue
                    got = run_gpu(ctx, g, x, sr, dtype, **kw)
                    assert got.tobytes() == want.tobytes(), (sr, dtype, complement, replace, acc)
    W.set_spmv_variant(ctx, -1)


@pytest.mark.parametrize("variant", [0, 1, 6, 7, 9, 14, 18, 32])
def test_float_plus_within_rounding_bound(ctx, oracle, variant):
    W.set_spmv_variant(ctx, variant)
    g = hub_graph(oracle, 5)
    rng = np.random.default_rng(6)
    rev = g.transpose()
    for dtype in ("f32", "f64"):
        x = (rng.random(g.n) * 2 - 1).astype(W.NP_DTYPES[dtype])
        for transpose in (False, True):
            got = run_gpu(ctx, g, x, "plus_times", dtype, transpose=transpose).astype(np.float64)
            want = run_oracle(oracle, g, x, "plus_times", dtype, transpose=transpose).astype(np.float64)
            m = rev if transpose else g
            absx = np.abs(x.astype(np.float64))
            abssum = oracle.spmv(m, absx, "plus_times", "f64")
            deg = np.diff(m.offsets.astype(np.int64))
            bound = 2 * deg * EPS[dtype] * abssum
            assert np.all(np.abs(got - want) <= bound + 1e-300)
    W.set_spmv_variant(ctx, -1)


@pytest.mark.parametrize("width", [2, 4, 5, 16, 64, 70])
def test_spmm_matches_the_oracle(ctx, oracle, width):
    g = hub_graph(oracle, 7, n=3000, m=20000)
    rng = np.random.default_rng(width)
    for sr, dtype in [("plus_times", "f64"), ("min_plus", "i64"), ("lor_land", "bool"),
                      ("plus_times", "f32"), ("min_plus", "u32")]:
        x = values(dtype, g.n * width, rng, ADD_OF[sr]).reshape(g.n, width)
        for transpose in (False, True):
            want = run_oracle(oracle, g, x, sr, dtype, transpose=transpose, width=width)
            got = run_gpu(ctx, g, x, sr, dtype, transpose=transpose, spmm=True)
            assert got.tobytes() == want.tobytes(), (sr, dtype, width, transpose)


@pytest.mark.parametrize("grouped", [0, 256])
def test_every_spmm_shape_matches_the_oracle(ctx, oracle, grouped):
    """Each launch shape the SpMM autotuner may pick (lanes per row 1..64,
    one column or 16-byte column groups a lane) gives the oracle's bits."""
    g = hub_graph(oracle, 11, n=2000, m=15000)
    rng = np.random.default_rng(grouped)
    try:
        for lanes in (1, 2, 4, 8, 16, 32, 64):
            W.set_spmm_shape(ctx, lanes | grouped)
            for sr, dtype, width in [("plus_times", "f64", 8), ("min_plus", "u32", 12), ("plus_times", "f32", 16),
                                     ("min_plus", "i64", 6), ("lor_land", "bool", 5)]:
                x = values(dtype, g.n * width, rng, ADD_OF[sr]).reshape(g.n, width)
                want = run_oracle(oracle, g, x, sr, dtype, width=width)
                got = run_gpu(ctx, g, x, sr, dtype, spmm=True)
                assert got.tobytes() == want.tobytes(), (lanes, grouped, sr, dtype, width)
    finally:
        W.set_spmm_shape(ctx, -1)


def test_spmm_autotuner_keeps_a_shape_and_the_bits(ctx):
    """On a graph of 2^22 entries or more the first SpMM tunes its shape; the
    tuned product equals the product under the default shape bit for bit."""
    off, tgt = W.rmat(19, 16, seed=12)
    n = len(off) - 1
    assert len(tgt) >= 1 << 22
    wf = W.edge_weights(off, tgt, "uniform01", seed=4)
    g = W.Graph(ctx, n, off, tgt, wf)
    x = np.random.default_rng(3).random((n, 16)).astype(np.float32)
    assert W.spmm_shape(g, "plus_times", "f32", 16) is None
    got = W.spmm(ctx, g, W.Product("plus_times"), W.Vector(ctx, x),
                 W.Vector(ctx, dtype="f32", id_space=n, width=16)).numpy()
    shape = W.spmm_shape(g, "plus_times", "f32", 16)
    assert shape is not None and shape[0] in (4, 2, 1, 16, 8) and isinstance(shape[1], bool)
    W.set_spmm_shape(ctx, 4 | 256)  # the default: 4 lanes of 4 columns
    try:
        want = W.spmm(ctx, g, W.Product("plus_times"), W.Vector(ctx, x),
                      W.Vector(ctx, dtype="f32", id_space=n, width=16)).numpy()
    finally:
        W.set_spmm_shape(ctx, -1)
    assert got.tobytes() == want.tobytes()
    g.close()


def test_rmat_scale18_three_semirings(ctx, oracle):
    """The benchmark's three products on R-MAT scale 18: min_plus over u32
    and or_and over bool bit-exact, plus_times f32 within its bound."""
    off, tgt = W.rmat(18, 16, seed=11)
    wf = W.edge_weights(off, tgt, "uniform01", seed=1)
    wi = W.edge_weights(off, tgt, "int", seed=2, p0=1, p1=255)
    n = len(off) - 1
    rng = np.random.default_rng(12)
    gf = oracle.Csr(n, off, tgt, wf)
    gi = oracle.Csr(n, off, tgt, wi)
    xf = ((rng.integers(0, 2**24, n) + 1) / 2.0**24).astype(np.float32)
    got = run_gpu(ctx, gf, xf, "plus_times", "f32").astype(np.float64)
    want = run_oracle(oracle, gf, xf, "plus_times", "f32").astype(np.float64)
    exact = oracle.spmv(gf, xf.astype(np.float64), "plus_times", "f64")
    deg = np.diff(off.astype(np.int64))
    assert np.all(np.abs(got - want) <= 2 * deg * EPS["f32"] * exact + 1e-30)
    xu = rng.integers(0, 2**24, n).astype(np.uint32)
    xu[rng.random(n) < 0.1] = U32INF
    got = run_gpu(ctx, gi, xu, "min_plus", "u32")
    want = run_oracle(oracle, gi, xu, "min_plus", "u32")
    assert got.tobytes() == want.tobytes()
    # The same through the reference's i64 type: u32 is its exact image.
    xi = np.where(xu == U32INF, I64MAX, xu.astype(np.int64))
    wi64 = run_oracle(oracle, gi, xi, "min_plus", "i64")
    assert np.array_equal(np.where(wi64 == I64MAX, U32INF, wi64).astype(np.uint32), got)
    xb = (rng.random(n) < 0.5).astype(np.uint8)
    got = run_gpu(ctx, gf, xb, "lor_land", "bool")
    want = run_oracle(oracle, gf, xb, "lor_land", "bool")
    assert got.tobytes() == want.tobytes()


def test_autotuner_choice_is_a_valid_variant(ctx, oracle):
    """Above 2^22 entries the first product tunes; the result matches the
    forced default variant bit for bit on integer-valued input."""
    off, tgt = W.rmat(19, 16, seed=2)
    n = len(off) - 1
    w = W.edge_weights(off, tgt, "int", seed=5, p0=1, p1=4)
    g = W.Graph(ctx, n, off, tgt, w)
    x = np.random.default_rng(3).integers(0, 4, n).astype(np.float32)
    xv = W.Vector(ctx, x)
    tuned = W.spmv(ctx, g, W.Product(), xv, W.Vector(ctx, np.zeros_like(x))).numpy()
    choice = W.spmv_variant(g, "plus_times", "f32")
    assert choice == 32  # float plus above 2^22 entries: the tiled kernel
    W.set_spmv_variant(ctx, 0)
    plain = W.spmv(ctx, g, W.Product(), xv, W.Vector(ctx, np.zeros_like(x))).numpy()
    W.set_spmv_variant(ctx, -1)
    assert tuned.tobytes() == plain.tobytes()


def test_tuned_float_plus_has_the_default_bits(ctx, oracle):
    """Random (not integer-valued) f32 and f64 input above 2^22 entries:
    float plus takes the tiled kernel by size (no timing decides), so the
    result is the forced tiled launch's bits, within the rounding bound of
    the oracle's left fold."""
    off, tgt = W.rmat(19, 16, seed=4)
    n = len(off) - 1
    w = W.edge_weights(off, tgt, "uniform01", seed=6)
    rng = np.random.default_rng(8)
    ref = oracle.Csr(n, off, tgt, w)
    rev = ref.transpose()
    for dtype in ("f32", "f64"):
        g = W.Graph(ctx, n, off, tgt, w)
        x = (rng.random(n) * 2 - 1).astype(W.NP_DTYPES[dtype])
        xv = W.Vector(ctx, x)
        for transpose in (False, True):
            prod = W.Product("plus_times", transpose=transpose)
            tuned = W.spmv(ctx, g, prod, xv, W.Vector(ctx, dtype=dtype, id_space=n)).numpy()
            choice = W.spmv_variant(g, "plus_times", dtype, transpose=transpose)
            assert choice == 32, (dtype, transpose, choice)
            W.set_spmv_variant(ctx, 32)
            forced = W.spmv(ctx, g, prod, xv, W.Vector(ctx, dtype=dtype, id_space=n)).numpy()
            W.set_spmv_variant(ctx, -1)
            assert tuned.tobytes() == forced.tobytes(), (dtype, transpose)
            m = rev if transpose else ref
            want = run_oracle(oracle, ref, x, "plus_times", dtype, transpose=transpose).astype(np.float64)
            abssum = oracle.spmv(m, np.abs(x.astype(np.float64)), "plus_times", "f64")
            deg = np.diff(m.offsets.astype(np.int64))
            assert np.all(np.abs(tuned.astype(np.float64) - want) <= 2 * deg * EPS[dtype] * abssum + 1e-300)
        g.close()


def test_tiled_kernel_across_waves_matches_the_oracle(ctx, oracle):
    """R-MAT scale 19 (8.4M entries, about four 512-entry tiles per wave, so
    rows cross tiles and waves, and hub rows span many waves): the tiled
    kernel forced for every exact monoid matches the oracle bit for bit,
    i64 plus with its prefix-overflow rule included."""
    off, tgt = W.rmat(19, 16, seed=21)
    n = len(off) - 1
    w = W.edge_weights(off, tgt, "int", seed=22, p0=1, p1=9)
    ref = oracle.Csr(n, off, tgt, w)
    g = W.Graph(ctx, n, off, tgt, w)
    rng = np.random.default_rng(23)
    W.set_spmv_variant(ctx, 32)
    try:
        for sr, dtype in [("min_plus", "u32"), ("min_plus", "i64"), ("lor_land", "bool"),
                          ("plus_times", "i64"), ("max_times", "f64"), ("plus_times", "f32"),
                          ("any_pair", "i64"), ("min_second", "f32"), ("plus_pair", "u32")]:
            x = values(dtype, n, rng, ADD_OF[sr])
            for transpose in (False, True):
                want = run_oracle(oracle, ref, x, sr, dtype, transpose=transpose)
                got = W.spmv(ctx, g, W.Product(sr, transpose=transpose), W.Vector(ctx, x, dtype=dtype),
                             W.Vector(ctx, dtype=dtype, id_space=n)).numpy()
                assert got.tobytes() == want.tobytes(), (sr, dtype, transpose)
        # A row whose i64 prefix sum overflows raises as the reference does.
        x = np.zeros(n, np.int64)
        x[tgt[off[0]:off[1]]] = I64MAX // 2
        with pytest.raises(W.WarrenError) as e:
            W.spmv(ctx, g, W.Product("plus_times"), W.Vector(ctx, x), W.Vector(ctx, np.zeros(n, np.int64)))
        assert e.value.name == "bigint_out_of_range"
    finally:
        W.set_spmv_variant(ctx, -1)
        g.close()


def test_repeated_launches_are_bit_identical(ctx, oracle):
    g = hub_graph(oracle, 9)
    x = np.random.default_rng(1).random(g.n).astype(np.float32)
    first = run_gpu(ctx, g, x, "plus_times", "f32", transpose=True)
    for _ in range(3):
        assert run_gpu(ctx, g, x, "plus_times", "f32", transpose=True).tobytes() == first.tobytes()


@pytest.mark.parametrize("variant", [8, 16, 18, 32])
def test_hot_column_variants_on_a_graph_wider_than_the_cache(ctx, oracle, variant):
    """R-MAT scale 16 has more gathered columns than the LDS cache holds, so
    the cached variant reads hot columns from LDS and the rest from x; every
    variant must give the default launch's bits on integer-valued input."""
    off, tgt = W.rmat(16, 16, seed=7)
    n = len(off) - 1
    w = W.edge_weights(off, tgt, "int", seed=8, p0=1, p1=9)
    g = W.Graph(ctx, n, off, tgt, w)
    rng = np.random.default_rng(9)
    for sr, dtype in [("plus_times", "f32"), ("plus_times", "f64"), ("min_plus", "u32"),
                      ("min_plus", "i64"), ("lor_land", "bool"), ("max_times", "f64")]:
        x = values(dtype, n, rng, ADD_OF[sr])
        for transpose in (False, True):
            prod = W.Product(sr, transpose=transpose)
            xv = W.Vector(ctx, x, dtype=dtype)
            W.set_spmv_variant(ctx, 0)
            want = W.spmv(ctx, g, prod, xv, W.Vector(ctx, dtype=dtype, id_space=n)).numpy()
            W.set_spmv_variant(ctx, variant)
            got = W.spmv(ctx, g, prod, xv, W.Vector(ctx, dtype=dtype, id_space=n)).numpy()
            W.set_spmv_variant(ctx, -1)
            assert got.tobytes() == want.tobytes(), (sr, dtype, transpose)
