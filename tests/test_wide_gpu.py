"""Graphs with u64 targets (TargetWidth::U64, graph/csr/target.rs:16-31):
upload with Csr::from_parts' checks, the reverse CSR, and products.

* Small graphs held with u64 targets give the same products as with u32
  targets -- the oracle's, bit for bit, for every semiring and dtype, plain,
  transposed, masked and accumulated; float plus here folds each row left
  to right, so it equals the oracle's f32 / f64 left fold exactly.
* An id space of 2^32 + 16 nodes (u64 targets by the reference's own rule),
  a few hundred edges: the forward and reverse CSRs and a plus_times
  product against a closed form computed from the edge list.
"""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu

SEMIRINGS = {"f32": ["plus_times", "min_plus", "max_times", "plus_second", "min_first"],
             "f64": ["plus_times", "max_plus", "min_times", "plus_pair", "max_second"],
             "i64": ["plus_times", "min_plus", "max_plus", "any_pair", "plus_first"],
             "bool": ["lor_land", "plus_times", "min_second"],
             "u32": ["min_plus", "plus_times", "min_second"]}


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def random_csr(oracle, rng, n, m):
    src = rng.integers(0, n, m)
    dst = np.where(rng.random(m) < 0.3, rng.integers(0, 4, m), rng.integers(0, n, m))
    w = rng.integers(1, 6, m).astype(np.float64)
    return oracle.csr_from_edges(n, src, dst, w)


def values(rng, dtype, n):
    if dtype == "bool":
        return (rng.random(n) < 0.4).astype(np.uint8)
    if dtype == "u32":
        v = rng.integers(0, 50, n).astype(np.uint32)
        v[rng.random(n) < 0.1] = 2**32 - 1
        return v
    if dtype == "i64":
        return rng.integers(-20, 20, n).astype(np.int64)
    return rng.standard_normal(n).astype(np.float32 if dtype == "f32" else np.float64)


@pytest.mark.parametrize("dtype", ["f32", "f64", "i64", "bool", "u32"])
def test_u64_targets_give_the_oracles_products(ctx, oracle, dtype):
    rng = np.random.default_rng(len(dtype))
    csr = random_csr(oracle, rng, 300, 2500)
    gw = W.Graph(ctx, csr.n, csr.offsets, csr.targets, csr.weights, target_width=8)
    assert gw.wide
    rev = csr.transpose()
    roff, rtgt, rw = gw.read_u64(reverse=True)
    assert np.array_equal(roff, rev.offsets) and np.array_equal(rtgt, rev.targets)
    assert np.array_equal(rw, rev.weights)
    for sr in SEMIRINGS[dtype]:
        for transpose in (False, True):
            for width in (1, 3):
                x = values(rng, dtype, csr.n * width)
                mask = (rng.random(csr.n) < 0.6).astype(np.float32)
                y0 = values(rng, dtype, csr.n * width)
                for kw in ({}, {"mask": mask}, {"mask": mask, "replace": True, "accumulate": "plus"}):
                    if dtype in ("bool",) and kw.get("accumulate"):
                        continue
                    xv = W.Vector(ctx, x.reshape(csr.n, width) if width > 1 else x, dtype=dtype)
                    yv = W.Vector(ctx, y0.reshape(csr.n, width) if width > 1 else y0, dtype=dtype)
                    mv = None if "mask" not in kw else W.Vector(ctx, kw["mask"])
                    p = W.Product(sr, transpose=transpose, mask=mv, replace=kw.get("replace", False),
                                  accumulate=kw.get("accumulate"))
                    try:
                        (W.spmm if width > 1 else W.spmv)(ctx, gw, p, xv, yv)
                        got = yv.numpy().reshape(-1)
                        err = None
                    except W.WarrenError as e:
                        got, err = None, e.name
                    try:
                        want = oracle.spmv(csr, x, sr, dtype, rev=rev if transpose else None, transpose=transpose,
                                           mask=kw.get("mask"), replace=kw.get("replace", False),
                                           accumulate=kw.get("accumulate"), y=y0.copy(), width=width)
                        werr = None
                    except oracle.OracleError as e:
                        want, werr = None, e.name
                    assert err == werr, (sr, transpose, width, kw, err, werr)
                    if err is None:
                        assert got.tobytes() == want.tobytes(), (sr, dtype, transpose, width, kw)


@pytest.mark.parametrize("width", [4, 8])
def test_upload_checks_in_the_references_order(ctx, width):
    """Csr::from_parts' checks (graph/csr/mod.rs:97-126), u32 and u64
    targets alike, the per-row ones on the device: the first row with a
    fault, a target outside the id space named before the row's order."""
    off = np.array([0, 2, 3, 5], np.uint64)
    for tgt, name in [([1, 0, 1, 0, 2], "row 0 not sorted"), ([0, 5, 1, 0, 1], "target outside id space in row 0"),
                      ([0, 1, 7, 2, 1], "target outside id space in row 1"), ([0, 1, 2, 2, 1], "row 2 not sorted"),
                      ([2, 1, 9, 0, 0], "row 0 not sorted"), ([0, 9, 1, 1, 0], "target outside id space in row 0")]:
        with pytest.raises(W.WarrenError) as e:
            W.Graph(ctx, 3, off, np.array(tgt), target_width=width)
        assert e.value.name == "invalid_csr" and name in str(e.value), (tgt, str(e.value))
    with pytest.raises(W.WarrenError) as e:
        W.Graph(ctx, 3, np.array([0, 2, 1, 3], np.uint64), np.array([0, 1, 2]), target_width=width)
    assert e.value.name == "invalid_csr" and "monotone at row 1" in str(e.value)
    g = W.Graph(ctx, 3, off, np.array([0, 2, 1, 0, 2]), target_width=width)
    g.close()


def test_u64_refusals(ctx):
    """The entry points that do not read u64 targets refuse such a graph."""
    off = np.array([0, 2, 3], np.uint64)
    g = W.Graph(ctx, 2, off, np.array([0, 1, 0], np.uint64), target_width=8)
    for call in (lambda: W.pagerank(ctx, g), lambda: W.bfs(ctx, g, [0]),
                 lambda: W.gather(ctx, g, [0], np.array([1.0]), "sum"),
                 lambda: W.spmv(ctx, g, W.Product("plus_times", relations=[0]), W.Vector(ctx, np.ones(2)),
                                W.Vector(ctx, np.zeros(2)))):
        with pytest.raises(W.WarrenError) as e:
            call()
        assert e.value.name in ("unsupported", "unknown_relation")


@pytest.mark.slow
def test_id_space_beyond_u32(ctx):
    """2^32 + 16 nodes: u64 targets by the reference's rule. Edges among the
    first rows and the last ones; y = A x over f32 against the edge list."""
    n = (1 << 32) + 16
    rng = np.random.default_rng(32)
    rows = np.concatenate([np.arange(0, 40), np.arange(n - 16, n)])
    src, dst = [], []
    for r in rows:
        k = int(rng.integers(1, 6))
        t = np.unique(np.concatenate([rng.integers(0, 100, k), rng.integers(n - 16, n, k)]).astype(np.uint64))
        src += [int(r)] * len(t)
        dst += t.tolist()
    src = np.array(src, np.uint64)
    dst = np.array(dst, np.uint64)
    m = len(src)
    counts = np.bincount(np.searchsorted(rows, src), minlength=len(rows))
    offsets = np.zeros(n + 1, np.uint64)
    # Rows 0..39 hold their entries first, then rows n-16..n-1.
    offsets[1:41] = np.cumsum(counts[:40])
    offsets[41:n - 15] = offsets[40]
    offsets[n - 15:] = offsets[40] + np.cumsum(counts[40:])
    assert offsets[-1] == m
    w = rng.integers(1, 4, m).astype(np.float64)
    print(f"u64 id space n={n} m={m}", flush=True)
    g = W.Graph(ctx, n, offsets, dst, w)
    assert g.wide
    x = np.zeros(n, np.float32)
    touched = np.unique(dst)
    x[touched] = rng.integers(1, 9, len(touched)).astype(np.float32)
    xs = dict(zip(touched.tolist(), x[touched].tolist()))
    xv = W.Vector(ctx, x)
    del x
    yv = W.Vector(ctx, dtype="f32", id_space=n)
    W.spmv(ctx, g, W.Product("plus_times"), xv, yv)
    y = yv.numpy()
    want = {}
    for s, t, ww in zip(src.tolist(), dst.tolist(), w.tolist()):
        want[s] = want.get(s, 0.0) + np.float32(ww) * np.float32(xs[t])
    got_rows = np.flatnonzero(y)
    assert set(got_rows.tolist()) <= set(rows.tolist())
    for r, v in want.items():
        assert y[r] == np.float32(v), (r, y[r], v)
    yt = W.Vector(ctx, dtype="f32", id_space=n)
    W.spmv(ctx, g, W.Product("plus_times", transpose=True), xv, yt)
    yt = yt.numpy()
    for t in touched[-5:]:
        wantt = np.float32(0.0)
        for s, tt, ww in sorted(zip(src.tolist(), dst.tolist(), w.tolist())):
            if tt == t:
                wantt = np.float32(wantt + np.float32(ww) * np.float32(xs[s] if s in xs else 0.0))
        assert yt[t] == wantt, (t, yt[t], wantt)
    g.close()
