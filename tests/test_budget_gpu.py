"""The compute memory budget on the device (ComputeMemory, compute/memory.rs;
WorkingSet admission, spectral/working_set.rs:114-128):

* an eigenproblem given a budget one byte short of its working set is
  refused before its first step, and one of exactly its working set runs
  (spectral/lanczos/tests.rs:406-415), whatever max_iter asks
  (tests.rs:392-404); nothing stays charged after either;
* a vector, a graph or a product scratch beyond the budget is refused with
  memory_limit and leaves the bytes in use as they were.
"""
import numpy as np
import pytest

import pg_warren_amd as W

pytestmark = pytest.mark.gpu


def path(oracle, n):
    src = np.concatenate([np.arange(n - 1), np.arange(1, n)])
    dst = np.concatenate([np.arange(1, n), np.arange(n - 1)])
    return oracle.csr_from_edges(n, src, dst, None)


@pytest.mark.parametrize("matrix,k,precision,max_iter", [
    ("random_walk", 3, "f64", None),          # tests.rs:406-415
    ("adjacency", 10, "f32", 2000),           # tests.rs:392-404: refused before 2,000 steps
    ("normalized_laplacian", 5, "f64", None),
])
def test_eigen_one_byte_short_of_its_working_set_is_refused(oracle, matrix, k, precision, max_iter):
    n = 200 if k == 3 else 500
    csr = path(oracle, n)
    ctx = W.Context(0)
    g = W.Graph(ctx, n, csr.offsets, csr.targets, None)
    kw = dict(matrix=matrix, which="largest", precision=precision, vectors=False, tol=0.0 if max_iter else 1e-6,
              max_iter=max_iter)
    W.eigen(ctx, g, k, **kw)  # first call: the graph's lazily built caches
    base = ctx.memory_in_use()
    vals, _, _, info = W.eigen(ctx, g, k, **kw)
    needed = info["working_set"]
    assert needed > 0 and ctx.memory_in_use() == base
    ctx.set_memory_limit(base + needed - 1)
    with pytest.raises(W.WarrenError) as e:
        W.eigen(ctx, g, k, **kw)
    assert e.value.name == "memory_limit"
    assert e.value.working_set == needed
    assert ctx.memory_in_use() == base
    ctx.set_memory_limit(base + needed)
    vals2, _, _, info2 = W.eigen(ctx, g, k, **kw)
    assert info2["iterations"] > 0 and info2["working_set"] == needed
    assert vals2.tobytes() == vals.tobytes()
    assert ctx.memory_in_use() == base
    ctx.set_memory_limit(0)
    g.close()
    ctx.close()


def test_working_set_counts_the_basis(oracle):
    """A larger basis needs n * size more elements: the reported working
    set grows by exactly the basis, rotation and projected-eigenproblem
    bytes (the device solve's workspace holds 2 size^2 + size + 1 f64)."""
    n = 300
    csr = path(oracle, n)
    ctx = W.Context(0)
    g = W.Graph(ctx, n, csr.offsets, csr.targets, None)
    sets = {}
    for size in (30, 40):
        W.eigen(ctx, g, 4, precision="f64", basis_size=size, vectors=False)
        sets[size] = W.eigen(ctx, g, 4, precision="f64", basis_size=size, vectors=False)[3]["working_set"]
    ws = lambda m: (2 * m * m + m + 1) * 8
    assert sets[40] - sets[30] == n * 10 * 8 + (40 * 40 - 30 * 30) * 8 + ws(40) - ws(30)
    g.close()
    ctx.close()


def test_vectors_and_graphs_beyond_the_budget_are_refused(oracle):
    ctx = W.Context(0, memory_limit=1 << 20)
    v = W.Vector(ctx, np.zeros(1000))
    base = ctx.memory_in_use()
    assert base >= 8000
    with pytest.raises(W.WarrenError) as e:
        W.Vector(ctx, np.zeros(1 << 17))  # exactly 1 MiB more than fits with v held
    assert e.value.name == "memory_limit"
    assert ctx.memory_in_use() == base
    rng = np.random.default_rng(3)
    csr = oracle.csr_from_edges(50000, rng.integers(0, 50000, 200000), rng.integers(0, 50000, 200000), None)
    with pytest.raises(W.WarrenError) as e:
        W.Graph(ctx, csr.n, csr.offsets, csr.targets, None)
    assert e.value.name == "memory_limit"
    assert ctx.memory_in_use() == base
    # Raised, the same requests fit.
    ctx.set_memory_limit(0)
    big = W.Vector(ctx, np.zeros(1 << 17))
    g = W.Graph(ctx, csr.n, csr.offsets, csr.targets, None)
    x = W.Vector(ctx, np.ones(csr.n))
    y = W.Vector(ctx, dtype="f64", id_space=csr.n)
    W.spmv(ctx, g, W.Product("plus_times"), x, y)
    assert np.array_equal(y.numpy(), oracle.spmv(csr, np.ones(csr.n), "plus_times", "f64"))
    for h in (big, x, y, g, v):
        h.close()
    ctx.close()


def test_a_product_whose_scratch_does_not_fit_is_refused(oracle):
    """A fresh context's first product allocates its partials scratch within
    the budget: with the budget at what the graph and vectors hold, the
    product is refused, the result untouched and nothing left charged."""
    rng = np.random.default_rng(8)
    n = 40000
    csr = oracle.csr_from_edges(n, rng.integers(0, n, 400000), rng.integers(0, n, 400000), None)
    ctx = W.Context(0)
    g = W.Graph(ctx, n, csr.offsets, csr.targets, None)
    x = W.Vector(ctx, np.ones(n, np.float32), dtype="f32")
    y = W.Vector(ctx, np.full(n, 7.0, np.float32), dtype="f32")
    base = ctx.memory_in_use()
    ctx.set_memory_limit(base)
    with pytest.raises(W.WarrenError) as e:
        W.spmv(ctx, g, W.Product("plus_times"), x, y)
    assert e.value.name == "memory_limit"
    assert np.all(y.numpy() == 7.0)
    assert ctx.memory_in_use() == base
    ctx.set_memory_limit(0)
    W.spmv(ctx, g, W.Product("plus_times"), x, y)
    assert np.array_equal(y.numpy(), oracle.spmv(csr, np.ones(n, np.float32), "plus_times", "f32"))
    for h in (x, y, g):
        h.close()
    ctx.close()
