"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/*.h declares, and the synthetic-input generators are
deterministic. No compute call needs a GPU here."""
import ctypes
import glob
import os
import re

import numpy as np

import pg_warren_amd as W
from pg_warren_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|uint64_t|void\s*\*|double)\s*\*?\s*(wg_\w+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_entry_point():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for name in sorted(names):
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert names == bound, (names - bound, bound - names)


def test_rmat_is_deterministic_and_clean():
    off, tgt = W.rmat(10, 16, seed=3)
    off2, tgt2 = W.rmat(10, 16, seed=3, threads=1)
    assert np.array_equal(off, off2) and np.array_equal(tgt, tgt2)
    n = 1 << 10
    assert len(off) == n + 1 and off[0] == 0 and off[-1] == len(tgt)
    for r in range(n):
        row = tgt[off[r]:off[r + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)  # sorted, no duplicates
        assert not np.any(row == r)  # no self loops
    off3, _ = W.rmat(10, 16, seed=4)
    assert not np.array_equal(off, off3)


def test_rmat_symmetrized_is_symmetric():
    off, tgt = W.rmat(9, 8, seed=5, symmetrize=True)
    n = 1 << 9
    edges = set()
    for r in range(n):
        for t in tgt[off[r]:off[r + 1]]:
            edges.add((r, int(t)))
    assert all((b, a) in edges for a, b in edges)


def test_edge_weights_ranges():
    off, tgt = W.rmat(10, 8, seed=2)
    w = W.edge_weights(off, tgt, "uniform01")
    assert np.all((w > 0) & (w <= 1)) and np.all(w == w.astype(np.float32))
    w = W.edge_weights(off, tgt, "int", p0=1, p1=255)
    assert np.all((w >= 1) & (w <= 255) & (w == np.round(w)))
    w = W.edge_weights(off, tgt, "lognormal", p0=0.0, p1=1.0)
    assert np.all(w > 0) and abs(np.mean(np.log(w))) < 0.1 and abs(np.std(np.log(w)) - 1) < 0.1


def test_context_creation_reports_a_status():
    """On a machine without a GPU the context fails with device_error (16);
    with one it succeeds."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    code = lib.wg_context_create(0, 0, ctypes.byref(h))
    assert code in (0, 16)
    if code == 0:
        lib.wg_context_destroy(h)
