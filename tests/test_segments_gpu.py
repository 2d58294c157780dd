"""The device segment loader (wg_graph_load_segments) and encoder against the
oracle's codec: decoded CSRs equal, every error the same variant with the
same fields, encoded bytes equal."""
import struct

import numpy as np
import pytest

import pg_warren_amd as W
from oracle import segments as S

pytestmark = pytest.mark.gpu
H = S.HEADER


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def expected_message(e):
    text = e.name
    if e.fields:
        text += ": " + ", ".join(f"{k} {v}" for k, v in e.fields.items())
    return text


def check_same(ctx, oracle, n, seg, reverse=None):
    """Device decode equals the oracle's, or fails with the same variant."""
    print(f"segments n={n} sparse={'rows' in seg and seg['rows'] is not None}", flush=True)
    try:
        want = S.decode_csr(n, seg.get("rows"), seg["offsets"], seg["targets"], seg.get("weights"))
    except S.SegmentError as e:
        with pytest.raises(W.WarrenError) as g:
            W.Graph.load_segments(ctx, n, seg, reverse)
        assert g.value.name == ("invalid_segment" if e.kind == "segment" else "invalid_csr")
        assert str(g.value).endswith(expected_message(e)), (str(g.value), expected_message(e))
        return None
    g = W.Graph.load_segments(ctx, n, seg, reverse)
    off, tgt, w = g.forward()
    assert np.array_equal(off, want[0]) and np.array_equal(tgt, want[1])
    if want[2] is not None:
        assert np.array_equal(w.view(np.uint64), want[2].view(np.uint64))
    ref = oracle.Csr(n, want[0], want[1], want[2])
    rref = ref.transpose()
    roff, rtgt, rw = g.reverse()
    assert np.array_equal(roff, rref.offsets) and np.array_equal(rtgt, rref.targets)
    return g


def csr_example():
    return S.encode_csr([0, 1, 1, 2], [2, 1], [1.5, 2.5])


def sparse_example():
    return S.encode_csr([0, 2, 3], [1, 999, 3], [0.5, 1.5, 2.5], rows=[3, 900])


def test_reference_examples_round_trip(ctx, oracle):
    g = check_same(ctx, oracle, 3, csr_example())
    enc = g.encode_segments()
    assert enc["offsets"] == csr_example()["offsets"] and enc["targets"] == csr_example()["targets"]
    assert enc["weights"] == csr_example()["weights"]
    g = check_same(ctx, oracle, 1000, sparse_example())
    # Encoded with dense offsets: the oracle's dense form of the same CSR.
    off, tgt, w = g.forward()
    assert g.encode_segments() == S.encode_csr(off, tgt, w)


def mutate(seg, key, at, value):
    b = bytearray(seg[key])
    b[at:at + len(value)] = value
    out = dict(seg)
    out[key] = bytes(b)
    return out


def test_every_error_as_the_oracle(ctx, oracle):
    seg, sp = csr_example(), sparse_example()
    cases = [
        (3, dict(seg, offsets=seg["offsets"][:10])),
        (3, mutate(seg, "offsets", 0, b"X")),
        (3, mutate(seg, "offsets", 4, b"\x09")),
        (3, mutate(seg, "offsets", 4, b"\x01")),
        (3, dict(seg, offsets=seg["targets"])),
        (3, mutate(seg, "offsets", 6, b"\x04")),
        (3, dict(seg, offsets=seg["offsets"] + b"\x00")),
        (3, mutate(seg, "offsets", 8, struct.pack("<Q", 2**64 - 1))),
        (3, mutate(seg, "offsets", 7, b"\x01")),
        (3, mutate(seg, "offsets", H + 8, struct.pack("<Q", 5))),
        (3, mutate(seg, "weights", 5, b"\x02")),
        (3, dict(seg, targets=S.encode(S.TARGETS, np.array([2, 1], "<u8")))),
        (1000, dict(sp, rows=sp["targets"])),
        (1000, dict(sp, rows=None)),
        (1000, mutate(sp, "rows", H, struct.pack("<I", 901))),
        (10, S.encode_csr([0, 2, 2], [1, 2], rows=[3, 5])),
        (10, S.encode_csr([0, 1, 2], [1, 2], rows=[3, 12])),
        (3, S.encode_csr([1, 1, 1, 2], [0, 1])),
        (3, S.encode_csr([0, 1, 1, 3], [0, 1])),
        (3, S.encode_csr([0, 2, 2, 3], [1, 5, 2])),
        (3, S.encode_csr([0, 2, 2, 3], [2, 1, 7])),
        (3, S.encode_csr([0, 2, 2, 3], [1, 1, 0])),
        (3, S.encode_csr([0, 2, 2, 3], [9, 1, 0])),
        (3, S.encode_csr([0, 2, 2, 3], [2, 1, 9])),
        (3, dict(S.encode_csr([0, 1, 1, 2], [2, 1]), weights=S.encode(S.WEIGHTS, np.array([1.0], "<f8")))),
        (5, S.encode_csr([0] * 6, [])),
    ]
    for n, case in cases:
        check_same(ctx, oracle, n, case)
    # The context stays usable after every refusal.
    check_same(ctx, oracle, 3, csr_example())


@pytest.mark.parametrize("scale,weighted,sparse", [(12, True, False), (14, False, True), (16, True, True)])
def test_rmat_segments(ctx, oracle, scale, weighted, sparse):
    off, tgt = W.rmat(scale, 8, seed=scale)
    n = len(off) - 1
    w = W.edge_weights(off, tgt, "uniform01", seed=3) if weighted else None
    seg = S.sparse_form(off, tgt, w) if sparse else S.encode_csr(off, tgt, w)
    g = check_same(ctx, oracle, n, seg)
    assert g.encode_segments() == S.encode_csr(off, tgt, w)
    # A faulty row deep in the graph: the first one is named.
    bad = tgt.copy()
    rows = np.flatnonzero(np.diff(off.astype(np.int64)) >= 2)
    r1, r2 = int(rows[len(rows) // 2]), int(rows[-1])
    bad[int(off[r1])], bad[int(off[r1]) + 1] = bad[int(off[r1]) + 1], bad[int(off[r1])]
    bad[int(off[r2])] = n + 7
    seg = S.sparse_form(off, bad, w) if sparse else S.encode_csr(off, bad, w)
    check_same(ctx, oracle, n, seg)


def test_reverse_segments_are_used(ctx, oracle):
    off, tgt = W.rmat(13, 8, seed=2)
    n = len(off) - 1
    w = W.edge_weights(off, tgt, "uniform01", seed=4)
    ref = oracle.Csr(n, off, tgt, w)
    rref = ref.transpose()
    g = W.Graph.load_segments(ctx, n, S.encode_csr(off, tgt, w),
                              S.encode_csr(rref.offsets, rref.targets, rref.weights))
    roff, rtgt, rw = g.reverse()
    assert np.array_equal(roff, rref.offsets) and np.array_equal(rtgt, rref.targets)
    assert g.encode_segments(reverse=True) == S.encode_csr(rref.offsets, rref.targets, rref.weights)
    # A transposed product reads it.
    x = np.random.default_rng(1).integers(0, 9, n).astype(np.float64)
    xv, yv = W.Vector(ctx, x), W.Vector(ctx, np.zeros(n))
    W.spmv(ctx, g, W.Product("min_plus", transpose=True), xv, yv)
    assert np.array_equal(yv.numpy(), oracle.spmv(ref, x, "min_plus", rev=rref, transpose=True))


def test_relation_csrs_store_sparse_offsets_where_they_pay(ctx):
    """graph/csr/graph_csr.rs:769-800: over 1,000 nodes relation a has rows 3
    and 500 and is written with sparse offsets both ways; relation b has a
    row from every node and is written dense; a one-relation graph is
    always dense. The segments decode back to the relation's CSR."""
    a_edges = [(500, 9, 1.0), (3, 7, 2.0), (500, 7, 3.0), (500, 9, 4.0)]
    src = [e[0] for e in a_edges] + list(range(1000))
    dst = [e[1] for e in a_edges] + [(v + 1) % 1000 for v in range(1000)]
    w = [e[2] for e in a_edges] + [1.0] * 1000
    rel = [0] * 4 + [1] * 1000
    g = W.Graph.build(ctx, 1000, np.array(src), np.array(dst), np.array(w), relations=np.array(rel),
                      relation_count=2)
    fa = g.encode_relation_segments(0)
    assert np.frombuffer(fa["rows"][16:], "<u4").tolist() == [3, 500]
    assert np.frombuffer(fa["offsets"][16:], "<u8").tolist() == [0, 1, 3]
    assert np.frombuffer(fa["targets"][16:], "<u4").tolist() == [7, 7, 9]
    assert np.frombuffer(fa["weights"][16:], "<f8").tolist() == [2.0, 3.0, 5.0]
    ra = g.encode_relation_segments(0, reverse=True)
    assert np.frombuffer(ra["rows"][16:], "<u4").tolist() == [7, 9]
    assert np.frombuffer(ra["offsets"][16:], "<u8").tolist() == [0, 2, 3]
    assert np.frombuffer(ra["targets"][16:], "<u4").tolist() == [3, 500, 500]
    for rev in (False, True):
        b = g.encode_relation_segments(1, reverse=rev)
        assert "rows" not in b and len(b["offsets"]) == 16 + 1001 * 8
    # Bytes equal to the oracle's encoding of the same CSR, and a round trip.
    ref = S.sparse_form([0] * 4 + [1] * 497 + [3] * 500, [7, 7, 9], [2.0, 3.0, 5.0])
    assert fa == ref
    back = W.Graph.load_segments(ctx, 1000, fa)
    off, tgt, wt = back.forward()
    assert off[501] == 3 and off[3] == 0 and off[4] == 1 and list(tgt) == [7, 7, 9]
    one = W.Graph.build(ctx, 1000, np.array([3]), np.array([7]), np.array([1.0]))
    assert "rows" not in one.encode_relation_segments(0)
    with pytest.raises(W.WarrenError) as e:
        g.encode_relation_segments(2)
    assert e.value.name == "unknown_relation"


@pytest.mark.parametrize("seed", range(3))
def test_random_relation_segments_match_the_oracle(ctx, oracle, seed):
    """Random relations over random id spaces: each relation's segments,
    both ways, equal the oracle's encoding in the reference's layout."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 3000))
    R = 3
    m = int(rng.integers(1, 4 * n))
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    # Relation 0 sparse-sized, the others spread.
    rel = np.where(rng.random(m) < 0.05, 0, rng.integers(1, R, m))
    src = np.where(rel == 0, rng.integers(0, max(n // 20, 1), m), src)
    w = rng.uniform(-2, 2, m)
    g = W.Graph.build(ctx, n, src, dst, w, relations=rel, relation_count=R)
    csrs = oracle.relation_csrs(n, src, dst, w, rel, R)
    for r in range(R):
        c = csrs[r]
        want_f = S.relation_form(n, c.offsets, c.targets, c.weights, relations=R)
        t = c.transpose()
        want_r = S.relation_form(n, t.offsets, t.targets, t.weights, relations=R)
        assert g.encode_relation_segments(r) == want_f, r
        assert g.encode_relation_segments(r, reverse=True) == want_r, r
