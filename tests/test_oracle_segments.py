"""The oracle's segment codec pinned by the reference's own bytes and errors
(graph/segment/mod.rs:371-531), plus round trips."""
import struct

import numpy as np
import pytest

from oracle import segments as S

H = S.HEADER


def csr_example():
    # mod.rs:375-384: edges 0 -> 2 (1.5), 2 -> 1 (2.5) over 3 nodes.
    return S.encode_csr([0, 1, 1, 2], [2, 1], [1.5, 2.5])


def sparse_example():
    # mod.rs:412-419: rows 3 and 900 of 1,000.
    return S.encode_csr([0, 2, 3], [1, 999, 3], [0.5, 1.5, 2.5], rows=[3, 900])


def err(fn, *a, **kw):
    with pytest.raises(S.SegmentError) as e:
        fn(*a, **kw)
    return e.value.name, e.value.fields


def test_a_segment_starts_with_its_header():
    seg = csr_example()
    assert seg["offsets"][:8] == b"WRSG\x02\x01\x08\x00"
    assert seg["offsets"][8:16] == struct.pack("<Q", 4)
    assert seg["targets"][4:8] == b"\x02\x02\x04\x00"
    assert len(seg["offsets"]) == H + 4 * 8


def test_round_trips():
    seg = csr_example()
    off, tgt, w = S.decode_csr(3, None, seg["offsets"], seg["targets"], seg["weights"])
    assert list(off) == [0, 1, 1, 2] and list(tgt) == [2, 1] and list(w) == [1.5, 2.5]
    sp = sparse_example()
    assert sp["rows"][:8] == b"WRSG\x02\x0b\x04\x00" and sp["rows"][8:16] == struct.pack("<Q", 2)
    assert len(sp["rows"]) == H + 8 and len(sp["offsets"]) == H + 3 * 8
    off, tgt, w = S.decode_csr(1000, sp["rows"], sp["offsets"], sp["targets"], sp["weights"])
    assert off[3] == 0 and off[4] == 2 and off[900] == 2 and off[901] == 3 and off[1000] == 3
    assert list(tgt) == [1, 999, 3]


def test_rows_segment_checks():
    # mod.rs:445-462
    sp = sparse_example()
    dec = lambda rows: S.decode_csr(1000, rows, sp["offsets"], sp["targets"], sp["weights"])
    assert err(dec, sp["targets"]) == ("SegmentKind", {"expected": 11, "found": 2})
    assert err(dec, None) == ("OffsetsLength", {"expected": 1001, "found": 3})
    bad = bytearray(sp["rows"])
    bad[H:H + 4] = struct.pack("<I", 901)
    assert err(dec, bytes(bad)) == ("StoredRowsNotAscending", {"index": 1})


def test_header_checks():
    # mod.rs:464-483, 520-531
    seg = csr_example()
    dec = lambda b: S.decode_csr(3, None, b, seg["targets"], None)
    assert err(dec, seg["offsets"][:10])[0] == "SegmentTruncated"
    assert err(dec, b"X" + seg["offsets"][1:])[0] == "SegmentMagic"
    b = bytearray(seg["offsets"]); b[4] = 9
    assert err(dec, bytes(b)) == ("SegmentFormat", {"found": 9})
    b[4] = 1
    assert err(dec, bytes(b)) == ("SegmentFormat", {"found": 1})  # before sparse offsets
    assert err(dec, seg["targets"]) == ("SegmentKind", {"expected": 1, "found": 2})
    b = bytearray(seg["offsets"]); b[6] = 4
    assert err(dec, bytes(b)) == ("SegmentWidth", {"expected": 8, "found": 4})
    assert err(dec, seg["offsets"] + b"\x00") == ("SegmentLength", {"expected": 32, "found": 33})
    b = bytearray(seg["offsets"]); b[8:16] = struct.pack("<Q", 2**64 - 1)
    assert err(dec, bytes(b))[0] == "SegmentLength"
    b = bytearray(seg["offsets"]); b[7] = 1
    assert err(dec, bytes(b)) == ("SegmentReserved", {"found": 1})
    # mod.rs:520-528: u64 targets for a small id space.
    wide = S.encode(S.TARGETS, np.array([2], "<u8"))
    assert err(S.decode_csr, 3, None, S.encode(S.OFFSETS, np.array([0, 1, 1, 1], "<u8")), wide) == \
        ("SegmentWidth", {"expected": 4, "found": 8})
    # mod.rs:530-538
    b = bytearray(seg["offsets"]); b[H + 8:H + 16] = struct.pack("<Q", 5)
    assert err(dec, bytes(b)) == ("OffsetsNotMonotone", {"row": 1})


def test_csr_invariants_in_order():
    # graph/csr/mod.rs:78-137 (from_parts) and :239-258 (stored rows).
    enc = S.encode_csr
    assert err(S.decode_csr, 3, None, **_kw(enc([1, 1, 1, 2], [0, 1]))) == ("FirstOffsetNotZero", {})
    assert err(S.decode_csr, 3, None, **_kw(enc([0, 1, 1, 3], [0, 1]))) == \
        ("FinalOffset", {"expected": 2, "found": 3})
    assert err(S.decode_csr, 3, None, **_kw(enc([0, 2, 2, 3], [1, 5, 2]))) == \
        ("TargetOutsideIdSpace", {"row": 0, "target": 5})
    assert err(S.decode_csr, 3, None, **_kw(enc([0, 2, 2, 3], [2, 1, 7]))) == ("RowNotSorted", {"row": 0})
    assert err(S.decode_csr, 3, None, **_kw(enc([0, 2, 2, 3], [1, 1, 0]))) == ("RowNotSorted", {"row": 0})
    # A row with both faults names the target; an earlier unsorted row wins.
    assert err(S.decode_csr, 3, None, **_kw(enc([0, 2, 2, 3], [9, 1, 0]))) == \
        ("TargetOutsideIdSpace", {"row": 0, "target": 9})
    assert err(S.decode_csr, 3, None, **_kw(enc([0, 2, 2, 3], [2, 1, 9]))) == ("RowNotSorted", {"row": 0})
    w = S.encode(S.WEIGHTS, np.array([1.0], "<f8"))
    seg = enc([0, 1, 1, 2], [2, 1])
    assert err(S.decode_csr, 3, None, seg["offsets"], seg["targets"], w) == \
        ("WeightsLength", {"expected": 2, "found": 1})
    sp = enc([0, 2, 2], [1, 2], rows=[3, 5])
    assert err(S.decode_csr, 10, sp["rows"], sp["offsets"], sp["targets"]) == ("StoredRowEmpty", {"row": 5})
    sp = enc([0, 1, 2], [1, 2], rows=[3, 12])
    assert err(S.decode_csr, 10, sp["rows"], sp["offsets"], sp["targets"]) == \
        ("StoredRowOutsideIdSpace", {"row": 12})


def _kw(seg):
    return {"offsets": seg["offsets"], "targets": seg["targets"]}


def test_random_round_trips_dense_and_sparse():
    rng = np.random.default_rng(3)
    for n in (1, 17, 500):
        deg = np.minimum(rng.integers(0, 6, n) * (rng.random(n) < 0.6), n)
        off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
        tgt = np.concatenate([np.sort(rng.choice(n, d, replace=False)) for d in deg]).astype(np.uint32) \
            if deg.sum() else np.zeros(0, np.uint32)
        w = rng.random(len(tgt))
        for seg in (S.encode_csr(off, tgt, w), S.sparse_form(off, tgt, w)):
            o, t, ww = S.decode_csr(n, seg.get("rows"), seg["offsets"], seg["targets"], seg["weights"])
            assert np.array_equal(o, off) and np.array_equal(t, tgt) and np.array_equal(ww, w)


def test_stores_sparse_thresholds():
    """graph/csr/sparse.rs:64-79: 1,000 nodes; 12 bytes a stored row at u32
    targets, 16 at u64, plus 8."""
    assert S.stores_sparse(0, 1000)
    assert S.stores_sparse(166, 1000)
    assert not S.stores_sparse(167, 1000)
    assert S.stores_sparse(124, 1000, target_bytes=8)
    assert not S.stores_sparse(125, 1000, target_bytes=8)
