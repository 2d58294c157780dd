"""The oracle's restatement of a CSR's segment codec (graph/segment/mod.rs,
graph/segment/csr.rs) and of the checks a decode makes
(Csr::from_segments, graph/csr/mod.rs:78-137, 239-258). Test
infrastructure only: it builds segment bytes for tests and names the first
violation a decode must report, in the reference's order.

Each segment: a 16-byte little-endian header (magic b"WRSG", format 2 for a
CSR's segments, kind, element width, a reserved 0 byte, the u64 element
count) and its payload. Kinds: offsets 1 (u64), targets 2 (u32 below a
2^32 id space), weights 3 (f64), rows 11 (the stored rows of sparse
offsets)."""
import struct

import numpy as np

HEADER = 16
OFFSETS, TARGETS, WEIGHTS, ROWS = 1, 2, 3, 11
FORMAT = 2


class SegmentError(Exception):
    """A decode's first violation: ``name`` is the GraphError variant
    (graph/error.rs), ``fields`` its values, ``kind`` 'segment' for the
    framing or 'csr' for the CSR's invariants."""

    def __init__(self, name, kind, **fields):
        self.name, self.kind, self.fields = name, kind, fields
        super().__init__(f"{name} {fields}")


def encode(kind, elements, width=None):
    """encode (mod.rs:300-316): the header and the little-endian payload."""
    a = np.ascontiguousarray(elements)
    width = a.dtype.itemsize if width is None else width
    return b"WRSG" + bytes([FORMAT, kind, width, 0]) + struct.pack("<Q", len(a)) + a.tobytes()


def encode_csr(offsets, targets, weights=None, rows=None):
    """encode_csr (csr.rs:130-137): the CSR's segments as a dict."""
    out = {"offsets": encode(OFFSETS, np.asarray(offsets, "<u8")),
           "targets": encode(TARGETS, np.asarray(targets, "<u4")),
           "weights": None if weights is None else encode(WEIGHTS, np.asarray(weights, "<f8"))}
    if rows is not None:
        out["rows"] = encode(ROWS, np.asarray(rows, "<u4"))
    return out


def stores_sparse(stored_rows, id_space, target_bytes=4):
    """stores_sparse (graph/csr/sparse.rs:37-41): sparse offsets -- a row id
    of the target width and a u64 offset per stored row, and the end -- when
    they take at most a quarter of the bytes of dense ones."""
    sparse = stored_rows * (target_bytes + 8) + 8
    dense = (id_space + 1) * 8
    return sparse * 4 <= dense


def relation_form(id_space, offsets, targets, weights=None, relations=2):
    """A relation's CSR encoded as the reference lays it out: sparse where
    stores_sparse says so (never for a graph's only relation)."""
    off = np.asarray(offsets, np.uint64)
    stored = int(np.count_nonzero(np.diff(off.astype(np.int64)) > 0))
    if relations > 1 and stores_sparse(stored, id_space):
        return sparse_form(off, targets, weights)
    return encode_csr(off, targets, weights)


def sparse_form(offsets, targets, weights=None):
    """The same CSR with sparse offsets: the ids of its non-empty rows and an
    offset per stored row plus the end (Csr::from_sparse_parts)."""
    off = np.asarray(offsets, np.uint64)
    deg = np.diff(off.astype(np.int64))
    rows = np.flatnonzero(deg > 0).astype(np.uint32)
    sparse = np.concatenate([[0], np.cumsum(deg[rows])]).astype(np.uint64)
    return encode_csr(sparse, targets, weights, rows=rows)


def _header(b, kind):
    """read_header (mod.rs:322-343)."""
    if b is None or len(b) < HEADER:
        raise SegmentError("SegmentTruncated", "segment")
    if b[:4] != b"WRSG":
        raise SegmentError("SegmentMagic", "segment")
    if b[4] != FORMAT:
        raise SegmentError("SegmentFormat", "segment", found=b[4])
    if b[5] != kind:
        raise SegmentError("SegmentKind", "segment", expected=kind, found=b[5])
    if b[7] != 0:
        raise SegmentError("SegmentReserved", "segment", found=b[7])
    return b[6], struct.unpack("<Q", b[8:16])[0]


def _elements(b, width, count, dtype):
    """expect_width and expect_length (mod.rs:346-365), then the payload."""
    if width != np.dtype(dtype).itemsize:
        raise SegmentError("SegmentWidth", "segment", expected=np.dtype(dtype).itemsize, found=width)
    expected = min(count * np.dtype(dtype).itemsize, 2**64 - 1)
    if expected != len(b) - HEADER:
        raise SegmentError("SegmentLength", "segment", expected=expected, found=len(b) - HEADER)
    return np.frombuffer(b[HEADER:], dtype=dtype)


def decode_csr(id_space, rows, offsets, targets, weights=None):
    """AnyCsr::decode (csr.rs:84-127) then Csr::from_segments: returns the
    CSR with dense offsets (offsets, targets, weights) or raises the first
    violation."""
    ow, oc = _header(offsets, OFFSETS)
    off = _elements(offsets, ow, oc, "<u8")
    w = None
    if weights is not None:
        ww, wc = _header(weights, WEIGHTS)
        w = _elements(weights, ww, wc, "<f8")
    rh = None if rows is None else _header(rows, ROWS)
    tw, tc = _header(targets, TARGETS)
    if id_space >= 2**32:
        raise SegmentError("Unsupported", "csr")
    tgt = _elements(targets, tw, tc, "<u4")
    stored = None if rh is None else _elements(rows, rh[0], rh[1], "<u4")
    # Csr::from_segments (graph/csr/mod.rs:78-137).
    expected = id_space + 1 if stored is None else len(stored) + 1
    if len(off) != expected:
        raise SegmentError("OffsetsLength", "csr", expected=expected, found=len(off))
    if off[0] != 0:
        raise SegmentError("FirstOffsetNotZero", "csr")
    row_id = np.arange(len(off) - 1, dtype=np.uint64) if stored is None else stored.astype(np.uint64)
    if stored is not None:
        # check_stored_rows (graph/csr/mod.rs:239-258).
        bad = np.flatnonzero(stored[1:] <= stored[:-1])
        if len(bad):
            raise SegmentError("StoredRowsNotAscending", "csr", index=int(bad[0]) + 1)
        if len(stored) and stored[-1] >= id_space:
            raise SegmentError("StoredRowOutsideIdSpace", "csr", row=int(stored[-1]))
        empty = np.flatnonzero(off[1:] <= off[:-1])
        if len(empty):
            raise SegmentError("StoredRowEmpty", "csr", row=int(stored[empty[0]]))
    down = np.flatnonzero(off[1:] < off[:-1])
    if len(down):
        raise SegmentError("OffsetsNotMonotone", "csr", row=int(row_id[down[0]]))
    if off[-1] != len(tgt):
        raise SegmentError("FinalOffset", "csr", expected=len(tgt), found=int(off[-1]))
    # Rows in order; within a row the target check comes first.
    pos_of = np.repeat(np.arange(len(off) - 1), np.diff(off.astype(np.int64)))
    out = np.flatnonzero(tgt >= id_space)
    unsorted = np.flatnonzero((pos_of[1:] == pos_of[:-1]) & (tgt[1:] <= tgt[:-1])) + 1
    first_out = pos_of[out[0]] if len(out) else len(off)
    first_unsorted = pos_of[unsorted[0]] if len(unsorted) else len(off)
    if first_out < len(off) and first_out <= first_unsorted:
        raise SegmentError("TargetOutsideIdSpace", "csr", row=int(row_id[first_out]), target=int(tgt[out[0]]))
    if first_unsorted < len(off):
        raise SegmentError("RowNotSorted", "csr", row=int(row_id[first_unsorted]))
    if w is not None and len(w) != len(tgt):
        raise SegmentError("WeightsLength", "csr", expected=len(tgt), found=len(w))
    if stored is None:
        dense = off.astype(np.uint64)
    else:
        deg = np.zeros(id_space, np.int64)
        deg[stored.astype(np.int64)] = np.diff(off.astype(np.int64))
        dense = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    return dense, tgt.astype(np.uint32), None if w is None else w.astype(np.float64)
