"""The oracle's gather (compute/gather/mod.rs), restated after the
reference's own naive version (compute/gather/reference.rs:103-177): join
the distinct (receiver, sender[, relation]) links of the selected edges
with the messages and fold each receiver's in collect order. Pure Python
over small inputs; test infrastructure only.

Edges are (relation, source, target) triples. Values are numpy scalars of
the messages' type, or bytes for byte messages (a list of bytes objects);
folds follow SemiringElement::apply: i64 sums checked, min and max in SQL
order (NaN above every number, the first operand kept on ties), bool or /
and; concat joins byte messages in collect order."""
import math

import numpy as np

COMBINERS = ["collect", "count", "sum", "min", "max", "first", "concat", "bool_or", "bool_and", "mode"]
I64_MIN, I64_MAX = -(2**63), 2**63 - 1


class GatherError(Exception):
    def __init__(self, name):
        super().__init__(name)
        self.name = name


def _kind(messages):
    if isinstance(messages, (list, tuple)) and all(isinstance(m, (bytes, bytearray)) for m in messages):
        return "bytes"
    dt = np.asarray(messages).dtype
    if dt == np.bool_:
        return "bool"
    return {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64", np.dtype(np.int64): "i64"}[dt]


def check(combiner, kind):
    """Combiner::check (combiner.rs:82-94): concat takes byte messages only,
    mode element messages only."""
    numeric = kind in ("f32", "f64", "i64")
    ok = {"collect": True, "count": True, "first": True, "sum": numeric, "min": numeric,
          "max": numeric, "bool_or": kind == "bool", "bool_and": kind == "bool",
          "concat": kind == "bytes", "mode": kind != "bytes"}[combiner]
    if not ok:
        raise GatherError("combiner_type")


def _nan(v):
    return isinstance(v, (float, np.floating)) and math.isnan(v)


def sql_lt(b, a):
    """b < a in SQL order (semiring.rs:264-278)."""
    if _nan(b):
        return False
    if _nan(a):
        return True
    return b < a


def sql_eq(a, b):
    return (_nan(a) and _nan(b)) or (not _nan(a) and not _nan(b) and a == b)


def _fold(combiner, kind, a, b):
    if combiner == "sum":
        if kind == "i64":
            r = int(a) + int(b)
            if r < I64_MIN or r > I64_MAX:
                raise GatherError("bigint_out_of_range")
            return np.int64(r)
        return a + b  # numpy scalar arithmetic in the messages' type
    if combiner == "min":
        return b if sql_lt(b, a) else a
    if combiner == "max":
        return b if sql_lt(a, b) else a
    if combiner == "bool_or":
        return bool(a) or bool(b)
    return bool(a) and bool(b)  # bool_and


def gather(n, edges, senders, messages, combiner, *, direction="in", relations=None,
           per_relation=False, selects=None, include_self=False):
    """Rows [(receiver, value)] in ascending receiver order; value a list for
    collect, an int for count, else one message."""
    kind = _kind(messages)
    check(combiner, kind)
    for s in senders:
        if s >= n:
            raise GatherError("sender_outside_id_space")
    links = set()
    for rel, src, dst in edges:
        if relations is not None and rel not in relations:
            continue
        r = rel if per_relation else 0
        if direction != "out":
            links.add((dst, src, r))
        if direction != "in":
            links.add((src, dst, r))
    links = sorted(links)
    sent = list(zip((int(s) for s in senders), messages))
    rows = []
    for receiver in range(n):
        if selects is not None and not selects[receiver]:
            continue
        order = ([receiver] if include_self else []) + [s for (r, s, _) in links if r == receiver]
        received = [m for s in order for (t, m) in sent if t == s]
        if not received:
            continue
        if combiner == "collect":
            value = received
        elif combiner == "count":
            value = len(received)
        elif combiner == "first":
            value = received[0]
        elif combiner == "concat":
            value = b"".join(received)  # Bytes::concat (combiner.rs:175-191)
        elif combiner == "mode":
            votes = [sum(1 for r in received if sql_eq(r, m)) for m in received]
            best = 0
            for i in range(1, len(received)):
                if votes[i] > votes[best] or (votes[i] == votes[best] and sql_lt(received[i], received[best])):
                    best = i
            value = received[best]
        else:
            value = received[0]
            for m in received[1:]:
                value = _fold(combiner, kind, value, m)
        rows.append((receiver, value))
    return rows
