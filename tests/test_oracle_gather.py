"""The oracle's gather pinned by the reference's known answers
(compute/gather/tests.rs:62-205) and a per-relation identity."""
import math

import numpy as np
import pytest

from oracle import gather as G

# tests.rs:62-67: relation 0: 0 -> 1, 0 -> 2, 1 -> 2, 2 -> 0; relation 1:
# 0 -> 1. Node 3 is isolated.
EDGES = [(0, 0, 1), (0, 0, 2), (0, 1, 2), (0, 2, 0), (1, 0, 1)]
SENT = [(2, 3.0), (0, 1.0), (2, 4.0), (1, 2.0), (3, 9.0)]


def run(sent, combiner, dtype=np.float64, **kw):
    s = [a for a, _ in sent]
    m = np.array([b for _, b in sent], dtype=dtype)
    return G.gather(4, EDGES, s, m, combiner, **kw)


def vals(rows):
    return [(r, list(v) if isinstance(v, list) else v) for r, v in rows]


def test_collect_in_each_direction():
    assert vals(run(SENT, "collect")) == [(0, [3.0, 4.0]), (1, [1.0]), (2, [1.0, 2.0])]
    assert vals(run(SENT, "collect", direction="out")) == [(0, [2.0, 3.0, 4.0]), (1, [3.0, 4.0]), (2, [1.0])]
    assert vals(run(SENT, "collect", direction="both")) == \
        [(0, [2.0, 3.0, 4.0]), (1, [1.0, 3.0, 4.0]), (2, [1.0, 2.0])]


def test_numeric_combiners():
    assert run(SENT, "count") == [(0, 2), (1, 1), (2, 2)]
    assert vals(run(SENT, "sum")) == [(0, 7.0), (1, 1.0), (2, 3.0)]
    assert vals(run(SENT, "min")) == [(0, 3.0), (1, 1.0), (2, 1.0)]
    assert vals(run(SENT, "max")) == [(0, 4.0), (1, 1.0), (2, 2.0)]
    assert vals(run(SENT, "first")) == [(0, 3.0), (1, 1.0), (2, 1.0)]
    assert vals(run(SENT, "sum", np.int64)) == [(0, 7), (1, 1), (2, 3)]
    assert vals(run(SENT, "max", np.float32)) == [(0, 4.0), (1, 1.0), (2, 2.0)]


def test_sql_order_and_overflow():
    nan = [(2, float("nan")), (2, 1.0)]
    assert vals(run(nan, "min")) == [(0, 1.0)]
    assert math.isnan(run(nan, "max")[0][1])
    with pytest.raises(G.GatherError) as e:
        run([(2, 2**63 - 1), (2, 1)], "sum", np.int64)
    assert e.value.name == "bigint_out_of_range"


def test_bool_combiners():
    sent = [(2, True), (0, False), (2, False), (1, True)]
    assert vals(run(sent, "bool_or", np.bool_)) == [(0, True), (1, False), (2, True)]
    assert vals(run(sent, "bool_and", np.bool_)) == [(0, False), (1, False), (2, False)]
    assert vals(run(sent, "first", np.bool_)) == [(0, True), (1, False), (2, False)]


def test_mode():
    sent = [(0, 5.0), (1, 5.0), (2, 7.0), (2, 5.0), (2, 7.0)]
    assert vals(run(sent, "mode")) == [(0, 7.0), (1, 5.0), (2, 5.0)]
    assert run(sent, "mode", direction="both")[0] == (0, 5.0)
    assert vals(run(sent, "mode", np.int64)) == [(0, 7), (1, 5), (2, 5)]
    flags = [(2, True), (2, False), (2, True), (0, False), (1, True)]
    assert vals(run(flags, "mode", np.bool_)) == [(0, True), (1, False), (2, False)]
    z = run([(2, float("nan")), (2, -0.0), (2, 0.0), (2, float("nan")), (2, 1.0), (2, 1.0)], "mode")
    assert z[0][1] == 0.0 and math.copysign(1.0, z[0][1]) < 0
    assert math.isnan(run([(2, float("nan")), (2, 1.0), (2, float("nan"))], "mode")[0][1])


def test_combiner_type_and_senders():
    with pytest.raises(G.GatherError) as e:
        run([(0, True)], "sum", np.bool_)
    assert e.value.name == "combiner_type"
    with pytest.raises(G.GatherError) as e:
        run([(4, 1.0)], "count")
    assert e.value.name == "sender_outside_id_space"


def test_per_relation_sends_once_per_relation():
    # 0 -> 1 is in both relations: per relation, 1 hears 0 twice.
    merged = dict(vals(run(SENT, "collect")))
    per = dict(vals(run(SENT, "collect", per_relation=True)))
    assert merged[1] == [1.0] and per[1] == [1.0, 1.0]
    # Both directions of one relation are still one link.
    both = dict(vals(run(SENT, "collect", direction="both", per_relation=True)))
    assert both[0] == [2.0, 2.0, 3.0, 4.0] and both[1] == [1.0, 1.0, 3.0, 4.0]


def test_byte_messages():
    """tests.rs:173-202: byte messages collect, concatenate, take the first
    and count; sum is refused for them, concat for numbers."""
    sent = [(2, b"c"), (0, b"a"), (2, b"d"), (1, b""), (3, b"z")]
    s = [a for a, _ in sent]
    m = [b for _, b in sent]
    assert G.gather(4, EDGES, s, m, "collect") == [(0, [b"c", b"d"]), (1, [b"a"]), (2, [b"a", b""])]
    assert G.gather(4, EDGES, s, m, "concat") == [(0, b"cd"), (1, b"a"), (2, b"a")]
    assert G.gather(4, EDGES, s, m, "first") == [(0, b"c"), (1, b"a"), (2, b"a")]
    assert G.gather(4, EDGES, s, m, "count") == [(0, 2), (1, 1), (2, 2)]
    for comb in ("sum", "mode", "bool_or"):
        with pytest.raises(G.GatherError) as e:
            G.gather(4, EDGES, s, m, comb)
        assert e.value.name == "combiner_type"
    with pytest.raises(G.GatherError) as e:
        run(SENT, "concat")
    assert e.value.name == "combiner_type"
