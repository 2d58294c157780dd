"""The device gather (wg_gather) against the oracle's gather: the reference's
known answers (compute/gather/tests.rs) and random graphs, messages,
relations, directions, masks and include_self, every combiner the device
takes; values compared bit for bit (floats in SQL order: NaN equals NaN)."""
import math

import numpy as np
import pytest

import pg_warren_amd as W
from oracle import gather as G

pytestmark = pytest.mark.gpu

EDGES = [(0, 0, 1), (0, 0, 2), (0, 1, 2), (0, 2, 0), (1, 0, 1)]
SENT = [(2, 3.0), (0, 1.0), (2, 4.0), (1, 2.0), (3, 9.0)]


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def graph_of(ctx, n, edges):
    rel, src, dst = (np.array(c, dtype=np.int64) for c in zip(*edges)) if edges else ([], [], [])
    nrel = int(max(rel) + 1) if len(edges) else 1
    return W.Graph.build(ctx, n, src, dst, relations=rel if nrel > 1 else None,
                         relation_count=max(nrel, 1))


def same_value(a, b):
    if isinstance(a, (float, np.floating)) and math.isnan(a):
        return isinstance(b, (float, np.floating)) and math.isnan(b)
    if isinstance(a, (float, np.floating)):
        return np.float64(a).tobytes() == np.float64(b).tobytes()
    return a == b


def compare(ctx, g, n, edges, senders, messages, combiner, **kw):
    mask = kw.pop("mask", None)
    try:
        want = G.gather(n, edges, senders, messages, combiner, selects=None if mask is None else mask != 0, **kw)
    except G.GatherError as e:
        with pytest.raises(W.WarrenError) as got:
            W.gather(ctx, g, senders, messages, combiner, mask=mask, **kw)
        assert got.value.name == e.name
        return
    recv, vals = W.gather(ctx, g, senders, messages, combiner, mask=mask, **kw)
    assert list(recv) == [r for r, _ in want], (combiner, kw)
    for (r, w), v in zip(want, vals):
        if combiner == "collect":
            assert len(w) == len(v) and all(same_value(a, b) for a, b in zip(w, v)), (r, w, v)
        else:
            assert same_value(w, v), (combiner, kw, r, w, v)


def test_known_answers(ctx):
    g = graph_of(ctx, 4, EDGES)
    s = [a for a, _ in SENT]
    m = np.array([b for _, b in SENT])
    r, v = W.gather(ctx, g, s, m, "collect")
    assert list(r) == [0, 1, 2] and [list(x) for x in v] == [[3.0, 4.0], [1.0], [1.0, 2.0]]
    r, v = W.gather(ctx, g, s, m, "collect", direction="out")
    assert [list(x) for x in v] == [[2.0, 3.0, 4.0], [3.0, 4.0], [1.0]]
    r, v = W.gather(ctx, g, s, m, "collect", direction="both")
    assert [list(x) for x in v] == [[2.0, 3.0, 4.0], [1.0, 3.0, 4.0], [1.0, 2.0]]
    assert list(W.gather(ctx, g, s, m, "count")[1]) == [2, 1, 2]
    assert list(W.gather(ctx, g, s, m, "sum")[1]) == [7.0, 1.0, 3.0]
    assert list(W.gather(ctx, g, s, m, "min")[1]) == [3.0, 1.0, 1.0]
    assert list(W.gather(ctx, g, s, m, "max")[1]) == [4.0, 1.0, 2.0]
    assert list(W.gather(ctx, g, s, m, "first")[1]) == [3.0, 1.0, 1.0]
    assert list(W.gather(ctx, g, s, m.astype(np.int64), "sum")[1]) == [7, 1, 3]
    with pytest.raises(W.WarrenError) as e:
        W.gather(ctx, g, [2, 2], np.array([2**63 - 1, 1], np.int64), "sum")
    assert e.value.name == "bigint_out_of_range"
    flags = np.array([True, False, True, False, True])
    r, v = W.gather(ctx, g, [2, 2, 2, 0, 1], flags, "mode")
    assert list(v) == [True, False, False]
    r, v = W.gather(ctx, g, [2] * 6, np.array([np.nan, -0.0, 0.0, np.nan, 1.0, 1.0]), "mode")
    assert v[0] == 0.0 and math.copysign(1.0, v[0]) < 0
    with pytest.raises(W.WarrenError) as e:
        W.gather(ctx, g, [0], np.array([True]), "sum")
    assert e.value.name == "combiner_type"
    with pytest.raises(W.WarrenError) as e:
        W.gather(ctx, g, [4], np.array([1.0]), "count")
    assert e.value.name == "sender_outside_id_space"


def test_known_answers_against_the_oracle(ctx):
    g = graph_of(ctx, 4, EDGES)
    s = [a for a, _ in SENT]
    for dtype in (np.float64, np.float32, np.int64):
        m = np.array([b for _, b in SENT]).astype(dtype)
        for comb in ("collect", "count", "sum", "min", "max", "first", "mode"):
            for direction in ("in", "out", "both"):
                for per in (False, True):
                    compare(ctx, g, 4, EDGES, s, m, comb, direction=direction, per_relation=per)


@pytest.mark.parametrize("seed", range(6))
def test_random_gathers(ctx, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 60))
    R = int(rng.integers(1, 4))
    m = int(rng.integers(0, 4 * n))
    edges = sorted({(int(rng.integers(0, R)), int(rng.integers(0, n)), int(rng.integers(0, n)))
                    for _ in range(m)})
    g = graph_of(ctx, n, edges) if edges else W.Graph(ctx, n, np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32))
    if not edges:
        edges = []
    k = int(rng.integers(0, 3 * n))
    senders = rng.integers(0, n, k)
    values = {"f64": rng.choice([0.5, -1.25, 3.0, np.nan, -0.0, 0.0, 1e300, 7.0], k),
              "f32": rng.choice([0.5, -1.25, 3.0, np.nan, -0.0, 2.0], k).astype(np.float32),
              "i64": rng.integers(-5, 6, k).astype(np.int64),
              "bool": rng.random(k) < 0.5}
    nrel = max(R, 1)
    print(f"gather seed={seed} n={n} edges={len(edges)} messages={k} relations={nrel}", flush=True)
    for trial in range(8):
        dt = ["f64", "f32", "i64", "bool"][trial % 4]
        combs = {"f64": ["collect", "count", "sum", "min", "max", "first", "mode"],
                 "f32": ["sum", "min", "max", "mode"], "i64": ["sum", "min", "mode", "count"],
                 "bool": ["bool_or", "bool_and", "first", "mode", "collect", "sum"]}[dt]
        direction = ["in", "out", "both"][trial % 3]
        rels = None if (trial % 2 or nrel == 1) else sorted(set(rng.integers(0, nrel, 2).tolist()))
        per = bool(trial % 3 == 1) and nrel > 1
        mask = (rng.random(n) < 0.7).astype(np.float64) if trial % 4 == 3 else None
        self_ = bool(trial % 2)
        for comb in combs:
            compare(ctx, g, n, edges, senders, values[dt], comb, direction=direction, relations=rels,
                    per_relation=per, mask=mask, include_self=self_)


def test_hubs_and_many_messages(ctx):
    """A receiver whose senders hold thousands of messages, folded in order."""
    rng = np.random.default_rng(42)
    n = 3000
    edges = sorted({(0, int(s), 7) for s in rng.integers(0, n, 2500)} |
                   {(0, int(a), int(b)) for a, b in rng.integers(0, n, (6000, 2))})
    g = graph_of(ctx, n, edges)
    senders = rng.integers(0, n, 20000)
    vals = rng.standard_normal(20000)
    for comb in ("sum", "collect", "mode", "max"):
        compare(ctx, g, n, edges, senders, vals, comb)


def test_long_segments_fold_in_order_and_raise(ctx):
    """Receivers with more than 512 messages (the one-lane ordered fold):
    f64 and i64 sums that overflow part-way raise what the oracle's left
    fold raises; finite ones, and min / max, equal it bit for bit."""
    rng = np.random.default_rng(5)
    n = 2000
    edges = sorted({(0, s, 3) for s in range(n) if s != 3} |
                   {(0, int(s), 9) for s in rng.integers(0, n, 900)} |
                   {(0, int(a), int(b)) for a, b in rng.integers(0, n, (3000, 2))})
    g = graph_of(ctx, n, edges)
    senders = np.arange(n)
    print(f"long-segment gather: n={n} edges={len(edges)}", flush=True)
    finite = rng.standard_normal(n) * 1e3
    for comb in ("sum", "min", "max"):
        compare(ctx, g, n, edges, senders, finite, comb)
    big = np.where(rng.random(n) < 0.5, 1e306, -1e305)
    compare(ctx, g, n, edges, senders, big, "sum")
    ints = rng.integers(2**58, 2**59, n).astype(np.int64)
    compare(ctx, g, n, edges, senders, ints, "sum")
    compare(ctx, g, n, edges, senders, -ints, "sum")


def test_byte_messages_known_answers(ctx):
    """compute/gather/tests.rs:173-202: byte messages collect, concatenate,
    take the first and count; a folding combiner is refused."""
    g = graph_of(ctx, 4, EDGES)
    sent = [(2, b"c"), (0, b"a"), (2, b"d"), (1, b""), (3, b"z")]
    s = [a for a, _ in sent]
    m = [b for _, b in sent]
    r, v = W.gather(ctx, g, s, m, "collect")
    assert list(r) == [0, 1, 2] and v == [[b"c", b"d"], [b"a"], [b"a", b""]]
    r, v = W.gather(ctx, g, s, m, "concat")
    assert list(r) == [0, 1, 2] and v == [b"cd", b"a", b"a"]
    assert W.gather(ctx, g, s, m, "first")[1] == [b"c", b"a", b"a"]
    assert list(W.gather(ctx, g, s, m, "count")[1]) == [2, 1, 2]
    for comb in ("sum", "min", "max", "bool_or", "bool_and", "mode"):
        with pytest.raises(W.WarrenError) as e:
            W.gather(ctx, g, s, m, comb)
        assert e.value.name == "combiner_type"


@pytest.mark.parametrize("seed", range(4))
def test_random_byte_gathers(ctx, seed):
    """Random graphs, relations, directions, masks and include_self with
    byte messages of 0 to 40 bytes (some of 300), against the oracle."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 50))
    R = int(rng.integers(1, 4))
    edges = sorted({(int(rng.integers(0, R)), int(rng.integers(0, n)), int(rng.integers(0, n)))
                    for _ in range(int(rng.integers(1, 4 * n)))})
    g = graph_of(ctx, n, edges)
    k = int(rng.integers(1, 3 * n))
    senders = rng.integers(0, n, k)
    msgs = [bytes(rng.integers(0, 256, int(rng.choice([0, 1, 3, 7, 40, 300]))).astype(np.uint8))
            for _ in range(k)]
    nrel = max(R, 1)
    print(f"byte gather seed={seed} n={n} edges={len(edges)} messages={k}", flush=True)
    for trial in range(6):
        direction = ["in", "out", "both"][trial % 3]
        rels = None if (trial % 2 or nrel == 1) else sorted(set(rng.integers(0, nrel, 2).tolist()))
        per = bool(trial % 3 == 1) and nrel > 1
        mask = (rng.random(n) < 0.7).astype(np.float64) if trial % 4 == 3 else None
        for comb in ("collect", "count", "first", "concat", "sum"):
            compare(ctx, g, n, edges, senders, msgs, comb, direction=direction, relations=rels,
                    per_relation=per, mask=mask, include_self=bool(trial % 2))


def test_byte_messages_hub(ctx):
    """A hub receiving 2,500 senders' strings (about 1 MB concatenated)."""
    rng = np.random.default_rng(9)
    n = 3000
    edges = sorted({(0, int(s), 7) for s in rng.integers(0, n, 2500)})
    g = graph_of(ctx, n, edges)
    senders = np.arange(n)
    msgs = [bytes(rng.integers(0, 256, int(rng.integers(0, 800))).astype(np.uint8)) for _ in range(n)]
    for comb in ("concat", "collect", "first"):
        compare(ctx, g, n, edges, senders, msgs, comb)
