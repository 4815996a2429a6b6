"""Breaking points (racon src/overlap.cpp:226-292): the oracle's walk pinned by
closed forms, the C ABI's room bound against the walk, and (GPU) the device
walk against the oracle walk over the same device edit operations."""
import random

import pytest

from oracle import oracle_py
from racon_amd import _lib


def _windows(t_begin, t_end, w):
    return len([i for i in range(0, t_end, w) if i > t_begin]) + 1


def test_walk_all_matches_closed_form():
    # n matches from target t_begin, query q_first: window k holds targets
    # [max(t_begin, k*W), min(t_end, (k+1)*W)) and the query moves in step
    for (t_begin, n, q_first, w) in [(0, 1000, 0, 500), (250, 1000, 7, 500), (500, 17, 3, 500),
                                     (999, 2, 0, 1000), (0, 64, 10, 8)]:
        t_end = t_begin + n
        pts = oracle_py.breaking_points(bytes(n), t_begin, t_end, q_first, w)
        want = []
        lo = t_begin
        while lo < t_end:
            hi = min(t_end, (lo // w + 1) * w)
            want.append((lo, q_first + lo - t_begin))
            want.append((hi, q_first + hi - t_begin))
            lo = hi
        assert pts == want


def test_walk_gaps_hand_case():
    # t_begin 0, W 4: M M D D | D M I M | then a window of deletions only
    ops = bytes([0, 0, 2, 2, 2, 0, 1, 0, 2, 2, 2, 2])
    pts = oracle_py.breaking_points(ops, 0, 12, 0, 4)
    # window 0 (t 0..3): matches at t0/q0, t1/q1 -> first (0,0), last (2,2)
    # window 1 (t 4..7): match t5/q2, insertion q3, match t6/q4 -> (5,2), (7,5)
    # window 2 (t 8..11): no match -> nothing
    assert pts == [(0, 0), (2, 2), (5, 2), (7, 5)]


def test_bound_matches_window_count():
    lib = _lib.load_library()
    rng = random.Random(7)
    spans = []
    for _ in range(300):
        w = rng.choice([1, 7, 64, 500, 1000])
        t_begin = rng.randrange(0, 5000)
        t_end = t_begin + rng.randrange(1, 5000)
        spans.append((w, t_begin, t_end))
    for w, tb, te in spans:
        arr = (_lib.RhBreakSpan * 1)()
        arr[0].t_begin, arr[0].t_end, arr[0].q_first = tb, te, 0
        assert lib.rh_breaking_points_bound(w, arr, 1) == 2 * _windows(tb, te, w)


def _mutate(rng, seq, rate):
    out = bytearray()
    for c in seq:
        r = rng.random()
        if r < rate / 3:
            continue
        if r < 2 * rate / 3:
            out.append(rng.choice(b"ACGT"))
        elif r < rate:
            out.append(c)
            out.append(rng.choice(b"ACGT"))
        else:
            out.append(c)
    return bytes(out)


@pytest.mark.gpu
def test_device_breaking_points_match_walk():
    rng = random.Random(11)
    pairs, spans = [], []
    for k in range(96):
        t = bytes(rng.choice(b"ACGT") for _ in range(rng.randrange(1, 3000)))
        q = _mutate(rng, t, 0.15) or b"A"
        pairs.append((q, t))
        t_begin = rng.choice([0, 1, 499, 500, 501, rng.randrange(0, 20000)])
        spans.append((t_begin, t_begin + len(t), rng.randrange(0, 5000)))
    b = _lib.AlignBatch()
    b.upload(pairs)
    b.run()
    ops, _ = b.download()
    for w in (500, 1000, 37):
        got = b.breaking_points(w, spans)
        for i, (tb, te, qf) in enumerate(spans):
            want = oracle_py.breaking_points(ops[i], tb, te, qf, w)
            assert got[i] == want, (i, w)
    b.close()
