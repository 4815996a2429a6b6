"""Pins the CPU restatement of edlib's NW path (the oracle for the aligner).

- distance: textbook Levenshtein DP (brute force) on seeded random pairs;
- path: consumes exactly |q| and |t|, costs exactly the distance, matches
  only equal bytes;
- preference: hand-derived known answers for the up > left > diagonal
  traceback order (a swapped order fails them);
- banded (edlib-like) == exact recursion on long pairs, which is what lets
  the device kernel compute exact scores without the band.
"""
import numpy as np
import pytest

import oracle_py
from synth import make_pair, mutate, random_seq


def levenshtein(q: bytes, t: bytes) -> int:
    prev = list(range(len(t) + 1))
    for i in range(1, len(q) + 1):
        cur = [i] + [0] * len(t)
        for j in range(1, len(t) + 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (q[i - 1] != t[j - 1]))
        prev = cur
    return prev[-1]


def check_path(q: bytes, t: bytes, d: int, ops: bytes) -> None:
    i = j = cost = 0
    for op in ops:
        if op in (0, 3):
            assert i < len(q) and j < len(t)
            assert (op == 0) == (q[i] == t[j])
            cost += op == 3
            i += 1
            j += 1
        elif op == 1:
            i += 1
            cost += 1
        elif op == 2:
            j += 1
            cost += 1
        else:
            raise AssertionError(f"bad op {op}")
    assert (i, j) == (len(q), len(t))
    assert cost == d


def test_distance_matches_brute_force():
    rng = np.random.default_rng(100)
    for _ in range(150):
        n = int(rng.integers(0, 40))
        t = random_seq(rng, n)
        q = mutate(rng, t, float(rng.uniform(0, 0.4))) if n else random_seq(rng, 3)
        d, ops, st = oracle_py.edlib_align(q, t)
        assert st == 0
        assert d == levenshtein(q, t)
        check_path(q, t, d, ops)


def test_empty_sides():
    assert oracle_py.edlib_align(b"", b"ACG")[:2] == (3, bytes([2, 2, 2]))
    assert oracle_py.edlib_align(b"ACG", b"")[:2] == (3, bytes([1, 1, 1]))


@pytest.mark.parametrize("q,t,expect", [
    # from the last cell: up (insertion) is preferred, then left (deletion),
    # then the diagonal
    (b"A", b"AA", bytes([0, 2])),        # 1M1D: left before diagonal at (0,1)
    (b"AA", b"A", bytes([0, 1])),        # 1M1I: up first at (1,0)
    (b"AC", b"CA", bytes([2, 0, 1])),    # D M I
    (b"ACGT", b"ACGT", bytes([0, 0, 0, 0])),
    (b"ACGT", b"AGGT", bytes([0, 3, 0, 0])),
])
def test_known_answers(q, t, expect):
    d, ops, _ = oracle_py.edlib_align(q, t)
    assert ops == expect
    assert d == levenshtein(q, t)


def test_banded_equals_exact_long_pairs():
    rng = np.random.default_rng(101)
    for n in (100, 700, 1500, 3000, 6000):
        q, t = make_pair(rng, n, error=0.12)
        a = oracle_py.edlib_align(q, t)
        b = oracle_py.edlib_align(q, t, exact=True)
        assert a == b
        check_path(q, t, a[0], a[1])


def test_banded_equals_exact_unequal_lengths():
    rng = np.random.default_rng(102)
    t = random_seq(rng, 2500)
    for q in (mutate(rng, t[:1800], 0.1), mutate(rng, t, 0.25), t[300:] + t[:50]):
        assert oracle_py.edlib_align(q, t) == oracle_py.edlib_align(q, t, exact=True)


def test_other_bytes_match_brute_force():
    # N runs, IUPAC codes and lower case on both sides: edlib compares bytes
    # as symbols of the pair's own alphabet, so the distance is Levenshtein's
    rng = np.random.default_rng(103)
    extra = np.frombuffer(b"NRYKMSWacgt", dtype=np.uint8)
    for _ in range(120):
        n = int(rng.integers(1, 45))
        t = bytearray(random_seq(rng, n))
        q = bytearray(mutate(rng, bytes(t), float(rng.uniform(0, 0.3))))
        for s in (q, t):
            for _ in range(int(rng.integers(0, 4))):
                s[int(rng.integers(0, len(s)))] = int(extra[rng.integers(0, len(extra))])
        q, t = bytes(q), bytes(t)
        d, ops, st = oracle_py.edlib_align(q, t)
        assert st == 0
        assert d == levenshtein(q, t)
        check_path(q, t, d, ops)
        assert (d, ops) == oracle_py.edlib_align(q, t, exact=True)[:2]
