"""Seeded synthetic windows and read pairs for the parity tests.

Layers are noisy copies of backbone spans (ONT-like substitutions,
insertions and deletions), so the POA graphs grow the way real windows do.
"""
from __future__ import annotations

import numpy as np

from racon_amd._lib import Layer, Window

BASES = np.frombuffer(b"ACGT", dtype=np.uint8)


def random_seq(rng: np.random.Generator, n: int) -> bytes:
    return BASES[rng.integers(0, 4, size=n)].tobytes()


def mutate(rng: np.random.Generator, seq: bytes, rate: float,
           split=(0.4, 0.2, 0.4)) -> bytes:
    """Applies substitutions/insertions/deletions at total `rate`."""
    out = bytearray()
    p_sub, p_ins, p_del = (rate * s for s in split)
    for c in seq:
        r = rng.random()
        if r < p_sub:
            choices = [b for b in b"ACGT" if b != c]
            out.append(choices[rng.integers(0, 3)])
        elif r < p_sub + p_ins:
            out.append(c)
            out.append(int(BASES[rng.integers(0, 4)]))
        elif r < p_sub + p_ins + p_del:
            continue
        else:
            out.append(c)
    if not out:
        out.append(int(BASES[rng.integers(0, 4)]))
    return bytes(out)


def random_quality(rng: np.random.Generator, n: int, lo: int = 8, hi: int = 30) -> bytes:
    return (rng.integers(lo, hi + 1, size=n) + 33).astype(np.uint8).tobytes()


def make_window(rng: np.random.Generator, length: int = 500, depth: int = 30,
                error: float = 0.12, quality: bool = True, backbone_quality: bool = False,
                full_fraction: float = 0.6, tgs: bool = True) -> Window:
    backbone = random_seq(rng, length)
    truth = mutate(rng, backbone, 0.05)
    bq = random_quality(rng, length) if backbone_quality else b"!" * length
    layers = []
    for _ in range(depth):
        if rng.random() < full_fraction:
            begin, end = 0, length - 1
        else:
            begin = int(rng.integers(0, length - 20))
            end = int(rng.integers(begin + 10, length))
        # layer bytes: a noisy copy of the corresponding truth span
        t0 = min(len(truth) - 1, begin)
        t1 = min(len(truth), max(t0 + 1, end + 1))
        seq = mutate(rng, truth[t0:t1], error)
        q = random_quality(rng, len(seq)) if quality else None
        layers.append(Layer(seq, q, begin, end))
    return Window(backbone, bq, layers, tgs)


def make_windows(seed: int, n: int, **kw) -> list:
    rng = np.random.default_rng(seed)
    return [make_window(rng, **kw) for _ in range(n)]


def make_pair(rng: np.random.Generator, length: int, error: float = 0.12):
    t = random_seq(rng, length)
    q = mutate(rng, t, error)
    return q, t


def oracle_args(w: Window):
    return (w.backbone, w.backbone_quality,
            [(l.sequence, l.quality, l.begin, l.end) for l in w.layers], w.tgs)
