"""Overlap alignment on the GPU vs the CPU oracle (edlib NW path)."""
import sys

import numpy as np
import pytest

from racon_amd import _lib
from synth import make_pair, mutate, random_seq

sys.path.insert(0, _lib.PKG_DIR + "/../oracle")
import oracle_py  # noqa: E402

pytestmark = pytest.mark.gpu


def check(pairs):
    b = _lib.AlignBatch()
    b.upload(pairs)
    b.run()
    ops, dist = b.download()
    b.close()
    for i, (q, t) in enumerate(pairs):
        d, oops, st = oracle_py.edlib_align(q, t)
        print(f"pair {i}: |q|={len(q)} |t|={len(t)} gpu d={dist[i]} oracle d={d}", flush=True)
        assert st == 0
        assert dist[i] == d, f"pair {i} distance"
        assert ops[i] == oops, f"pair {i} path differs"


def test_tiny_pairs():
    rng = np.random.default_rng(1)
    pairs = [(b"A", b"A"), (b"A", b"C"), (b"AC", b"A"), (b"ACGT", b"TGCA"), (b"", b"ACG"),
             (b"ACG", b"")]
    for n in (1, 3, 10, 63, 64, 65, 127, 200):
        pairs.append(make_pair(rng, n))
    check(pairs)


def test_leaf_sized_pairs():
    rng = np.random.default_rng(2)
    check([make_pair(rng, int(n)) for n in rng.integers(300, 2000, size=12)])


def test_hirschberg_pairs():
    rng = np.random.default_rng(3)
    check([make_pair(rng, int(n)) for n in (3000, 5000, 8000)])


def test_multi_strip_query():
    # queries longer than 64 blocks (4096 rows) use several strips
    rng = np.random.default_rng(4)
    t = random_seq(rng, 9000)
    q = mutate(rng, t, 0.12)
    check([(q, t), (t[:4200], q[:4100])])


def test_unequal_and_other_bytes():
    rng = np.random.default_rng(5)
    t = random_seq(rng, 1500)
    q = mutate(rng, t, 0.15)
    q2 = q[:700] + b"NNNN" + q[700:]
    check([(q[:900], t), (q2, t), (q2.replace(b"A", b"R", 3), t)])


def _plant(seq: bytes, at, run: bytes) -> bytes:
    """`seq` with `run` written over it at each position of `at` (clipped)."""
    b = bytearray(seq)
    for p in at:
        for k, c in enumerate(run):
            if 0 <= p + k < len(b):
                b[p + k] = c
    return bytes(b)


@pytest.mark.parametrize("length", [600, 3000, 9000])
def test_target_side_other_bytes(length):
    # N runs and IUPAC codes in the target (contigs with gaps), at 64-column
    # chunk boundaries and in the first and last chunk; lengths cover leaf
    # tracebacks, the Hirschberg split and multi-strip (> 4096-row) queries
    rng = np.random.default_rng(100 + length)
    t = random_seq(rng, length)
    q = mutate(rng, t, 0.12)
    edges = [0, 1, 62, 63, 64, 65, 126, 127, 128, 129, length // 2, length - 65, length - 64,
             length - 2, length - 1]
    pairs = [
        (q, _plant(t, edges, b"N")),
        (q, _plant(t, edges[::2], b"NNNNN")),
        (q, _plant(t, edges, b"R")),
        (_plant(q, edges[1::3], b"Y"), _plant(t, edges, b"RYKMSWN")),
        (_plant(q, [length // 3], b"N" * 40), _plant(t, [length // 3 + 5], b"N" * 40)),
        (q, t.lower() if length == 600 else _plant(t, [63, 127], b"acgt")),
    ]
    check(pairs)


@pytest.mark.parametrize("band", [0, 1, 37, 100000])
def test_band_guess_does_not_change_the_path(band):
    # the first band guess of a pair doubles until the banded distance fits
    # it (exact from then on): a guess of 1 forces many doublings, a huge one
    # computes the whole matrix; every path must be the oracle's
    rng = np.random.default_rng(200 + band)
    pairs = [make_pair(rng, int(n), error=0.12) for n in (90, 1500, 5000, 9000)]
    t = random_seq(rng, 6000)
    pairs.append((mutate(rng, t[:4500], 0.2), t))          # length difference 1500+
    pairs.append((random_seq(rng, 2500), random_seq(rng, 3000)))  # unrelated: d ~ 0.55 L
    pairs.append((mutate(rng, t, 0.3), t))
    b = _lib.AlignBatch()
    b.set_band(band)
    b.upload(pairs)
    b.run()
    ops, dist = b.download()
    st = b.stats()
    b.close()
    assert st.rejected == 0
    full = sum(((len(q) + 63) // 64 + 63) // 64 * (len(t) + 63) for q, t in pairs)
    print(f"band {band}: {st.steps} wave steps ({full} per unbanded pass)", flush=True)
    for i, (q, t) in enumerate(pairs):
        d, oops, _ = oracle_py.edlib_align(q, t)
        assert dist[i] == d, f"pair {i} distance"
        assert ops[i] == oops, f"pair {i} path"
