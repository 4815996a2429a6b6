"""Windows at the edges of the device path, against the CPU oracle:

- layers longer than the main pass allows (-w 14000 and -w 20000 windows):
  they run in a pass of their own with wide DP rows and a shallow LDS ring;
- more than 8,190 layers in one window (deep amplicon-like coverage);
- a column with IUPAC codes (15 distinct bytes): the default kernel's
  aligned sets (9 bytes per column) overflow and the window reruns in the
  wide kernel (17 bytes per column);
- a column with more distinct bytes than even the wide kernel holds: the
  window is returned as RH_WINDOW_REJECTED (backbone, for the caller's CPU
  path) while every other window of the batch is polished as usual.
"""
import sys

import numpy as np
import pytest

from racon_amd import _lib
from racon_amd._lib import Layer, Window
from synth import make_window, make_windows, mutate, oracle_args, random_quality, random_seq

sys.path.insert(0, _lib.PKG_DIR + "/../oracle")
import oracle_py  # noqa: E402

pytestmark = pytest.mark.gpu


def run(windows, scores=(3, -5, -4)):
    b = _lib.PoaBatch(*scores, trim=True)
    b.upload(windows)
    b.run(_lib.RH_POA_BATCHED)
    out, status = b.download()
    st = b.stats()
    b.close()
    return out, status, st


def check_against_oracle(windows, out, status, scores=(3, -5, -4), skip=()):
    for i, w in enumerate(windows):
        if i in skip:
            continue
        cons, ost, _, _ = oracle_py.window_consensus(*oracle_args(w), True, *scores)
        print(f"window {i}: backbone {len(w.backbone)}, {len(w.layers)} layers: gpu "
              f"{len(out[i])} bp status {status[i]} / oracle {len(cons)} bp status {ost}",
              flush=True)
        assert status[i] == ost, f"window {i} status"
        assert out[i] == cons, f"window {i} consensus"


def test_long_layer_windows_run_in_their_own_pass():
    rng = np.random.default_rng(61)
    wins = make_windows(62, 6, length=500, depth=12)
    # -w 14000 and -w 20000 windows: two full-span layers and two partial ones
    for length in (14000, 20000):
        w = make_window(rng, length=length, depth=4, full_fraction=0.5)
        wins.insert(3, w)
    out, status, st = run(wins)
    assert st.rejected == 0
    check_against_oracle(wins, out, status)


def test_more_than_8190_layers():
    rng = np.random.default_rng(63)
    length = 400
    backbone = random_seq(rng, length)
    truth = mutate(rng, backbone, 0.03)
    layers = []
    for _ in range(8300):
        begin = int(rng.integers(0, length - 40))
        end = begin + int(rng.integers(20, 40))
        seq = mutate(rng, truth[begin:end + 1], 0.08)
        layers.append(Layer(seq, random_quality(rng, len(seq)), begin, end))
    wins = [Window(backbone, b"!" * length, layers, True)] + make_windows(64, 3, length=300,
                                                                             depth=10)
    out, status, st = run(wins)
    assert st.rejected == 0
    check_against_oracle(wins, out, status)


def column_window(rng, length, codes):
    """A window whose layers are the backbone with one different byte each
    at column 60: the column's aligned set needs one member per byte."""
    backbone = random_seq(rng, length)
    layers = []
    for c in codes:
        seq = bytearray(backbone)
        seq[60] = c if c != backbone[60] else ord("X")
        layers.append(Layer(bytes(seq), None, 0, length - 1))
    return Window(backbone, b"!" * length, layers, True)


def test_iupac_column_window_reruns_in_the_wide_kernel():
    rng = np.random.default_rng(67)
    # A, C, G, T with N and the ten other IUPAC codes: 15 bytes in one column
    wide = column_window(rng, 120, b"ACGTNRYKMSWBDHV")
    wins = make_windows(68, 4, length=300, depth=10)
    wins.insert(2, wide)
    out, status, st = run(wins)
    print(f"retried {st.retried}, rejected {st.rejected}", flush=True)
    assert st.rejected == 0 and st.retried >= 1
    check_against_oracle(wins, out, status)


def test_wider_column_window_is_rejected_not_fatal():
    rng = np.random.default_rng(65)
    length = 120
    # twenty distinct bytes in one column: more than the wide kernel holds
    wide = column_window(rng, length, b"ACGTNRYKMSWBDHVEFIJL")
    wins = make_windows(66, 4, length=300, depth=10)
    wins.insert(2, wide)
    out, status, st = run(wins)
    assert status[2] == _lib.RH_WINDOW_REJECTED
    assert out[2] == wide.backbone
    assert st.rejected == 1
    check_against_oracle(wins, out, status, skip={2})
