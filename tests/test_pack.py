"""Bulk window packing (the rh_window / rh_layer arrays handed to
rh_poa_upload) against the fields of the Python windows, read back through
the pointers as the library reads them."""
import ctypes

from racon_amd._lib import Layer, RhLayer, RhWindow, Window, pack_windows
from synth import make_windows


def read_back(arr, n):
    out = []
    wins = ctypes.cast(arr.ctypes.data, ctypes.POINTER(RhWindow))
    for i in range(n):
        w = wins[i]
        bb = ctypes.string_at(w.backbone, w.backbone_length)
        bq = (ctypes.string_at(w.backbone_quality, w.backbone_length)
              if w.backbone_quality else None)
        layers = []
        lp = ctypes.cast(w.layers, ctypes.POINTER(RhLayer))
        for k in range(w.n_layers):
            l = lp[k]
            seq = ctypes.string_at(l.sequence, l.length)
            q = ctypes.string_at(l.quality, l.length) if l.quality else None
            layers.append(Layer(seq, q, l.begin, l.end))
        out.append(Window(bb, bq, layers, bool(w.tgs)))
    return out


def test_pack_round_trip():
    wins = make_windows(31, 5, length=120, depth=7)
    wins += make_windows(32, 3, length=80, depth=5, quality=False, tgs=False)
    wins += make_windows(33, 2, length=60, depth=0)  # windows without layers
    arr, keep = pack_windows(wins)
    assert read_back(arr, len(wins)) == wins


def test_pack_empty_and_mixed_quality():
    arr, keep = pack_windows([])
    assert len(arr) == 1
    w = Window(b"ACGT", b"!!!!", [Layer(b"AC", None, 0, 1), Layer(b"GT", b"++", 2, 3)], True)
    arr, keep = pack_windows([w, w])
    assert read_back(arr, 2) == [w, w]


def test_pack_joined_buffers(monkeypatch):
    # the interpreter-independent path: parts joined into one buffer
    from racon_amd import _lib
    monkeypatch.setattr(_lib, "_DIRECT", False)
    wins = make_windows(34, 4, length=90, depth=6)
    wins += make_windows(35, 2, length=70, depth=3, quality=False)
    arr, keep = pack_windows(wins)
    assert read_back(arr, len(wins)) == wins


def test_pack_bytearray_and_memoryview_parts():
    # non-bytes parts are converted before their addresses are taken: the
    # library must read the sequence, not the object header
    seq = b"ACGTACGTAC"
    qual = b"++++++++++"
    w = Window(bytearray(b"ACGTACGTACGT"), memoryview(b"!!!!!!!!!!!!"),
               [Layer(bytearray(seq), memoryview(qual), 0, 9),
                Layer(memoryview(seq), bytearray(qual), 1, 10)], True)
    arr, keep = pack_windows([w])
    back = read_back(arr, 1)[0]
    assert back.backbone == b"ACGTACGTACGT"
    assert back.backbone_quality == b"!!!!!!!!!!!!"
    assert [(l.sequence, l.quality) for l in back.layers] == [(seq, qual), (seq, qual)]


def test_pack_refuses_str_parts():
    import pytest
    w = Window(b"ACGT", b"!!!!", [Layer("ACGT", None, 0, 3)], True)
    with pytest.raises(TypeError):
        pack_windows([w])
