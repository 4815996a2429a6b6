"""Readers of the binary dumps written by `racon --dump-windows/--dump-overlaps`.

Window dump: "RHW1", u64 n, then per window u32 length, u8 tgs, 3 pad bytes,
backbone, backbone quality, u32 layers, and per layer (insertion order)
u32 length, u32 begin, u32 end, u8 has_quality, bytes, [quality bytes].
Overlap dump: "RHA1", u64 n, then per pair u32 |q|, u32 |t|, q, t.
"""
from __future__ import annotations

import struct
from typing import List, Tuple

from ._lib import Layer, Window


def read_windows(path: str) -> List[Window]:
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:4] != b"RHW1":
        raise ValueError(f"{path}: not a window dump")
    (n,) = struct.unpack_from("<Q", buf, 4)
    pos = 12
    out = []
    for _ in range(n):
        length, tgs = struct.unpack_from("<IB", buf, pos)
        pos += 8
        backbone = buf[pos:pos + length]
        pos += length
        quality = buf[pos:pos + length]
        pos += length
        (n_layers,) = struct.unpack_from("<I", buf, pos)
        pos += 4
        layers = []
        for _ in range(n_layers):
            ll, b, e, hq = struct.unpack_from("<IIIB", buf, pos)
            pos += 13
            seq = buf[pos:pos + ll]
            pos += ll
            q = None
            if hq:
                q = buf[pos:pos + ll]
                pos += ll
            layers.append(Layer(seq, q, b, e))
        out.append(Window(backbone, quality, layers, bool(tgs)))
    return out


def read_overlaps(path: str) -> List[Tuple[bytes, bytes]]:
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:4] != b"RHA1":
        raise ValueError(f"{path}: not an overlap dump")
    (n,) = struct.unpack_from("<Q", buf, 4)
    pos = 12
    out = []
    for _ in range(n):
        ql, tl = struct.unpack_from("<II", buf, pos)
        pos += 8
        q = buf[pos:pos + ql]
        pos += ql
        t = buf[pos:pos + tl]
        pos += tl
        out.append((q, t))
    return out
