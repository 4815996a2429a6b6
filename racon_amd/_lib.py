"""ctypes binding of libracon_hip.so (include/racon_hip.h).

The product is the C ABI; this module only lets Python drive it (tests,
benchmark, smoke run). Loading fails loudly when the library has not been
built: there is no fallback implementation behind it.
"""
from __future__ import annotations

import ctypes
import itertools
import operator
import os
import sys
from dataclasses import dataclass
from typing import List, Optional, Sequence

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# RACON_HIP_LIBRARY names another build of the same library (e.g. one
# compiled with other kernel occupancy settings, for comparisons)
LIB_PATH = os.environ.get("RACON_HIP_LIBRARY") or os.path.join(PKG_DIR, "lib", "libracon_hip.so")
RACON_BIN = os.path.join(PKG_DIR, "bin", "racon")

RH_OK = 0
RH_POA_BATCHED = 0
RH_POA_PLAIN = 1
RH_POA_STATIC = 2
RH_WINDOW_UNPOLISHED = 0
RH_WINDOW_POLISHED = 1
RH_WINDOW_POLISHED_UNTRIMMED = 2
RH_WINDOW_REJECTED = 3


class RhLayer(ctypes.Structure):
    _fields_ = [
        ("sequence", ctypes.c_char_p),
        ("quality", ctypes.c_char_p),
        ("length", ctypes.c_uint32),
        ("begin", ctypes.c_uint32),
        ("end", ctypes.c_uint32),
    ]


class RhWindow(ctypes.Structure):
    _fields_ = [
        ("backbone", ctypes.c_char_p),
        ("backbone_quality", ctypes.c_char_p),
        ("backbone_length", ctypes.c_uint32),
        ("n_layers", ctypes.c_uint32),
        ("layers", ctypes.POINTER(RhLayer)),
        ("tgs", ctypes.c_uint32),
    ]


class RhPoaStats(ctypes.Structure):
    _fields_ = [
        ("pack_ms", ctypes.c_double),
        ("h2d_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("d2h_ms", ctypes.c_double),
        ("cells", ctypes.c_uint64),
        ("windows", ctypes.c_uint64),
        ("gpu_windows", ctypes.c_uint64),
        ("retried", ctypes.c_uint64),
        ("input_bytes", ctypes.c_uint64),
        ("slots", ctypes.c_uint32),
        ("launches", ctypes.c_uint32),
        ("rejected", ctypes.c_uint64),
        ("band_reruns", ctypes.c_uint64),
    ]


class RhPair(ctypes.Structure):
    _fields_ = [
        ("query", ctypes.c_char_p),
        ("query_length", ctypes.c_uint32),
        ("target", ctypes.c_char_p),
        ("target_length", ctypes.c_uint32),
    ]


class RhBreakSpan(ctypes.Structure):
    _fields_ = [
        ("t_begin", ctypes.c_uint32),
        ("t_end", ctypes.c_uint32),
        ("q_first", ctypes.c_uint32),
    ]


class RhLayerSource(ctypes.Structure):
    _fields_ = [
        ("overlap", ctypes.c_uint32),
        ("query_length", ctypes.c_uint32),
        ("first_window", ctypes.c_uint64),
        ("quality", ctypes.c_uint64),
        ("mirrored", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class RhLayerRecord(ctypes.Structure):
    _fields_ = [
        ("window", ctypes.c_uint32),
        ("overlap", ctypes.c_uint32),
        ("q_begin", ctypes.c_uint32),
        ("q_end", ctypes.c_uint32),
        ("begin", ctypes.c_uint32),
        ("end", ctypes.c_uint32),
    ]


RH_NO_QUALITY = 0xFFFFFFFFFFFFFFFF


class RhAlignStats(ctypes.Structure):
    _fields_ = [
        ("pack_ms", ctypes.c_double),
        ("h2d_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("d2h_ms", ctypes.c_double),
        ("pairs", ctypes.c_uint64),
        ("cells", ctypes.c_uint64),
        ("ops", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
        ("steps", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
    ]


class RhPoaTuning(ctypes.Structure):
    _fields_ = [
        ("ring_rows", ctypes.c_uint32 * 2),
        ("kernel_ms", ctypes.c_double * 2),
        ("default_ms", ctypes.c_double * 2),
        ("candidates", ctypes.c_uint32),
    ]


RH_TIMEOUT = 1


class RhRequestInfo(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_uint32),
        ("windows", ctypes.c_uint32),
        ("device_bytes", ctypes.c_uint64),
        ("output_bound", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
        ("shared_launches", ctypes.c_uint32),
        ("arrival_s", ctypes.c_double),
        ("admitted_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
    ]


class RhServerStats(ctypes.Structure):
    _fields_ = [
        ("pool_bytes", ctypes.c_uint64),
        ("pool_in_use", ctypes.c_uint64),
        ("window_capacity", ctypes.c_uint64),
        ("scratch_bytes", ctypes.c_uint64),
        ("submitted", ctypes.c_uint64),
        ("admitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("shared_launches", ctypes.c_uint64),
        ("windows_launched", ctypes.c_uint64),
        ("retried", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
    ]


# Every entry point of include/racon_hip.h with its ctypes signature.
_SIGNATURES = {
    "rh_last_error": (ctypes.c_char_p, []),
    "rh_version": (ctypes.c_int32, []),
    "rh_device_count": (ctypes.c_int32, []),
    "rh_poa_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int8, ctypes.c_int8,
                                       ctypes.c_int8, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "rh_poa_destroy": (None, [ctypes.c_void_p]),
    "rh_poa_upload": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32,
                                       ctypes.POINTER(RhWindow)]),
    "rh_poa_run": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "rh_poa_submit": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "rh_poa_wait": (ctypes.c_int32, [ctypes.c_void_p]),
    "rh_poa_capacity": (ctypes.c_int32, [ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "rh_poa_device_bytes": (ctypes.c_uint64, [ctypes.c_void_p]),
    "rh_poa_output_bound": (ctypes.c_uint64, [ctypes.c_void_p]),
    "rh_poa_download": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_uint8)]),
    "rh_poa_get_stats": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(RhPoaStats)]),
    "rh_poa_set_tuning": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_uint32]),
    "rh_poa_autotune": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(RhPoaTuning)]),
    "rh_poa_set_band": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32]),
    "rh_poa_phase_ticks": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "rh_poa_window_costs": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]),
    "rh_server_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int8, ctypes.c_int8,
                                          ctypes.c_int8, ctypes.c_int32, ctypes.c_uint64,
                                          ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "rh_server_warm": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32,
                                        ctypes.POINTER(RhWindow)]),
    "rh_server_submit": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32,
                                          ctypes.POINTER(RhWindow),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "rh_server_hold": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "rh_server_wait": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32]),
    "rh_server_request_info": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint64,
                                                ctypes.POINTER(RhRequestInfo)]),
    "rh_server_fetch": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_char_p,
                                         ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_uint8)]),
    "rh_server_get_stats": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(RhServerStats)]),
    "rh_server_destroy": (None, [ctypes.c_void_p]),
    "rh_align_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "rh_align_destroy": (None, [ctypes.c_void_p]),
    "rh_align_upload": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32,
                                         ctypes.POINTER(RhPair)]),
    "rh_align_run": (ctypes.c_int32, [ctypes.c_void_p]),
    "rh_align_set_band": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32]),
    "rh_align_output_bound": (ctypes.c_uint64, [ctypes.c_void_p]),
    "rh_align_download": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64,
                                           ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_int32)]),
    "rh_align_get_stats": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(RhAlignStats)]),
    "rh_breaking_points_bound": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.POINTER(RhBreakSpan),
                                                   ctypes.c_uint32]),
    "rh_align_breaking_points": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32,
                                                  ctypes.POINTER(RhBreakSpan),
                                                  ctypes.POINTER(ctypes.c_uint64),
                                                  ctypes.POINTER(ctypes.c_uint32),
                                                  ctypes.POINTER(ctypes.c_uint32),
                                                  ctypes.c_uint64]),
    "rh_layers_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_uint32, ctypes.c_double,
                                          ctypes.c_uint64, ctypes.c_uint64, ctypes.c_char_p,
                                          ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "rh_layers_add": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.POINTER(RhLayerSource), ctypes.c_uint32]),
    "rh_layers_finish": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "rh_layers_download": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(RhLayerRecord),
                                            ctypes.c_uint64]),
    "rh_layers_destroy": (None, [ctypes.c_void_p]),
    # device pointers are passed as integers (c_void_p)
    "rh_align_run_noalloc": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_uint32,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
}

# rh_dev_pair / rh_dev_align_result of include/racon_hip.h as numpy dtypes
DEV_PAIR_FIELDS = [("query_offset", "<u8"), ("target_offset", "<u8"), ("ops_offset", "<u8"),
                   ("query_length", "<u4"), ("target_length", "<u4")]
DEV_ALIGN_RESULT_FIELDS = [("n_ops", "<u4"), ("edit_distance", "<i4"), ("status", "<i4"),
                           ("reserved", "<u4")]
NOALLOC_SCRATCH_BYTES = 64 << 20

_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libracon_hip.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"libracon_hip.so not found at {path}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    for name, (restype, argtypes) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def _check(rc: int, what: str) -> None:
    if rc != RH_OK:
        msg = load_library().rh_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


@dataclass
class Layer:
    sequence: bytes
    quality: Optional[bytes]
    begin: int
    end: int


@dataclass
class Window:
    backbone: bytes
    backbone_quality: bytes
    layers: List[Layer]
    tgs: bool = True


def _address(blob: bytes) -> int:
    return ctypes.cast(ctypes.c_char_p(blob), ctypes.c_void_p).value or 0


# CPython keeps a bytes object's data at a fixed offset from the object
# (the address id() returns); checked once here, else the parts are joined
# into one buffer first.
_BYTES_DATA = sys.getsizeof(b"") - 1
_probe = b"ACGT"
_DIRECT = _address(_probe) == id(_probe) + _BYTES_DATA


def _as_bytes(p):
    """`p` as an immutable bytes object (None stays None). bytearray and
    memoryview are copied; anything else (str, lists) is refused, since its
    memory is not a byte string."""
    if p is None or type(p) is bytes:
        return p
    if isinstance(p, (bytearray, memoryview)):
        return bytes(p)
    raise TypeError(f"sequence and quality parts must be bytes, not {type(p).__name__}")


def _pointers(parts: Sequence[Optional[bytes]]):
    """Address of each part's bytes (0 for None) without copying them where
    the interpreter allows; returns (addresses, lengths, objects to keep).
    The id() shortcut is only taken for exact `bytes` objects: every other
    part is converted by _as_bytes first."""
    import numpy as np
    if not all(p is None or type(p) is bytes for p in parts):
        parts = [_as_bytes(p) for p in parts]
    n = len(parts)
    present = np.fromiter(map(operator.is_not, parts, itertools.repeat(None)), dtype=bool,
                          count=n)
    if _DIRECT:
        addr = np.fromiter(map(id, parts), dtype=np.uint64, count=n) + np.uint64(_BYTES_DATA)
        lens = np.zeros(n, dtype=np.uint64)
        lens[present] = np.fromiter(map(len, itertools.compress(parts, present)),
                                    dtype=np.uint64, count=int(present.sum()))
        keep = parts
    else:
        lens = np.fromiter((len(p) if p is not None else 0 for p in parts), dtype=np.uint64,
                           count=n)
        keep = b"".join(p for p in parts if p is not None) + b"\0"
        addr = np.zeros(n, dtype=np.uint64)
        if n > 1:
            np.cumsum(lens[:-1], out=addr[1:])
        addr += np.uint64(_address(keep))
    return np.where(present, addr, np.uint64(0)), lens, keep


def pack_windows(windows: Sequence[Window]):
    """The rh_window / rh_layer arrays of `windows` built in bulk (numpy
    records laid out as the C structs, pointers into a few joined buffers).
    Returns (window records, objects that must outlive the upload call)."""
    import numpy as np
    layer_dt = np.dtype([("sequence", "<u8"), ("quality", "<u8"), ("length", "<u4"),
                         ("begin", "<u4"), ("end", "<u4")], align=True)
    window_dt = np.dtype([("backbone", "<u8"), ("backbone_quality", "<u8"),
                          ("backbone_length", "<u4"), ("n_layers", "<u4"),
                          ("layers", "<u8"), ("tgs", "<u4")], align=True)
    assert layer_dt.itemsize == ctypes.sizeof(RhLayer)
    assert window_dt.itemsize == ctypes.sizeof(RhWindow)
    n = len(windows)
    flat = [l for w in windows for l in w.layers]
    layers = np.zeros(max(1, len(flat)), dtype=layer_dt)
    get_begin, get_end = operator.attrgetter("begin"), operator.attrgetter("end")
    s_ptr, s_len, s_blob = _pointers(list(map(operator.attrgetter("sequence"), flat)))
    q_ptr, _, q_blob = _pointers(list(map(operator.attrgetter("quality"), flat)))
    if flat:
        layers["sequence"] = s_ptr
        layers["quality"] = q_ptr
        layers["length"] = s_len
        layers["begin"] = np.fromiter(map(get_begin, flat), dtype=np.uint32, count=len(flat))
        layers["end"] = np.fromiter(map(get_end, flat), dtype=np.uint32, count=len(flat))
    arr = np.zeros(max(1, n), dtype=window_dt)
    b_ptr, b_len, b_blob = _pointers([w.backbone for w in windows])
    bq_ptr, _, bq_blob = _pointers([w.backbone_quality for w in windows])
    n_layers = np.fromiter((len(w.layers) for w in windows), dtype=np.uint64, count=n)
    first = np.zeros(n, dtype=np.uint64)
    if n > 1:
        np.cumsum(n_layers[:-1], out=first[1:])
    if n:
        arr["backbone"][:n] = b_ptr
        arr["backbone_quality"][:n] = bq_ptr
        arr["backbone_length"][:n] = b_len
        arr["n_layers"][:n] = n_layers
        arr["layers"][:n] = np.uint64(layers.ctypes.data) + first * np.uint64(layer_dt.itemsize)
        arr["tgs"][:n] = np.fromiter((1 if w.tgs else 0 for w in windows), dtype=np.uint32,
                                     count=n)
    return arr, (layers, s_blob, q_blob, b_blob, bq_blob)


class PoaBatch:
    """A device batch of windows (rh_poa_*)."""

    def __init__(self, match: int = 3, mismatch: int = -5, gap: int = -4,
                 trim: bool = True, device: int = 0):
        self.lib = load_library()
        self.handle = ctypes.c_void_p()
        _check(self.lib.rh_poa_create(device, match, mismatch, gap, int(trim),
                                      ctypes.byref(self.handle)), "rh_poa_create")
        self._keep = []
        self.n = 0

    def close(self) -> None:
        if self.handle:
            self.lib.rh_poa_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, slots: int = 0, rows: int = 0, ring_rows: int = 0) -> None:
        _check(self.lib.rh_poa_set_tuning(self.handle, slots, rows, ring_rows),
               "rh_poa_set_tuning")

    def set_band(self, band: int) -> None:
        """rh_poa_set_band: 256 for racon -b's banded approximation, 0 exact."""
        _check(self.lib.rh_poa_set_band(self.handle, band), "rh_poa_set_band")

    def autotune(self) -> dict:
        """rh_poa_autotune on the uploaded batch: the fastest LDS ring depth
        per layer-length bin, kept for this batch's later runs."""
        t = RhPoaTuning()
        _check(self.lib.rh_poa_autotune(self.handle, ctypes.byref(t)), "rh_poa_autotune")
        return {"ring_rows": list(t.ring_rows), "kernel_ms": list(t.kernel_ms),
                "default_ms": list(t.default_ms), "candidates": t.candidates}

    def upload(self, windows: Sequence[Window]) -> None:
        arr, keep = pack_windows(windows)
        self.upload_packed(arr, len(windows), keep)

    def upload_packed(self, arr, n: int, keep=None) -> None:
        """rh_poa_upload of an rh_window array made by pack_windows (the
        caller's window list, built once)."""
        _check(self.lib.rh_poa_upload(self.handle, n,
                                      ctypes.cast(arr.ctypes.data, ctypes.POINTER(RhWindow))),
               "rh_poa_upload")
        self._keep = (arr, keep)
        self.n = n

    def run(self, mode: int = RH_POA_BATCHED) -> None:
        _check(self.lib.rh_poa_run(self.handle, mode), "rh_poa_run")

    def submit(self, mode: int = RH_POA_BATCHED) -> None:
        """Queues the kernels and returns (rh_poa_submit)."""
        _check(self.lib.rh_poa_submit(self.handle, mode), "rh_poa_submit")

    def wait(self) -> None:
        """Blocks until a submitted run is done (rh_poa_wait)."""
        _check(self.lib.rh_poa_wait(self.handle), "rh_poa_wait")

    def device_bytes(self) -> int:
        return int(self.lib.rh_poa_device_bytes(self.handle))

    def output_buffers(self, capacity: int = 0):
        """Buffers for download_into: (out, capacity, offsets, lengths, status)."""
        cap = max(1, capacity or self.lib.rh_poa_output_bound(self.handle))
        return (ctypes.create_string_buffer(cap), cap, (ctypes.c_uint64 * max(1, self.n))(),
                (ctypes.c_uint32 * max(1, self.n))(), (ctypes.c_uint8 * max(1, self.n))())

    def download_into(self, bufs) -> None:
        """rh_poa_download into buffers from output_buffers (no Python
        objects per window)."""
        out, cap, offsets, lengths, status = bufs
        _check(self.lib.rh_poa_download(self.handle, out, cap, offsets, lengths, status),
               "rh_poa_download")

    @staticmethod
    def split(bufs, n: int):
        out, _, offsets, lengths, status = bufs
        raw = out.raw
        return ([raw[offsets[i]:offsets[i] + lengths[i]] for i in range(n)],
                [status[i] for i in range(n)])

    def download(self):
        bufs = self.output_buffers()
        self.download_into(bufs)
        return self.split(bufs, self.n)

    PHASES = ("include", "subsort", "rows", "column0", "dp", "traceback", "insert", "sort",
              "consensus", "setup")

    def phase_seconds(self) -> dict:
        """Device time per kernel phase of the last run, summed over windows."""
        ticks = (ctypes.c_uint64 * len(self.PHASES))()
        rate = ctypes.c_uint64()
        _check(self.lib.rh_poa_phase_ticks(self.handle, ticks, len(self.PHASES),
                                           ctypes.byref(rate)), "rh_poa_phase_ticks")
        return {k: ticks[i] / max(1, rate.value) for i, k in enumerate(self.PHASES)}

    def tick_rate(self) -> int:
        """Device constant-clock ticks per second."""
        ticks = (ctypes.c_uint64 * len(self.PHASES))()
        rate = ctypes.c_uint64()
        _check(self.lib.rh_poa_phase_ticks(self.handle, ticks, len(self.PHASES),
                                           ctypes.byref(rate)), "rh_poa_phase_ticks")
        return max(1, rate.value)

    def window_costs(self):
        """Per uploaded window of the last run: (DP cells, device ticks)."""
        cells = (ctypes.c_uint64 * self.n)()
        ticks = (ctypes.c_uint64 * self.n)()
        _check(self.lib.rh_poa_window_costs(self.handle, cells, ticks, self.n),
               "rh_poa_window_costs")
        return list(cells), list(ticks)

    def stats(self) -> RhPoaStats:
        st = RhPoaStats()
        _check(self.lib.rh_poa_get_stats(self.handle, ctypes.byref(st)), "rh_poa_get_stats")
        return st


class AlignBatch:
    """A device batch of pairwise NW alignments (rh_align_*)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.handle = ctypes.c_void_p()
        _check(self.lib.rh_align_create(device, ctypes.byref(self.handle)), "rh_align_create")
        self.n = 0
        self._keep = None

    def close(self) -> None:
        if self.handle:
            self.lib.rh_align_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, pairs: Sequence[tuple]) -> None:
        arr = (RhPair * max(1, len(pairs)))()
        for i, (q, t) in enumerate(pairs):
            arr[i].query = q
            arr[i].query_length = len(q)
            arr[i].target = t
            arr[i].target_length = len(t)
        _check(self.lib.rh_align_upload(self.handle, len(pairs), arr), "rh_align_upload")
        self._keep = (arr, list(pairs))
        self.n = len(pairs)

    def run(self) -> None:
        _check(self.lib.rh_align_run(self.handle), "rh_align_run")

    def set_band(self, band: int) -> None:
        """First band guess of every pair (0: automatic)."""
        _check(self.lib.rh_align_set_band(self.handle, band), "rh_align_set_band")

    def download(self):
        cap = self.lib.rh_align_output_bound(self.handle)
        ops = ctypes.create_string_buffer(max(1, cap))
        offsets = (ctypes.c_uint64 * max(1, self.n))()
        n_ops = (ctypes.c_uint32 * max(1, self.n))()
        dist = (ctypes.c_int32 * max(1, self.n))()
        _check(self.lib.rh_align_download(self.handle, ops, cap, offsets, n_ops, dist),
               "rh_align_download")
        raw = ops.raw
        return ([raw[offsets[i]:offsets[i] + n_ops[i]] for i in range(self.n)],
                [dist[i] for i in range(self.n)])

    def breaking_points(self, window_length: int, spans: Sequence[tuple]):
        """Device breaking points of the last run: spans[i] = (t_begin, t_end,
        q_first) of pair i; returns per pair a list of (target, query) points."""
        n = self.n
        arr = (RhBreakSpan * max(1, n))()
        for i, (tb, te, qf) in enumerate(spans):
            arr[i].t_begin, arr[i].t_end, arr[i].q_first = tb, te, qf
        cap = self.lib.rh_breaking_points_bound(window_length, arr, n)
        points = (ctypes.c_uint32 * max(2, 2 * cap))()
        offsets = (ctypes.c_uint64 * max(1, n))()
        counts = (ctypes.c_uint32 * max(1, n))()
        _check(self.lib.rh_align_breaking_points(self.handle, window_length, arr, offsets, counts,
                                                 points, cap), "rh_align_breaking_points")
        out = []
        for i in range(n):
            o = 2 * offsets[i]
            out.append([(points[o + 2 * k], points[o + 2 * k + 1]) for k in range(counts[i])])
        return out

    def stats(self) -> RhAlignStats:
        st = RhAlignStats()
        _check(self.lib.rh_align_get_stats(self.handle, ctypes.byref(st)),
               "rh_align_get_stats")
        return st
class LayerSet:
    """Window layers from device breaking points (rh_layers_*): racon's
    span and mean-quality filters and the grouping by window of
    src/polisher.cpp:403-457."""

    def __init__(self, window_length: int, quality_threshold: float, n_windows: int,
                 capacity: int, qualities: bytes = b"", device: int = 0):
        self.lib = load_library()
        self.handle = ctypes.c_void_p()
        self._qualities = bytes(qualities)
        _check(self.lib.rh_layers_create(device, window_length, quality_threshold, n_windows,
                                         capacity, self._qualities or None,
                                         len(self._qualities), ctypes.byref(self.handle)),
               "rh_layers_create")

    def close(self) -> None:
        if self.handle:
            self.lib.rh_layers_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, batch: "AlignBatch", sources: Sequence[tuple]) -> None:
        """sources[i] = (overlap, query_length, first_window, quality offset or
        None, mirrored) for pair i of the batch's last breaking points."""
        arr = (RhLayerSource * max(1, len(sources)))()
        for i, (ov, qlen, fw, qoff, mirrored) in enumerate(sources):
            arr[i].overlap, arr[i].query_length, arr[i].first_window = ov, qlen, fw
            arr[i].quality = RH_NO_QUALITY if qoff is None else qoff
            arr[i].mirrored = 1 if mirrored else 0
        _check(self.lib.rh_layers_add(self.handle, batch.handle, arr, len(sources)),
               "rh_layers_add")

    def finish(self) -> list:
        """Layers grouped by window: (window, overlap, q_begin, q_end, begin, end)."""
        n = ctypes.c_uint64()
        _check(self.lib.rh_layers_finish(self.handle, ctypes.byref(n)), "rh_layers_finish")
        recs = (RhLayerRecord * max(1, n.value))()
        _check(self.lib.rh_layers_download(self.handle, recs, n.value), "rh_layers_download")
        return [(r.window, r.overlap, r.q_begin, r.q_end, r.begin, r.end)
                for r in recs[:n.value]]


