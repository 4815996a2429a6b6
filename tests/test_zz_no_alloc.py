"""The allocation-free aligner entry point (rh_align_run_noalloc) against the
batch entry point (rh_align_upload / run / download) on the same pairs,
called twice in a row, the second time with a smaller input: the scratch
array keeps the first call's contents, so the second call must not depend
on them. Device buffers for the inputs and outputs are the caller's (the
test allocates them with the HIP runtime); the entry point allocates nothing.
"""
import ctypes

import numpy as np
import pytest

from racon_amd import _lib
from synth import make_pair

pytestmark = pytest.mark.gpu


def batch_align(pairs):
    b = _lib.AlignBatch()
    b.upload(pairs)
    b.run()
    ops, dist = b.download()
    b.close()
    return ops, dist


class Hip:
    """The caller's side of the device buffers: hipMalloc / hipMemcpy /
    hipFree of the HIP runtime libracon_hip.so itself loaded (the test
    allocates; the entry point under test does not)."""

    def __init__(self):
        _lib.load_library()
        # the soname libracon_hip.so was linked against: the same loaded copy
        self.rt = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.held = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(1, nbytes)) == 0
        self.held.append(p)
        return p

    def to_device(self, data: bytes):
        p = self.alloc(len(data))
        self.copy_in(p, data)
        return p

    def copy_in(self, p, data: bytes):
        buf = ctypes.create_string_buffer(data, len(data))
        assert self.rt.hipMemcpy(p, buf, len(data), 1) == 0  # host to device

    def to_host(self, p, nbytes) -> bytes:
        buf = ctypes.create_string_buffer(max(1, nbytes))
        assert self.rt.hipMemcpy(buf, p, nbytes, 2) == 0  # device to host
        return buf.raw[:nbytes]

    def free(self):
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


def device_batch(hip, pairs):
    """Caller-owned device copies of the pairs (arena, rh_dev_pair records)
    and the output buffers. Returns the rh_align_run_noalloc arguments after
    (device, stream), a reader of the results (ops, distances, statuses) and
    a clearer that refills the outputs with junk."""
    n = len(pairs)
    blob = bytearray()
    recs = np.zeros(n, dtype=np.dtype(_lib.DEV_PAIR_FIELDS, align=True))
    out = 0
    for i, (q, t) in enumerate(pairs):
        recs[i]["query_offset"] = len(blob)
        blob += q
        recs[i]["target_offset"] = len(blob)
        blob += t
        recs[i]["ops_offset"] = out
        recs[i]["query_length"] = len(q)
        recs[i]["target_length"] = len(t)
        out += len(q) + len(t)
    res_dt = np.dtype(_lib.DEV_ALIGN_RESULT_FIELDS, align=True)
    arena = hip.to_device(bytes(blob) + b"\0")
    d_pairs = hip.to_device(recs.tobytes())
    ops = hip.to_device(b"\xff" * max(1, out))
    results = hip.to_device(b"\0" * (n * res_dt.itemsize))
    args = (n, d_pairs, arena, max(len(q) for q, _ in pairs), max(len(t) for _, t in pairs), ops,
            results)

    def read():
        res = np.frombuffer(hip.to_host(results, n * res_dt.itemsize), dtype=res_dt)
        ops_h = hip.to_host(ops, out)
        got_ops = [ops_h[int(recs[i]["ops_offset"]):int(recs[i]["ops_offset"]) +
                         int(res[i]["n_ops"])] for i in range(n)]
        return (got_ops, [int(d) for d in res["edit_distance"]],
                [int(s) for s in res["status"]])

    def clear():
        hip.copy_in(ops, b"\xff" * max(1, out))
        hip.copy_in(results, b"\x5a" * (n * res_dt.itemsize))

    return args, read, clear


def noalloc_align(pairs, hip):
    args, read, _ = device_batch(hip, pairs)
    lib = _lib.load_library()
    rc = lib.rh_align_run_noalloc(0, None, *args)
    assert rc == 0, lib.rh_last_error().decode()
    assert hip.rt.hipDeviceSynchronize() == 0
    return read()


def test_noalloc_matches_batch_entry_twice():
    hip = Hip()
    rng = np.random.default_rng(77)
    first = [make_pair(rng, int(n)) for n in (5, 64, 65, 300, 1200, 3000, 5000, 9000)]
    first += [make_pair(rng, int(n)) for n in rng.integers(100, 2500, size=40)]
    second = [make_pair(rng, int(n)) for n in (40, 700, 2000)]
    try:
        for pairs in (first, second):
            want_ops, want_dist = batch_align(pairs)
            got_ops, got_dist, status = noalloc_align(pairs, hip)
            print(f"{len(pairs)} pairs: statuses {set(status)}", flush=True)
            assert status == [0] * len(pairs)
            assert got_dist == want_dist
            assert got_ops == want_ops
    finally:
        hip.free()


def test_noalloc_refuses_oversized_pairs():
    lib = _lib.load_library()
    # 4 Mbp x 4 Mbp pairs need a slot beyond the scratch: refused before any
    # launch (the device pointers are never read)
    rc = lib.rh_align_run_noalloc(0, None, 1, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                  1 << 22, 1 << 22, ctypes.c_void_p(16), ctypes.c_void_p(16))
    assert rc == -1
    assert b"scratch" in lib.rh_last_error()
