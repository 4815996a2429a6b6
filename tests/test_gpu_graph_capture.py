"""The allocation-free aligner entry point under HIP graph capture. The call
puts two kernels on the caller's stream (the queue-counter fill and the batch
kernel; no allocation, no memset), so it can be captured into a graph. Every
replay, with an eager call on other pairs that reuses the same static scratch
in between, must give the eager call's alignments. Replays run on the HIP
runtime libracon_hip.so itself loaded.
"""
import ctypes

import numpy as np
import pytest

from racon_amd import _lib
from synth import make_pair
from test_zz_no_alloc import Hip, batch_align, device_batch

pytestmark = pytest.mark.gpu

HIP_STREAM_CAPTURE_MODE_RELAXED = 2


def graph_api(rt):
    vp = ctypes.c_void_p
    for name, args in (("hipStreamCreate", [ctypes.POINTER(vp)]),
                       ("hipStreamSynchronize", [vp]),
                       ("hipStreamDestroy", [vp]),
                       ("hipStreamBeginCapture", [vp, ctypes.c_int]),
                       ("hipStreamEndCapture", [vp, ctypes.POINTER(vp)]),
                       ("hipGraphInstantiate", [ctypes.POINTER(vp), vp, vp, ctypes.c_char_p,
                                                ctypes.c_size_t]),
                       ("hipGraphLaunch", [vp, vp]),
                       ("hipGraphExecDestroy", [vp]),
                       ("hipGraphDestroy", [vp])):
        getattr(rt, name).argtypes = args


def test_noalloc_call_captured_in_a_graph_replays_exactly():
    hip = Hip()
    rt = hip.rt
    graph_api(rt)
    lib = _lib.load_library()
    rng = np.random.default_rng(91)
    pairs = [make_pair(rng, int(n)) for n in (70, 640, 1500, 3000, 5200)]
    pairs += [make_pair(rng, int(n)) for n in rng.integers(100, 2000, size=24)]
    other = [make_pair(rng, int(n)) for n in (900, 2500)]
    want_ops, want_dist = batch_align(pairs)
    other_ops, other_dist = batch_align(other)
    stream = ctypes.c_void_p()
    graph = ctypes.c_void_p()
    exe = ctypes.c_void_p()
    try:
        assert rt.hipStreamCreate(ctypes.byref(stream)) == 0
        args, read, clear = device_batch(hip, pairs)
        oargs, oread, oclear = device_batch(hip, other)
        # eager call first: it also looks up the scratch array's address,
        # which is not a stream operation
        assert lib.rh_align_run_noalloc(0, stream, *args) == 0, lib.rh_last_error().decode()
        assert rt.hipStreamSynchronize(stream) == 0
        assert read() == (want_ops, want_dist, [0] * len(pairs))
        print("eager call matches the batch entry point", flush=True)

        assert rt.hipStreamBeginCapture(stream, HIP_STREAM_CAPTURE_MODE_RELAXED) == 0
        rc = lib.rh_align_run_noalloc(0, stream, *args)
        assert rt.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
        assert rc == 0, lib.rh_last_error().decode()
        assert rt.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
        print("captured and instantiated", flush=True)

        for replay in range(3):
            clear()
            oclear()
            # eager work between replays, on the same scratch
            assert lib.rh_align_run_noalloc(0, stream, *oargs) == 0
            assert rt.hipStreamSynchronize(stream) == 0
            assert oread() == (other_ops, other_dist, [0, 0])
            assert rt.hipGraphLaunch(exe, stream) == 0
            assert rt.hipStreamSynchronize(stream) == 0
            got = read()
            print(f"replay {replay}: {len(got[0])} pairs, first distances {got[1][:4]}",
                  flush=True)
            assert got == (want_ops, want_dist, [0] * len(pairs)), f"replay {replay}"
    finally:
        if exe:
            rt.hipGraphExecDestroy(exe)
        if graph:
            rt.hipGraphDestroy(graph)
        if stream:
            rt.hipStreamSynchronize(stream)
            rt.hipStreamDestroy(stream)
        hip.free()
