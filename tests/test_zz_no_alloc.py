"""The allocation-free aligner entry point (rh_align_run_noalloc) against the
batch entry point (rh_align_upload / run / download) on the same pairs,
called twice in a row, the second time with a smaller input: the scratch
array keeps the first call's contents, so the second call must not depend
on them. Device buffers for the inputs and outputs are the caller's (torch
allocates them here); the entry point itself allocates nothing.
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


def noalloc_align(pairs, torch):
    n = len(pairs)
    dev = torch.device("cuda", 0)
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
    arena = torch.tensor(np.frombuffer(bytes(blob) + b"\0", dtype=np.uint8), device=dev)
    d_pairs = torch.tensor(recs.view(np.uint8), device=dev)
    ops = torch.full((max(1, out),), 255, dtype=torch.uint8, device=dev)
    res_dt = np.dtype(_lib.DEV_ALIGN_RESULT_FIELDS, align=True)
    results = torch.zeros(n * res_dt.itemsize, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev)
    lib = _lib.load_library()
    rc = lib.rh_align_run_noalloc(0, ctypes.c_void_p(stream.cuda_stream), n,
                                  ctypes.c_void_p(d_pairs.data_ptr()),
                                  ctypes.c_void_p(arena.data_ptr()),
                                  max(len(q) for q, _ in pairs), max(len(t) for _, t in pairs),
                                  ctypes.c_void_p(ops.data_ptr()),
                                  ctypes.c_void_p(results.data_ptr()))
    assert rc == 0, lib.rh_last_error().decode()
    stream.synchronize()
    res = results.cpu().numpy().view(res_dt)
    ops_h = ops.cpu().numpy().tobytes()
    got_ops = [ops_h[int(recs[i]["ops_offset"]):int(recs[i]["ops_offset"]) + int(res[i]["n_ops"])]
               for i in range(n)]
    return got_ops, [int(d) for d in res["edit_distance"]], [int(s) for s in res["status"]]


def test_noalloc_matches_batch_entry_twice():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(77)
    first = [make_pair(rng, int(n)) for n in (5, 64, 65, 300, 1200, 3000, 5000, 9000)]
    first += [make_pair(rng, int(n)) for n in rng.integers(100, 2500, size=40)]
    second = [make_pair(rng, int(n)) for n in (40, 700, 2000)]
    for pairs in (first, second):
        want_ops, want_dist = batch_align(pairs)
        got_ops, got_dist, status = noalloc_align(pairs, torch)
        print(f"{len(pairs)} pairs: statuses {set(status)}", flush=True)
        assert status == [0] * len(pairs)
        assert got_dist == want_dist
        assert got_ops == want_ops


def test_noalloc_refuses_oversized_pairs():
    lib = _lib.load_library()
    # 4 Mbp x 4 Mbp pairs need a slot beyond the scratch: refused before any
    # launch (the device pointers are never read)
    rc = lib.rh_align_run_noalloc(0, None, 1, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                  1 << 22, 1 << 22, ctypes.c_void_p(16), ctypes.c_void_p(16))
    assert rc == -1
    assert b"scratch" in lib.rh_last_error()
