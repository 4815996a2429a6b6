"""Serving loop (rh_server_* through racon_amd.serving): admission by device
budget, first come first served, windows of several requests in one launch,
and every request's result the same alone and in the mix.

CPU tests check the argument handling of the C ABI (no device needed); the
GPU tests run the native server.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from racon_amd import _lib
from racon_amd.serving import PolishServer, percentile
from synth import make_windows, oracle_args


def test_percentile_nearest_rank():
    v = [5.0, 1.0, 3.0, 2.0, 4.0]
    assert percentile(v, 50) == 3.0
    assert percentile(v, 99) == 5.0
    assert percentile(v, 20) == 1.0
    assert percentile([7.0], 99) == 7.0


def test_server_arguments_are_checked():
    lib = _lib.load_library()
    h = ctypes.c_void_p()
    # no executors, and a budget below 1 MiB: refused before any device call
    assert lib.rh_server_create(0, 3, -5, -4, 1, 1 << 30, 0, 2048, ctypes.byref(h)) == -1
    assert b"executors" in lib.rh_last_error()
    assert lib.rh_server_create(0, 3, -5, -4, 1, 1000, 2, 2048, ctypes.byref(h)) == -1
    assert not h.value
    with pytest.raises(ValueError):
        PolishServer(0)


@pytest.mark.skipif(_lib.load_library().rh_device_count() > 0, reason="a device is present")
def test_server_without_device_fails_loudly():
    with pytest.raises(RuntimeError, match="rh_server_create"):
        PolishServer(1 << 30)


def batch_alone(ws):
    b = _lib.PoaBatch(3, -5, -4, trim=True)
    b.upload(ws)
    b.run()
    out = b.download()
    b.close()
    return out


@pytest.mark.gpu
def test_gpu_mix_matches_alone():
    sys.path.insert(0, _lib.PKG_DIR + "/../oracle")
    import oracle_py

    reqs = [make_windows(600, 3, length=200, depth=10),
            make_windows(601, 6, length=300, depth=14),
            make_windows(602, 2, length=120, depth=1),   # < 3 sequences: no kernel
            make_windows(603, 4, length=500, depth=20)]
    alone = [batch_alone(ws) for ws in reqs]
    with PolishServer(1 << 30, executors=2, launch_windows=5) as srv:
        # back to back: the first launch takes request 0 and part of request 1
        hs = srv.submit_many(reqs[:3])
        hs.append(srv.submit(reqs[3]))  # arrives while launches are in flight
        for h in hs:
            h.wait(120)
        st = srv.stats()
    print(f"stats {st}", flush=True)
    assert st["completed"] == 4 and st["pool_in_use"] == 0
    assert st["shared_launches"] >= 1
    assert st["windows_launched"] == 3 + 6 + 4  # the shallow windows need no kernel
    for k, h in enumerate(hs):
        print(f"request {k}: {len(h.windows)} windows, {h.launches} launches "
              f"({h.shared_launches} shared)", flush=True)
        assert (h.consensus, h.status) == (alone[k][0], alone[k][1]), f"request {k}"
        assert h.done >= h.admitted >= h.arrival
    assert hs[0].shared_launches >= 1
    # and the mix against the oracle on the small request
    for w, cons, st in zip(reqs[0], hs[0].consensus, hs[0].status):
        ocons, ost, _, _ = oracle_py.window_consensus(*oracle_args(w), True, 3, -5, -4)
        assert (cons, st) == (ocons, ost)


@pytest.mark.gpu
def test_gpu_budget_admits_in_arrival_order():
    reqs = [make_windows(610 + k, n, length=500, depth=30) for k, n in enumerate([60, 60, 4, 60])]
    with PolishServer(1 << 30, executors=1) as probe:
        sizes = [probe.submit(ws).wait(120).device_bytes for ws in reqs]
    # a pool that holds the largest request but never two of the big ones
    pool = max(sizes) + min(sizes) + 4096
    budget = pool * 32 // 31 + 4096
    with PolishServer(budget, executors=2, launch_windows=16) as srv:
        hs = [srv.submit(ws) for ws in reqs]
        for h in hs:
            h.wait(120)
        assert srv.stats()["pool_in_use"] == 0
    print("sizes", sizes, "admitted", [h.admitted for h in hs], "done", [h.done for h in hs],
          flush=True)
    for a, b in zip(hs, hs[1:]):
        assert b.admitted >= a.admitted  # first come, first served
    assert hs[1].admitted >= hs[0].done  # request 1 waited for request 0's bytes
    for ws, h in zip(reqs, hs):
        assert (h.consensus, h.status) == batch_alone(ws)


@pytest.mark.gpu
def test_gpu_request_over_budget_is_refused_and_empty_requests_complete():
    big = make_windows(620, 90, length=500, depth=30)  # ~2 MB
    with PolishServer(1 << 20, executors=1) as srv:
        with pytest.raises(RuntimeError, match="exceeds the server's budget"):
            srv.submit(big)
        empty = srv.submit([])
        shallow = srv.submit(make_windows(621, 2, length=80, depth=1))
        small = srv.submit(big[:1])
        assert empty.wait(60).consensus == []
        assert shallow.wait(60).status == [_lib.RH_WINDOW_UNPOLISHED] * 2
        assert shallow.consensus == [w.backbone for w in shallow.windows]
        assert (small.wait(60).consensus, small.status) == batch_alone(big[:1])


# ---- cfg5-sized requests: contigs of 20 kbp to 2 Mbp at 20-100x ----

CFG5_REQUESTS = [  # (seed, contig bp, depth): 40, ~400 and ~4,000 windows of 500 bp
    (1051, 20000, 20),
    (1052, 200000, 60),
    (1053, 2000000, 100),
]


def cfg5_request_windows(workdir, seed, contig, depth):
    """Windows of one polishing job: simgen contig + reads, then the product
    CLI's input step and GPU overlap alignment (--dump-windows)."""
    from racon_amd.dumps import read_windows

    prefix = os.path.join(workdir, f"req{seed}")
    bindir = os.path.join(_lib.PKG_DIR, "bin")
    subprocess.run([os.path.join(bindir, "simgen"), "contig", str(seed), prefix, str(contig),
                    str(depth), "8000", "4000", "0.12", "0.4:0.2:0.4"], check=True,
                   stdout=subprocess.DEVNULL)
    dump = prefix + ".windows.bin"
    subprocess.run([os.path.join(bindir, "racon"), "-t", "16", "-w", "500", "-m", "3", "-x", "-5",
                    "-g", "-4", "-q", "10", "-e", "0.3", "--dump-windows", dump,
                    prefix + ".reads.fastq", prefix + ".paf", prefix + ".contigs.fasta"],
                   check=True, stdout=subprocess.DEVNULL)
    return read_windows(dump)


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_gpu_cfg5_sized_requests_alone_and_mixed(tmp_path):
    sys.path.insert(0, _lib.PKG_DIR + "/../oracle")
    import oracle_py

    reqs = []
    for seed, contig, depth in CFG5_REQUESTS:
        ws = cfg5_request_windows(str(tmp_path), seed, contig, depth)
        print(f"request seed {seed}: {contig} bp at {depth}x -> {len(ws)} windows", flush=True)
        reqs.append(ws)
    assert [len(ws) for ws in reqs][0] >= 40 and len(reqs[2]) >= 3900
    alone = [batch_alone(ws) for ws in reqs]
    print("alone done", flush=True)
    with PolishServer(8 << 30, executors=4, launch_windows=2048) as srv:
        # the big job first; the small ones (and a second copy of the middle
        # one) arrive together while it runs
        order = [2, 0, 1, 1]
        hs = [srv.submit(reqs[2])] + srv.submit_many([reqs[0], reqs[1], reqs[1]])
        for h in hs:
            h.wait(300)
        st = srv.stats()
    print(f"mixed stats {st}", flush=True)
    for k, h in zip(order, hs):
        print(f"request {k}: latency {h.latency:.3f} s, {h.launches} launches "
              f"({h.shared_launches} shared)", flush=True)
        assert (h.consensus, h.status) == alone[k], f"request {k}"
    assert st["shared_launches"] >= 1
    # the oracle on a sample of every request's windows
    rng = np.random.default_rng(5)
    for k, ws in enumerate(reqs):
        sample = rng.choice(len(ws), size=min(12, len(ws)), replace=False)
        for i in sample:
            ocons, ost, _, _ = oracle_py.window_consensus(*oracle_args(ws[i]), True, 3, -5, -4)
            assert (alone[k][0][i], alone[k][1][i]) == (ocons, ost), f"request {k} window {i}"
