"""Serving loop: admission by budget, co-batching, per-request reassembly.

The scheduler runs on the CPU with a stand-in executor (a deterministic
function of each window, with a delay so launches overlap); the GPU test
checks that every request's consensus is the same served alone and in a mix.
"""
import sys
import threading
import time

import pytest

from racon_amd import _lib
from racon_amd.serving import PolishServer, percentile, window_bytes
from synth import make_windows, oracle_args


def fake_consensus(w):
    return w.backbone[::-1] + bytes([65 + len(w.layers) % 26])


class FakeExecutor:
    def __init__(self, delay=0.01, fail=False):
        self.delay = delay
        self.fail = fail
        self.calls = 0

    def __call__(self, windows):
        self.calls += 1
        time.sleep(self.delay)
        if self.fail:
            raise RuntimeError("device rejected the launch")
        return [fake_consensus(w) for w in windows], [1] * len(windows)


def requests_of(seed, sizes, length=60, depth=4):
    return [make_windows(seed + i, n, length=length, depth=depth) for i, n in enumerate(sizes)]


def test_window_bytes_counts_sequences_and_qualities():
    w = make_windows(1, 1, length=50, depth=3)[0]
    expect = len(w.backbone) + len(w.backbone_quality) + sum(
        len(l.sequence) + len(l.quality) for l in w.layers)
    assert window_bytes(w) == expect


def test_mix_matches_alone_and_launches_mix_requests():
    reqs = requests_of(100, [5, 1, 9, 3, 7, 2])
    with PolishServer(10**9, executors=2, launch_windows=6,
                      executor_factory=lambda: FakeExecutor(0.05)) as srv:
        handles = []
        threads = []

        def client(ws, delay):
            time.sleep(delay)
            handles.append(srv.submit(ws))

        for k, ws in enumerate(reqs):
            t = threading.Thread(target=client, args=(ws, 0.005 * k))
            t.start()
            threads.append(t)
        for t in threads:
            t.join()
        for h in handles:
            h.wait(10)
    assert len(handles) == len(reqs)
    for h in handles:
        assert h.consensus == [fake_consensus(w) for w in h.windows]
        assert h.status == [1] * len(h.windows)
        assert h.done >= h.admitted >= h.arrival
    assert srv.in_use() == 0
    assert sum(n for _, n, _ in srv.launch_log) == sum(len(r) for r in reqs)
    assert max(n for _, n, _ in srv.launch_log) <= 6
    # windows of different requests share launches
    assert any(k > 1 for _, _, k in srv.launch_log)


def test_budget_admits_in_arrival_order():
    reqs = requests_of(200, [4, 4, 1, 4])
    sizes = [sum(window_bytes(w) for w in ws) for ws in reqs]
    budget = max(sizes) + min(sizes) - 1  # never two of the big ones at once
    with PolishServer(budget, executors=2, launch_windows=2,
                      executor_factory=lambda: FakeExecutor(0.01)) as srv:
        hs = [srv.submit(ws) for ws in reqs]
        assert srv.in_use() <= budget
        for h in hs:
            h.wait(10)
    # first come, first served: no request is admitted before an earlier one
    for a, b in zip(hs, hs[1:]):
        assert b.admitted >= a.admitted
    # request 1 waited for request 0's bytes to come back
    assert hs[1].admitted >= hs[0].done
    for h in hs:
        assert h.consensus == [fake_consensus(w) for w in h.windows]


def test_submit_many_shares_the_first_launch():
    reqs = requests_of(700, [2, 2, 3])
    with PolishServer(10**9, executors=1, launch_windows=5,
                      executor_factory=lambda: FakeExecutor(0.01)) as srv:
        hs = srv.submit_many(reqs)
        for h in hs:
            h.wait(10)
    assert [n for _, n, _ in srv.launch_log] == [5, 2]
    assert srv.launch_log[0][2] == 3
    assert hs[2].launches == {0, 1}
    for h in hs:
        assert h.consensus == [fake_consensus(w) for w in h.windows]


def test_warm_runs_every_executor_once():
    made = []

    def factory():
        made.append(FakeExecutor(0))
        return made[-1]

    with PolishServer(10**6, executors=3, executor_factory=factory) as srv:
        srv.warm(make_windows(800, 2, length=40, depth=2))
        assert [ex.calls for ex in made] == [1, 1, 1]
        srv.submit(make_windows(801, 1, length=40, depth=2)).wait(10)
    assert sum(ex.calls for ex in made) == 4
    assert srv.launch_log and srv.launch_log[0][0] == 0


def test_request_over_budget_is_refused():
    ws = make_windows(300, 3, length=80, depth=4)
    with PolishServer(window_bytes(ws[0]), executors=1,
                      executor_factory=lambda: FakeExecutor(0)) as srv:
        with pytest.raises(ValueError, match="exceeds the budget"):
            srv.submit(ws)
        srv.submit(ws[:1]).wait(10)


def test_empty_request_and_close_drains():
    with PolishServer(10**6, executors=1, executor_factory=lambda: FakeExecutor(0.01)) as srv:
        empty = srv.submit([])
        busy = srv.submit(make_windows(400, 3, length=40, depth=2))
    assert empty.wait(0).consensus == []
    assert busy.wait(0).consensus == [fake_consensus(w) for w in busy.windows]
    with pytest.raises(RuntimeError, match="closed"):
        srv.submit([])


def test_executor_failure_reaches_every_request_of_the_launch():
    with PolishServer(10**6, executors=1, launch_windows=8,
                      executor_factory=lambda: FakeExecutor(0, fail=True)) as srv:
        hs = [srv.submit(ws) for ws in requests_of(500, [2, 3])]
        for h in hs:
            with pytest.raises(RuntimeError, match="rejected"):
                h.wait(10)
    assert srv.in_use() == 0


def test_percentile_nearest_rank():
    v = [5.0, 1.0, 3.0, 2.0, 4.0]
    assert percentile(v, 50) == 3.0
    assert percentile(v, 99) == 5.0
    assert percentile(v, 20) == 1.0
    assert percentile([7.0], 99) == 7.0


@pytest.mark.gpu
def test_gpu_mix_matches_alone():
    sys.path.insert(0, _lib.PKG_DIR + "/../oracle")
    import oracle_py

    reqs = [make_windows(600, 3, length=200, depth=10),
            make_windows(601, 6, length=300, depth=14),
            make_windows(602, 1, length=120, depth=2),   # < 3 sequences
            make_windows(603, 4, length=500, depth=20)]
    alone = []
    for ws in reqs:
        b = _lib.PoaBatch(3, -5, -4, trim=True)
        b.upload(ws)
        b.run()
        alone.append(b.download())
        b.close()
    with PolishServer(10**9, executors=2, launch_windows=5) as srv:
        # one instant: the first launch takes request 0 and part of request 1
        hs = srv.submit_many(reqs[:3])
        hs.append(srv.submit(reqs[3]))  # arrives while launches are in flight
        for h in hs:
            h.wait(120)
    assert srv.launch_log[0][2] == 2
    for k, h in enumerate(hs):
        print(f"request {k}: {len(h.windows)} windows, launches {sorted(h.launches)}",
              flush=True)
        assert (h.consensus, h.status) == (alone[k][0], alone[k][1]), f"request {k}"
    # and the mix against the oracle on the small request
    for w, cons, st in zip(reqs[0], hs[0].consensus, hs[0].status):
        ocons, ost, _, _ = oracle_py.window_consensus(*oracle_args(w), True, 3, -5, -4)
        assert (cons, st) == (ocons, ost)
