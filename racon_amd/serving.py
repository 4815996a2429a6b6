"""Continuous-batching polishing server over the POA batch API.

Requests are polishing jobs already cut into windows (racon's input step:
Polisher::initialize, src/polisher.cpp:192-457, and the breaking points of
find_overlap_breaking_points, src/polisher.cpp:462-484). The server admits
them by a device-memory budget, co-batches the windows of every admitted
request into launches of the persistent POA kernel, and hands each request
its consensus list when its last window is back, with its latency.

Scheduling:
  * admission is first come, first served: a request enters when the bytes
    of its windows fit in the budget left by the admitted, unfinished ones
    (no later request overtakes a waiting one); a request larger than the
    whole budget is refused at submit;
  * each executor (one PoaBatch, i.e. one HIP stream and its own device
    scratch) takes the next windows of the admitted requests in admission
    order, up to `launch_windows`, whenever it is free; requests that arrive
    while launches are in flight join the next launch (continuous batching);
  * with two or more executors, one launch's packing and copies overlap
    another's kernel, and the kernels of two launches share the device
    (the second fills the CUs the first's tail frees).

A window's consensus depends on that window alone (one window per wave,
no state shared between windows), so every request's output is the same
alone and in any mix; the GPU tests check it.
"""
from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

from ._lib import RH_POA_BATCHED, PoaBatch, Window


def window_bytes(w: Window) -> int:
    """Bytes of one window's input (sequences and qualities), the unit of
    the admission budget."""
    n = len(w.backbone) + len(w.backbone_quality or b"")
    for layer in w.layers:
        n += len(layer.sequence) + (len(layer.quality) if layer.quality else 0)
    return n


@dataclass
class Request:
    """One polishing job and, once done, its results and timings."""
    rid: int
    windows: List[Window]
    bytes: int
    arrival: float
    admitted: Optional[float] = None
    done: Optional[float] = None
    consensus: List[Optional[bytes]] = field(default_factory=list)
    status: List[int] = field(default_factory=list)
    launches: set = field(default_factory=set)
    error: Optional[str] = None
    _left: int = 0
    _next: int = 0
    _event: threading.Event = field(default_factory=threading.Event)

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    def wait(self, timeout: Optional[float] = None) -> "Request":
        if not self._event.wait(timeout):
            raise TimeoutError(f"request {self.rid} not done")
        if self.error is not None:
            raise RuntimeError(f"request {self.rid}: {self.error}")
        return self


class GpuExecutor:
    """Runs one launch of windows on the device (rh_poa_upload, rh_poa_run,
    rh_poa_download on this executor's own batch and stream)."""

    def __init__(self, match: int, mismatch: int, gap: int, trim: bool, device: int = 0):
        self.batch = PoaBatch(match, mismatch, gap, trim=trim, device=device)

    def __call__(self, windows: Sequence[Window]):
        self.batch.upload(windows)
        self.batch.run(RH_POA_BATCHED)
        return self.batch.download()

    def close(self) -> None:
        self.batch.close()


class PolishServer:
    """Admits requests by budget and serves their windows with `executors`
    workers. `executor_factory()` returns a callable windows -> (consensus
    list, status list); the default runs the HIP kernel."""

    def __init__(self, budget_bytes: int, executors: int = 2, launch_windows: int = 4096,
                 scores=(3, -5, -4), trim: bool = True,
                 executor_factory: Optional[Callable[[], Callable]] = None):
        if budget_bytes <= 0 or executors <= 0 or launch_windows <= 0:
            raise ValueError("budget, executors and launch size must be positive")
        self.budget = budget_bytes
        self.launch_windows = launch_windows
        self._lock = threading.Condition()
        self._pending: collections.deque = collections.deque()   # not yet admitted
        self._admitted: collections.deque = collections.deque()  # windows left to launch
        self._in_use = 0
        self._closed = False
        self._next_rid = 0
        self._launch_id = 0
        self.completed: List[Request] = []
        self.launch_log: List[tuple] = []  # (launch id, windows, requests)
        if executor_factory is None:
            def executor_factory():
                return GpuExecutor(*scores, trim=trim)
        self._executors = [executor_factory() for _ in range(executors)]
        self._threads = [threading.Thread(target=self._worker, args=(ex,), daemon=True)
                         for ex in self._executors]
        for t in self._threads:
            t.start()

    # ---- client side ----
    def submit(self, windows: Sequence[Window]) -> Request:
        return self.submit_many([windows])[0]

    def submit_many(self, jobs: Sequence[Sequence[Window]]) -> List[Request]:
        """Submits several requests at one instant (in list order)."""
        jobs = [list(ws) for ws in jobs]
        sizes = [sum(window_bytes(w) for w in ws) for ws in jobs]
        with self._lock:
            if self._closed:
                raise RuntimeError("server closed")
            for size in sizes:
                if size > self.budget:
                    raise ValueError(f"request of {size} bytes exceeds the budget of "
                                     f"{self.budget} bytes")
            now = time.perf_counter()
            reqs = []
            for ws, size in zip(jobs, sizes):
                req = Request(self._next_rid, ws, size, now)
                self._next_rid += 1
                req.consensus = [None] * len(ws)
                req.status = [0] * len(ws)
                req._left = len(ws)
                self._pending.append(req)
                reqs.append(req)
            self._admit()
            self._lock.notify_all()
        return reqs

    def warm(self, windows: Sequence[Window]) -> None:
        """Runs every executor once on `windows`, one after the other, before
        requests arrive: device scratch is sized on first use (and grows when
        a launch needs more), so warming with the largest windows expected
        keeps allocations out of the served launches."""
        for ex in self._executors:
            ex(list(windows))

    def in_use(self) -> int:
        with self._lock:
            return self._in_use

    def close(self) -> None:
        """Finishes every submitted request, then stops the workers."""
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for t in self._threads:
            t.join()
        for ex in self._executors:
            if hasattr(ex, "close"):
                ex.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- scheduler (callers hold the lock) ----
    def _admit(self) -> None:
        while self._pending and self._in_use + self._pending[0].bytes <= self.budget:
            req = self._pending.popleft()
            req.admitted = time.perf_counter()
            if req.windows:
                self._in_use += req.bytes
                self._admitted.append(req)
            else:  # nothing to launch: done on admission
                req.done = req.admitted
                self.completed.append(req)
                req._event.set()

    def _take(self):
        """Next launch: windows of the admitted requests in admission order."""
        items = []
        while self._admitted and len(items) < self.launch_windows:
            req = self._admitted[0]
            k = min(len(req.windows) - req._next, self.launch_windows - len(items))
            items.extend((req, req._next + i) for i in range(k))
            req._next += k
            if req._next == len(req.windows):
                self._admitted.popleft()
        return items

    def _finish(self, req: Request) -> None:
        req.done = time.perf_counter()
        self._in_use -= req.bytes
        self.completed.append(req)
        req._event.set()
        self._admit()

    def _worker(self, executor) -> None:
        while True:
            with self._lock:
                while not self._admitted and not (self._closed and not self._pending):
                    self._lock.wait()
                if not self._admitted:
                    return  # closed and drained
                items = self._take()
                launch = self._launch_id
                self._launch_id += 1
                self.launch_log.append((launch, len(items), len({r.rid for r, _ in items})))
            error = None
            try:
                out, status = executor([req.windows[i] for req, i in items])
            except Exception as e:  # reported on every request of the launch
                error = str(e)
            with self._lock:
                for k, (req, i) in enumerate(items):
                    req.launches.add(launch)
                    if error is None:
                        req.consensus[i] = out[k]
                        req.status[i] = status[k]
                    elif req.error is None:
                        req.error = error
                    req._left -= 1
                    if req._left == 0:
                        self._finish(req)
                self._lock.notify_all()


def percentile(values: Sequence[float], q: float) -> float:
    """Nearest-rank percentile (q in 0..100)."""
    if not values:
        return float("nan")
    s = sorted(values)
    k = max(0, min(len(s) - 1, int(-(-q * len(s) // 100)) - 1))
    return s[k]
