"""Python handle of the native serving loop (rh_server_* of include/racon_hip.h).

Requests are polishing jobs already cut into windows (racon's input step:
Polisher::initialize, src/polisher.cpp:192-457, and the breaking points of
find_overlap_breaking_points, src/polisher.cpp:462-484). The server itself
is C++ inside libracon_hip.so (racon_amd/csrc/hip/rh_serve.cpp):

  * submit packs a request on the caller's thread and queues it; admission
    is first come, first served into a device pool of `budget_bytes`, where
    the request's windows stay resident until its last window is back (no
    later request overtakes a waiting one; a request larger than the pool
    is refused at submit);
  * `executors` native threads, each with its own HIP stream and share of
    the persistent kernel's resident waves, launch the next windows of the
    admitted requests in admission order, up to `launch_windows` per launch;
    requests that arrive while kernels run join the next launch, and the
    kernels of different executors run side by side on the device.

A window's consensus depends on that window alone (one window per wave,
no state shared between windows), so every request's output is the same
alone and in any mix; the GPU tests check it.
"""
from __future__ import annotations

import ctypes
import time
from typing import List, Optional, Sequence

from ._lib import (RH_OK, RH_TIMEOUT, RhRequestInfo, RhServerStats, RhWindow, Window, _check,
                   load_library, pack_windows)


class Request:
    """One submitted request; `wait()` fills in its results and timings
    (seconds on the server's clock)."""

    def __init__(self, server: "PolishServer", rid: int, windows: Sequence[Window]):
        self._server = server
        self.rid = rid
        self.windows = windows
        self.consensus: Optional[List[bytes]] = None
        self.status: Optional[List[int]] = None
        self.arrival = self.admitted = self.done = None
        self.launches = self.shared_launches = 0
        self.device_bytes = 0

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    def wait(self, timeout: Optional[float] = None) -> "Request":
        if self.consensus is not None:
            return self
        lib, handle = self._server.lib, self._server.handle
        deadline = None if timeout is None else time.perf_counter() + timeout
        while True:
            step = 200 if deadline is None else max(
                0, min(200, int((deadline - time.perf_counter()) * 1e3)))
            rc = lib.rh_server_wait(handle, self.rid, step)
            if rc == RH_OK:
                break
            if rc != RH_TIMEOUT:
                _check(rc, f"request {self.rid}")
            if deadline is not None and time.perf_counter() >= deadline:
                raise TimeoutError(f"request {self.rid} not done")
        info = RhRequestInfo()
        _check(lib.rh_server_request_info(handle, self.rid, ctypes.byref(info)),
               "rh_server_request_info")
        n = info.windows
        cap = max(1, info.output_bound)
        out = ctypes.create_string_buffer(cap)
        offsets = (ctypes.c_uint64 * max(1, n))()
        lengths = (ctypes.c_uint32 * max(1, n))()
        status = (ctypes.c_uint8 * max(1, n))()
        _check(lib.rh_server_fetch(handle, self.rid, out, cap, offsets, lengths, status),
               "rh_server_fetch")
        raw = out.raw
        self.consensus = [raw[offsets[i]:offsets[i] + lengths[i]] for i in range(n)]
        self.status = [status[i] for i in range(n)]
        self.arrival, self.admitted, self.done = info.arrival_s, info.admitted_s, info.done_s
        self.launches, self.shared_launches = info.launches, info.shared_launches
        self.device_bytes = info.device_bytes
        return self


class PolishServer:
    """A native serving loop on one device (see the module docstring)."""

    def __init__(self, budget_bytes: int, executors: int = 4, launch_windows: int = 2048,
                 scores=(3, -5, -4), trim: bool = True, device: int = 0):
        if budget_bytes <= 0 or executors <= 0 or launch_windows <= 0:
            raise ValueError("budget, executors and launch size must be positive")
        self.lib = load_library()
        self.handle = ctypes.c_void_p()
        m, x, g = scores
        _check(self.lib.rh_server_create(device, m, x, g, int(trim), budget_bytes, executors,
                                         launch_windows, ctypes.byref(self.handle)),
               "rh_server_create")

    def submit(self, windows: Sequence[Window]) -> Request:
        arr, keep = pack_windows(windows)
        rid = ctypes.c_uint64()
        _check(self.lib.rh_server_submit(self.handle, len(windows),
                                         ctypes.cast(arr.ctypes.data, ctypes.POINTER(RhWindow)),
                                         ctypes.byref(rid)),
               "rh_server_submit")
        del keep  # the server copied the inputs before returning
        return Request(self, rid.value, windows)

    def submit_many(self, jobs: Sequence[Sequence[Window]], admit_timeout: float = 10.0
                    ) -> List[Request]:
        """Submits several requests at one instant (in list order): no launch
        starts until all of them are admitted (or admit_timeout passed, when
        the pool cannot hold them all), so they share the next launches."""
        _check(self.lib.rh_server_hold(self.handle, 1), "rh_server_hold")
        try:
            reqs = [self.submit(ws) for ws in jobs]
            info = RhRequestInfo()
            deadline = time.perf_counter() + admit_timeout
            for r in reqs:
                while time.perf_counter() < deadline:
                    _check(self.lib.rh_server_request_info(self.handle, r.rid,
                                                           ctypes.byref(info)),
                           "rh_server_request_info")
                    if info.state >= 1:
                        break
                    time.sleep(0.001)
        finally:
            _check(self.lib.rh_server_hold(self.handle, 0), "rh_server_hold")
        return reqs

    def warm(self, windows: Sequence[Window]) -> None:
        """Runs the windows once on every executor before requests arrive, so
        that device scratch is sized up front (a later growth reallocates,
        which waits for the whole device)."""
        arr, keep = pack_windows(windows)
        _check(self.lib.rh_server_warm(self.handle, len(windows),
                                       ctypes.cast(arr.ctypes.data, ctypes.POINTER(RhWindow))),
               "rh_server_warm")
        del keep

    def stats(self) -> dict:
        st = RhServerStats()
        _check(self.lib.rh_server_get_stats(self.handle, ctypes.byref(st)), "rh_server_get_stats")
        return {name: getattr(st, name) for name, _ in RhServerStats._fields_}

    def close(self) -> None:
        """Finishes every submitted request, then stops the executors."""
        if self.handle:
            self.lib.rh_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def percentile(values: Sequence[float], q: float) -> float:
    """Nearest-rank percentile (q in 0..100)."""
    if not values:
        return float("nan")
    s = sorted(values)
    k = max(0, min(len(s) - 1, int(-(-q * len(s) // 100)) - 1))
    return s[k]
