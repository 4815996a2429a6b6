#!/usr/bin/env python3
"""Benchmark of racon's polishing hot path on one MI355X.

Workload (cfg2): one seeded random 5 Mbp contig with 50x
ONT-like reads (lognormal 8 +- 4 kbp, 12% errors at 0.4:0.2:0.4), PAF from
true positions, window 500, scores 3/-5/-4. The inputs are generated on the
box (racon_amd/bin/simgen, seed 1002), and the product CLI aligns the
overlaps on the GPU and writes the windows. A step is racon's polish phase
over every window through the C ABI: rh_poa_upload (host pack + H2D),
rh_poa_run (the persistent POA kernel) and rh_poa_download (D2H). Steps
alternate between two batches (step k+1's transfers beside step k's
kernels, as the CLI's chunks run), all inside the clock; --sequential times
them one after the other on one batch. The kernel-only time (HIP events on
the batch's stream) is reported beside it.
Also reported: the static-grid comparator (one workgroup per window, one
launch), the plain mode (one launch per window), the overlap aligner, and
the CPU baseline (racon's CPU path restated, 1 and T threads). Printed: one
JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "polished windows/s + POA GCUPS, batched vs plain, 1 MI355X, bit-exact"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 39.3  # 256 CUs x 64 lanes x 2.4 GHz int32
POA_OPS_PER_CELL = 6    # linear gap: diag add, vertical add, 2 max, horizontal add + max

CONFIGS = {
    1: dict(name="cfg1: 50 kbp contig, 30x ONT 12%, w500", seed=1001,
            sim=["contig", "{seed}", "{prefix}", "50000", "30", "8000", "4000", "0.12",
                 "0.4:0.2:0.4"],
            racon=["-w", "500", "-m", "3", "-x", "-5", "-g", "-4", "-q", "10", "-e", "0.3"],
            files=("reads.fastq", "paf", "contigs.fasta")),
    2: dict(name="cfg2: 5 Mbp contig, 50x ONT 12%, w500 (~10k windows x ~50 layers)",
            seed=1002,
            sim=["contig", "{seed}", "{prefix}", "5000000", "50", "8000", "4000", "0.12",
                 "0.4:0.2:0.4"],
            racon=["-w", "500", "-m", "3", "-x", "-5", "-g", "-4", "-q", "10", "-e", "0.3"],
            files=("reads.fastq", "paf", "contigs.fasta")),
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)

def queue_report(windows, batch, slots, kernel_ms):
    """How well the work queue's cost order balances the resident waves.

    Replays the queue (windows in the library's order: estimated cost
    (backbone + layer bytes) x (longest layer + 1), largest first, each to the
    first free wave) with the measured per-window device times, and compares
    its makespan with the same replay in measured-time order and with the
    perfect split (total / waves)."""
    import heapq
    cells, ticks = batch.window_costs()
    rate = batch.tick_rate()
    est = []
    for w in windows:
        n_bound = len(w.backbone) + sum(len(l.sequence) for l in w.layers)
        longest = max((len(l.sequence) for l in w.layers), default=0)
        est.append(n_bound * (longest + 1))
    ran = [i for i, t in enumerate(ticks) if t > 0]
    if not ran or not slots:
        return {}
    secs = [ticks[i] / rate for i in range(len(ticks))]

    def replay(order):
        heap = [0.0] * slots
        for i in order:
            t = heapq.heappop(heap)
            heapq.heappush(heap, t + secs[i])
        return max(heap)

    by_est = sorted(ran, key=lambda i: -est[i])
    by_time = sorted(ran, key=lambda i: -secs[i])
    total = sum(secs[i] for i in ran)
    srt = sorted(secs[i] for i in ran)
    rep = {
        "kernel_s": round(kernel_ms / 1e3, 4),
        "replay_cost_order_s": round(replay(by_est), 4),
        "replay_time_order_s": round(replay(by_time), 4),
        "perfect_split_s": round(total / slots, 4),
        "window_s_mean": round(total / len(ran), 4),
        "window_s_p99": round(srt[int(0.99 * (len(srt) - 1))], 4),
        "window_s_max": round(srt[-1], 4),
    }
    log("queue: " + ", ".join(f"{k} {v}" for k, v in rep.items()))
    for i in by_time[:5]:
        w = windows[i]
        lens = [len(l.sequence) for l in w.layers]
        log(f"  heavy window {i}: {secs[i]:.3f} s, {cells[i] / 1e6:.1f} M cells, "
            f"{len(lens)} layers, backbone {len(w.backbone)}, layer mean "
            f"{sum(lens) / max(1, len(lens)):.0f} max {max(lens, default=0)}")
    return rep



def prepare(cfg_id: int, workdir: str, threads: int):
    """Generates the config's inputs and dumps windows/overlaps with the CLI."""
    cfg = CONFIGS[cfg_id]
    os.makedirs(workdir, exist_ok=True)
    prefix = os.path.join(workdir, f"cfg{cfg_id}")
    sim = os.path.join(ROOT, "racon_amd", "bin", "simgen")
    racon = os.path.join(ROOT, "racon_amd", "bin", "racon")
    reads, paf, contigs = (f"{prefix}.{x}" for x in cfg["files"])
    if not os.path.exists(reads):
        args = [a.format(seed=cfg["seed"], prefix=prefix) for a in cfg["sim"]]
        t0 = time.time()
        subprocess.run([sim] + args, check=True)
        log(f"generated {cfg['name']} in {time.time() - t0:.1f} s")
    wdump, odump = f"{prefix}.windows.bin", f"{prefix}.overlaps.bin"
    report = f"{prefix}.report.json"
    if not os.path.exists(wdump):
        t0 = time.time()
        subprocess.run([racon, "-t", str(threads)] + cfg["racon"] +
                       ["--dump-windows", wdump, "--dump-overlaps", odump, "--report", report,
                        reads, paf, contigs], check=True, stdout=subprocess.DEVNULL)
        log(f"racon initialize (GPU overlap alignment) + dump in {time.time() - t0:.1f} s")
    return wdump, odump, (reads, paf, contigs)


def serving_bench(windows, args, scores) -> dict:
    """Poisson arrivals of polishing requests into the serving loop.

    A request stands for a contig of log-uniform 20 kbp - 2 Mbp polished at
    a depth uniform in 20-100: that many consecutive windows of the loaded
    config (window length 500), each keeping a random subset of at most
    `depth` layers. Latency is submit to last consensus (and, "from
    schedule", from the Poisson arrival time: under overload the submitting
    loop falls behind it); windows/s is over the span from the first arrival
    to the last completion."""
    import math
    import random

    from racon_amd import Window
    from racon_amd.serving import PolishServer, percentile

    rng = random.Random(1005)

    def make_request():
        contig = math.exp(rng.uniform(math.log(20e3), math.log(2e6)))
        n_w = max(1, math.ceil(contig / 500))
        start = rng.randrange(len(windows))
        depth = rng.randint(20, 100)
        out = []
        for i in range(n_w):
            w = windows[(start + i) % len(windows)]
            keep = sorted(rng.sample(range(len(w.layers)), min(depth, len(w.layers))))
            out.append(Window(w.backbone, w.backbone_quality, [w.layers[k] for k in keep], w.tgs))
        return out

    budget = int(args.serving_budget * (1 << 30))
    rates = [float(r) for r in args.serving.split(",") if r]
    report = {"metric": "serving: p50/p99 request latency and windows/s under Poisson arrivals",
              "unit": "s", "n_gpus": 1, "data": "synthetic",
              "config": {"workload": f"requests cut from cfg{args.config} windows "
                                     "(contig 20 kbp-2 Mbp log-uniform, depth 20-100)",
                         "budget_bytes": budget, "executors": args.serving_executors,
                         "launch_windows": args.serving_launch},
              "rates": []}
    with PolishServer(budget, executors=args.serving_executors,
                      launch_windows=args.serving_launch, scores=scores) as srv:
        # every executor sizes its scratch on a full launch of the largest windows
        size = [len(w.backbone) + sum(len(l.sequence) for l in w.layers) for w in windows]
        heavy = sorted(range(len(windows)), key=lambda i: -size[i])[:args.serving_launch]
        srv.warm([windows[i] for i in heavy])
        for lam in rates:
            n_req = max(6, int(lam * args.serving_seconds))
            reqs = [make_request() for _ in range(n_req)]
            gaps = [rng.expovariate(lam) for _ in range(n_req)]
            st0 = srv.stats()
            handles = []
            lags = []  # submit call after its scheduled arrival (client behind)
            t0 = time.perf_counter()
            due = t0
            for ws, gap in zip(reqs, gaps):
                due += gap
                time.sleep(max(0.0, due - time.perf_counter()))
                lags.append(max(0.0, time.perf_counter() - due))
                handles.append(srv.submit(ws))
            for h in handles:
                h.wait()
            st1 = srv.stats()
            first = min(h.arrival for h in handles)
            span = max(h.done for h in handles) - first
            lat = [h.latency for h in handles]
            lat_sched = [h.latency + g for h, g in zip(handles, lags)]
            n_win = sum(len(ws) for ws in reqs)
            row = {"rate": lam, "requests": n_req, "windows": n_win,
                   "windows_per_s": round(n_win / span, 1),
                   "p50_s": round(percentile(lat, 50), 4), "p99_s": round(percentile(lat, 99), 4),
                   "p50_from_schedule_s": round(percentile(lat_sched, 50), 4),
                   "p99_from_schedule_s": round(percentile(lat_sched, 99), 4),
                   "mean_wait_s": round(sum(h.admitted - h.arrival for h in handles) / n_req, 4),
                   "launches": st1["launches"] - st0["launches"],
                   "mixed_launches": st1["shared_launches"] - st0["shared_launches"],
                   "kernel_s": round((st1["kernel_ms"] - st0["kernel_ms"]) / 1e3, 3)}
            log(f"serving rate {lam}/s: {n_req} requests, {n_win} windows, "
                f"p50 {row['p50_s']} s, p99 {row['p99_s']} s, {row['windows_per_s']} windows/s, "
                f"{row['mixed_launches']}/{row['launches']} launches mixed")
            report["rates"].append(row)
        report["server"] = {k: srv.stats()[k] for k in ("pool_bytes", "scratch_bytes",
                                                         "window_capacity")}
    return report


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(windows, gpu_out, gpu_status, threads, per_thread, reps, scores):
    """racon's CPU path, restated (the oracle's scalar C++ POA, one window per
    task as Polisher::polish does, src/polisher.cpp:486-504), timed on this
    host with 1 and `threads` worker threads over the first windows of the
    same workload; median of `reps` runs each, every sampled window checked
    byte-identical to the GPU output."""
    import concurrent.futures
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    m, x, g = scores

    def one(i):
        w = windows[i]
        cons, status, _, _ = oracle_py.window_consensus(
            w.backbone, w.backbone_quality,
            [(l.sequence, l.quality, l.begin, l.end) for l in w.layers], w.tgs, True, m, x, g)
        if cons != gpu_out[i] or status != gpu_status[i]:
            raise RuntimeError(f"window {i}: GPU output differs from the oracle")
        return i

    legs = {}
    for t in sorted({1, threads}):
        n = min(len(windows), per_thread * t)
        times = []
        for r in range(reps):
            t0 = time.perf_counter()
            with concurrent.futures.ThreadPoolExecutor(t) as ex:
                list(ex.map(one, range(n)))
            times.append(time.perf_counter() - t0)
            log(f"cpu -t {t}: {n} windows in {times[-1]:.2f} s (run {r + 1}/{reps})")
        med = statistics.median(times)
        legs[t] = {"windows": n, "seconds": [round(s, 3) for s in times],
                   "median_s": round(med, 3), "windows_per_s": round(n / med, 3)}
    top = legs[threads]
    return {"value": top["windows_per_s"], "unit": "windows/s", "cores": threads,
            "kind": "port",
            "sample": (f"first {top['windows']} of {len(windows)} windows of the same workload "
                       f"(-t {threads}) and first {legs[1]['windows']} (-t 1); racon CPU path, "
                       f"restated: the oracle's scalar C++ POA (upstream spoa would use SIMD), "
                       f"one window per task, median of {reps} runs; every sampled window "
                       f"byte-identical to the GPU output"),
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "t1": legs[1], f"t{threads}": top}


def profile_traffic(path):
    """HBM bytes per POA launch from the committed rocprofv3 summary (line
    'traffic_bytes_per_launch: N'), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=sorted(CONFIGS))
    ap.add_argument("--workdir", default="/tmp/racon_amd_bench")
    ap.add_argument("--plain-windows", type=int, default=48,
                    help="windows timed in the one-launch-per-window mode")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="worker threads of the CPU baseline (0: the host's CPUs, at most 16)")
    ap.add_argument("--cpu-windows-per-thread", type=int, default=240,
                    help="CPU baseline sample: windows per worker thread")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-static", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--skip-banded", action="store_true",
                    help="skip the banded (-b) comparison run")
    ap.add_argument("--skip-serving-probe", action="store_true",
                    help="skip the 4-second serving-loop overload run")
    ap.add_argument("--sequential", action="store_true",
                    help="polish-phase steps one after the other on one batch (default: two "
                         "batches, each step's transfers beside the other step's kernels)")
    ap.add_argument("--kernel-only", action="store_true",
                    help="profiling runs: time only rh_poa_run (inputs resident), skip the "
                         "comparators and the CPU baseline")
    ap.add_argument("--align-only", action="store_true",
                    help="profiling runs: only the overlap aligner, `steps` launches over the "
                         "config's overlaps (inputs resident)")
    ap.add_argument("--profile", default=os.path.join(ROOT, "profiles", "poa_cfg2_counters.txt"),
                    help="committed rocprofv3 summary with the POA kernel's HBM bytes per launch")
    ap.add_argument("--serving", default="", metavar="RATES",
                    help="only run the serving loop: Poisson request arrivals at these rates "
                         "(requests/s, comma separated), requests cut from the config's windows")
    ap.add_argument("--serving-seconds", type=float, default=8.0,
                    help="arrival span per rate (at least 6 requests per rate)")
    ap.add_argument("--serving-budget", type=float, default=8.0,
                    help="admission budget in GiB of window bytes")
    ap.add_argument("--serving-executors", type=int, default=4,
                    help="concurrent launches (one POA batch and HIP stream each)")
    ap.add_argument("--serving-launch", type=int, default=2048,
                    help="most windows per launch")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (torch), to scale "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE readings, then exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; windows would be "
                                   "sharded across devices (independent windows, host-side "
                                   "gather), but no multi-GPU figure is produced", "n_gpus": args.gpus}))
        return 2

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "reads": n * 4, "writes": n * 4}))
        return 0

    import racon_amd
    from racon_amd._lib import pack_windows
    from racon_amd.dumps import read_overlaps, read_windows
    lib = racon_amd.load_library()
    if lib.rh_device_count() < 1:
        raise RuntimeError("no HIP device visible")

    wdump, odump, _ = prepare(args.config, args.workdir, args.threads)
    t0 = time.time()
    windows = read_windows(wdump)
    pairs = read_overlaps(odump)
    log(f"loaded {len(windows)} windows, {len(pairs)} overlaps in {time.time() - t0:.1f} s")
    m, x, g = 3, -5, -4
    if args.serving:
        print(json.dumps(serving_bench(windows, args, (m, x, g))), flush=True)
        return 0
    if args.align_only:
        ab = racon_amd.AlignBatch()
        ab.upload(pairs)
        times = []
        for i in range(args.warmup + args.steps):
            ab.run()
            if i >= args.warmup:
                times.append(ab.stats().kernel_ms)
        ast = ab.stats()
        ab.close()
        ms = sum(times) / len(times)
        log(f"align: {len(pairs)} overlaps, kernel {ms:.1f} ms per launch ({times})")
        print(json.dumps({"metric": "overlap alignment kernel", "unit": "ms", "kernel_ms": ms,
                          "pairs": len(pairs), "cells_full_matrix": int(ast.cells),
                          "wave_steps": int(ast.steps),
                          "gcups_full_matrix": round(ast.cells / (ms / 1e3) / 1e9, 1)}),
              flush=True)
        return 0

    # the caller's rh_window array (racon's Window objects), built once
    n_windows = len(windows)
    arr, keep = pack_windows(windows)

    # ---- the library's autotuner: LDS ring depth (and so waves per CU) per bin ----
    batch = racon_amd.PoaBatch(m, x, g, trim=True)
    batch.upload_packed(arr, n_windows, keep)
    tuned = None
    if not args.no_tune:
        batch.run(racon_amd.RH_POA_BATCHED)  # first launch: one-time costs
        tuned = batch.autotune()
        log(f"autotune: {tuned}")

    # ---- the polish phase: pack + H2D + kernels + D2H per step ----
    bufs = batch.output_buffers(2 * batch.lib.rh_poa_output_bound(batch.handle))
    # steps alternate between two batches, as the CLI's chunks do: step k + 1
    # is packed and copied while step k's kernels run, and step k's results
    # come back while step k + 1's run; the clock covers every step's pack,
    # H2D, kernels and D2H, the first upload and the last download included
    batches = [batch]
    if not args.kernel_only and not args.sequential:
        second = racon_amd.PoaBatch(m, x, g, trim=True)
        second.upload_packed(arr, n_windows, keep)
        if tuned is not None:
            second.run(racon_amd.RH_POA_BATCHED)
            log(f"autotune (second batch): {second.autotune()}")
        batches.append(second)
    all_bufs = [bufs] + [b.output_buffers(2 * b.lib.rh_poa_output_bound(b.handle))
                         for b in batches[1:]]
    step_ms, parts = [], []
    last_bufs = bufs
    seq_ms = []

    def record(b, i, ms):
        st = b.stats()
        parts.append((st.pack_ms, st.h2d_ms, st.kernel_ms, st.d2h_ms))
        log(f"step {i + 1}/{args.steps}: {ms:.1f} ms (pack {st.pack_ms:.1f}, H2D "
            f"{st.h2d_ms:.1f}, kernel {st.kernel_ms:.1f}, D2H {st.d2h_ms:.1f} ms; slots "
            f"{st.slots}, retried {st.retried})")

    def sequential_steps(n, timed):
        for i in range(n):
            t0 = time.perf_counter()
            if args.kernel_only:
                batch.run(racon_amd.RH_POA_BATCHED)
            else:
                batch.upload_packed(arr, n_windows, keep)
                batch.run(racon_amd.RH_POA_BATCHED)
                batch.download_into(bufs)
            if timed:
                step_ms.append((time.perf_counter() - t0) * 1e3)
                record(batch, i, step_ms[-1])

    def pipelined_steps(n, timed):
        t0 = time.perf_counter()
        batches[0].upload_packed(arr, n_windows, keep)
        batches[0].submit(racon_amd.RH_POA_BATCHED)
        for k in range(n):
            cur, nxt = batches[k % 2], batches[(k + 1) % 2]
            if k + 1 < n:
                nxt.upload_packed(arr, n_windows, keep)  # beside step k's kernels
            cur.wait()
            if k + 1 < n:
                nxt.submit(racon_amd.RH_POA_BATCHED)
            cur.download_into(all_bufs[k % 2])  # beside step k + 1's kernels
            if timed:
                record(cur, k, (time.perf_counter() - t0) * 1e3 / (k + 1))
        if timed:
            step_ms.extend([(time.perf_counter() - t0) * 1e3 / n] * n)
        return all_bufs[(n - 1) % 2]

    steps = sequential_steps if len(batches) == 1 else pipelined_steps
    for i in range(args.warmup):
        steps(len(batches), False)  # (every batch once)
        log(f"warmup {i + 1}/{args.warmup} done")
    out_bufs = steps(args.steps, True)
    if out_bufs is not None:
        last_bufs = out_bufs
    if len(batches) > 1:
        # the same steps one after the other, for the record
        for _ in range(2):
            t0 = time.perf_counter()
            batch.upload_packed(arr, n_windows, keep)
            batch.run(racon_amd.RH_POA_BATCHED)
            batch.download_into(bufs)
            seq_ms.append((time.perf_counter() - t0) * 1e3)
        log(f"sequential steps: {seq_ms}")
        for b in batches[1:]:
            b.close()
    st = batch.stats()
    phases = batch.phase_seconds()
    total_phase = sum(phases.values()) or 1.0
    log("phase share of wave time: " + ", ".join(
        f"{k} {100 * v / total_phase:.1f}%" for k, v in phases.items()))
    busy_frac = total_phase / (st.slots * (st.kernel_ms / 1e3)) if st.slots else 0.0
    log(f"wave busy fraction {busy_frac:.3f} ({st.slots} resident waves)")
    sched = queue_report(windows, batch, st.slots, st.kernel_ms)
    if args.kernel_only:
        batch.download_into(bufs)
        last_bufs = bufs
    gpu_out, gpu_status = batch.split(last_bufs, n_windows)
    cells = st.cells
    ms_per_step = sum(step_ms) / len(step_ms)
    kernel_avg = sum(p[2] for p in parts) / len(parts)
    windows_per_s = n_windows / (ms_per_step / 1e3)
    kernel_wps = n_windows / (kernel_avg / 1e3)
    poa_gcups = cells / (kernel_avg / 1e3) / 1e9
    input_bytes = st.input_bytes
    batch.close()

    extra = {}
    if not args.kernel_only:
        # ---- static grid: one workgroup per window, one launch, no queue ----
        if not args.skip_static:
            sb = racon_amd.PoaBatch(m, x, g, trim=True)
            sb.upload_packed(arr, n_windows, keep)
            sb.run(racon_amd.RH_POA_STATIC)  # sizes the slots
            sms = []
            for _ in range(2):
                sb.run(racon_amd.RH_POA_STATIC)
                sms.append(sb.stats().kernel_ms)
            sst = sb.stats()
            sout, _ = sb.download()
            sb.close()
            if sout != gpu_out:
                raise RuntimeError("static-grid and batched outputs differ")
            static_ms = min(sms)
            extra["static_kernel_ms"] = round(static_ms, 3)
            extra["static_launches"] = int(sst.launches)
            extra["static_windows_per_s_kernel"] = round(n_windows / (static_ms / 1e3), 2)
            extra["batched_over_static_kernel"] = round(static_ms / kernel_avg, 3)
            log(f"static grid: kernel {static_ms:.1f} ms ({sst.launches} launches) vs batched "
                f"{kernel_avg:.1f} ms")

        # ---- racon -b: the banded approximation (not bit-exact; reported beside) ----
        if not args.skip_banded:
            bb = racon_amd.PoaBatch(m, x, g, trim=True)
            bb.set_band(256)
            bb.upload_packed(arr, n_windows, keep)
            bb.run(racon_amd.RH_POA_BATCHED)  # sizes the slots
            bms = []
            for _ in range(2):
                bb.run(racon_amd.RH_POA_BATCHED)
                bms.append(bb.stats().kernel_ms)
            bst = bb.stats()
            bout, _ = bb.download()
            bb.close()
            band_ms = min(bms)
            same = sum(1 for a, c in zip(bout, gpu_out) if a == c)
            extra["banded"] = {
                "band": 256, "kernel_ms": round(band_ms, 3),
                "windows_per_s_kernel": round(n_windows / (band_ms / 1e3), 2),
                "speedup_over_exact_kernel": round(kernel_avg / band_ms, 3),
                "cells": int(bst.cells), "band_reruns": int(bst.band_reruns),
                "windows_equal_exact": same, "windows": n_windows}
            log(f"banded (-b): kernel {band_ms:.1f} ms vs exact {kernel_avg:.1f} ms, "
                f"{same}/{n_windows} windows equal, {bst.band_reruns} rerun exactly")

        # ---- plain: one launch per window, same kernel body ----
        subset = windows[:args.plain_windows]
        plain = racon_amd.PoaBatch(m, x, g, trim=True)
        plain.upload(subset)
        plain.run(racon_amd.RH_POA_PLAIN)  # warm
        t0 = time.perf_counter()
        plain.run(racon_amd.RH_POA_PLAIN)
        plain_s = time.perf_counter() - t0
        plain_out, _ = plain.download()
        plain.close()
        plain_wps = len(subset) / plain_s
        if plain_out != gpu_out[:len(subset)]:
            raise RuntimeError("plain and batched outputs differ")
        log(f"plain: {len(subset)} windows in {plain_s:.2f} s -> {plain_wps:.1f} windows/s")
        extra["plain_windows_per_s"] = round(plain_wps, 3)
        extra["batched_over_plain"] = round(windows_per_s / plain_wps, 2)

        # ---- serving loop under overload (Poisson arrivals, 64 requests/s:
        # about twice what it can serve, so windows/s is its capacity) ----
        if not args.skip_serving_probe:
            probe = argparse.Namespace(**vars(args))
            probe.serving = "64"
            probe.serving_seconds = 4.0
            rep = serving_bench(windows, probe, (m, x, g))
            row = rep["rates"][0]
            extra["serving_overload"] = {
                k: row[k] for k in ("rate", "requests", "windows", "windows_per_s", "p50_s",
                                    "p99_s", "p50_from_schedule_s", "p99_from_schedule_s",
                                    "launches", "mixed_launches")}
            extra["serving_overload"]["over_kernel_only"] = round(
                row["windows_per_s"] / kernel_wps, 3)
            extra["serving_overload"]["executors"] = args.serving_executors
            extra["serving_overload"]["launch_windows"] = args.serving_launch

        # ---- overlap alignment kernel ----
        ab = racon_amd.AlignBatch()
        ab.upload(pairs)
        ab.run()  # warm
        ab.run()
        ast = ab.stats()
        ab.close()
        align_gcups = ast.cells / (ast.kernel_ms / 1e3) / 1e9
        log(f"align: {len(pairs)} overlaps, kernel {ast.kernel_ms:.1f} ms, "
            f"{align_gcups:.1f} GCUPS (full-matrix equivalent)")
        extra["align_kernel_ms"] = round(ast.kernel_ms, 3)
        extra["align_gcups_full_matrix"] = round(align_gcups, 2)
        extra["align_cells_full_matrix"] = int(ast.cells)
        # cells the banded passes computed: at most 64 blocks x 64 rows per wave step
        extra["align_wave_steps"] = int(ast.steps)
        extra["align_cells_computed_max"] = int(ast.steps) * 4096
        extra["align_roofline_valu_frac"] = round(
            ast.cells * 20 / 32 / (ast.kernel_ms / 1e3) / 1e12 / VALU_PEAK_TOPS, 5)

    # ---- CPU baseline: racon's CPU path restated, 1 and T threads ----
    cpu = None
    if not args.skip_cpu and not args.kernel_only:
        try:
            avail = len(os.sched_getaffinity(0))
        except AttributeError:
            avail = os.cpu_count() or 1
        threads = args.cpu_threads or max(1, min(16, avail))
        cpu = cpu_baseline(windows, gpu_out, gpu_status, threads, args.cpu_windows_per_thread,
                           args.cpu_reps, (m, x, g))

    valu_tops = cells * POA_OPS_PER_CELL / (kernel_avg / 1e3) / 1e12
    traffic = profile_traffic(args.profile)
    pack_ms, h2d_ms, _, d2h_ms = (sum(p[k] for p in parts) / len(parts) for k in range(4))
    result = {
        "metric": METRIC,
        "value": round(windows_per_s, 2),
        "unit": "windows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": CONFIGS[args.config]["name"], "windows": n_windows,
                   "gpu_windows": int(st.gpu_windows), "overlaps": len(pairs),
                   "window_length": 500, "scores": [m, x, g], "seed": CONFIGS[args.config]["seed"],
                   "step": ("rh_poa_run only (inputs resident)" if args.kernel_only else
                            ("polish phase, one step after the other: " if args.sequential
                             else "polish phase, steps pipelined over two batches (step k+1's "
                             "pack + H2D and step k's D2H beside the kernels; all inside the "
                             "clock): ") +
                            "rh_poa_upload (pack + H2D) + rh_poa_run (kernels) + "
                            "rh_poa_download (D2H)")},
        "step_ms_parts": {"pack": round(pack_ms, 3), "h2d": round(h2d_ms, 3),
                          "kernel": round(kernel_avg, 3), "d2h": round(d2h_ms, 3)},
        "sequential_ms_per_step": round(min(seq_ms), 3) if seq_ms else None,
        "kernel_ms": round(kernel_avg, 3),
        "kernel_windows_per_s": round(kernel_wps, 2),
        "poa_gcups": round(poa_gcups, 3),
        "poa_cells": int(cells),
        "input_bytes": int(input_bytes),
        "roofline": {"bound": "valu", "achieved": round(valu_tops, 4), "peak": VALU_PEAK_TOPS,
                     "unit": "TOP/s", "frac": round(valu_tops / VALU_PEAK_TOPS, 5),
                     "traffic": traffic,
                     "ops_per_cell": POA_OPS_PER_CELL},
        "roofline_hbm": ({"achieved": round(traffic / (kernel_avg / 1e3) / 1e9, 2),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(traffic / (kernel_avg / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
                          "traffic_source": os.path.relpath(args.profile, ROOT)}
                         if traffic else None),
        "cpu_baseline": cpu,
        "poa_phase_share": {k: round(v / total_phase, 4) for k, v in phases.items()},
        "poa_wave_busy_frac": round(busy_frac, 4),
        "poa_queue": sched,
        "poa_resident_waves": st.slots,
        "autotune": tuned,
    }
    result.update(extra)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
