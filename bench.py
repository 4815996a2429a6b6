#!/usr/bin/env python3
"""Benchmark of racon's polishing hot path on one MI355X.

Workload (cfg2): one seeded random 5 Mbp contig with 50x
ONT-like reads (lognormal 8 +- 4 kbp, 12% errors at 0.4:0.2:0.4), PAF from
true positions, window 500, scores 3/-5/-4. The inputs are generated on the
box (racon_amd/bin/simgen, seed 1002), the product CLI aligns the overlaps on
the GPU and writes the windows, and the windows are uploaded once: a step is
one pass of the persistent POA kernel over every window, inputs resident in
HBM. Printed: one JSON line (see the keys below).
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
POA_BYTES_PER_CELL = 2  # one 16-bit traceback record written per DP cell

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
    `depth` layers. Latency is submit to last consensus; windows/s is over
    the span from the first arrival to the last completion."""
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
            launches0 = len(srv.launch_log)
            handles = []
            t0 = time.perf_counter()
            due = t0
            for ws, gap in zip(reqs, gaps):
                due += gap
                time.sleep(max(0.0, due - time.perf_counter()))
                handles.append(srv.submit(ws))
            for h in handles:
                h.wait()
            t1 = max(h.done for h in handles)
            lat = [h.latency for h in handles]
            n_win = sum(len(ws) for ws in reqs)
            log_slice = srv.launch_log[launches0:]
            row = {"rate": lam, "requests": n_req, "windows": n_win,
                   "windows_per_s": round(n_win / (t1 - t0), 1),
                   "p50_s": round(percentile(lat, 50), 4), "p99_s": round(percentile(lat, 99), 4),
                   "mean_wait_s": round(sum(h.admitted - h.arrival for h in handles) / n_req, 4),
                   "launches": len(log_slice),
                   "mixed_launches": sum(1 for _, _, k in log_slice if k > 1)}
            log(f"serving rate {lam}/s: {n_req} requests, {n_win} windows, "
                f"p50 {row['p50_s']} s, p99 {row['p99_s']} s, {row['windows_per_s']} windows/s, "
                f"{row['mixed_launches']}/{row['launches']} launches mixed")
            report["rates"].append(row)
    return report


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=sorted(CONFIGS))
    ap.add_argument("--workdir", default="/tmp/racon_amd_bench")
    ap.add_argument("--plain-windows", type=int, default=48,
                    help="windows timed in the one-launch-per-window mode")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the single-core CPU baseline sample")
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--traffic-bytes", type=float, default=0.0,
                    help="HBM bytes per POA launch measured with rocprofv3 (FETCH_SIZE x2 + "
                         "WRITE_SIZE, copy-calibrated) to report in roofline.traffic")
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

    # ---- batched persistent kernel, windows resident in HBM ----
    batch = racon_amd.PoaBatch(m, x, g, trim=True)
    batch.upload(windows)
    up = batch.stats()
    log(f"uploaded {up.input_bytes / 1e6:.1f} MB (pack {up.pack_ms:.0f} ms, "
        f"H2D {up.h2d_ms:.0f} ms)")
    # autotune the LDS ring depth (DP rows kept on chip per wave) on the
    # resident batch; 0 = the occupancy-derived default
    tuned = {}
    if not args.no_tune:
        # one untimed launch first, so the first candidate is not charged
        # with the first launch's one-time costs
        batch.run(racon_amd.RH_POA_BATCHED)
        for ring in (0, 4, 8):
            batch.set_tuning(ring_rows=ring)
            t0 = time.perf_counter()
            batch.run(racon_amd.RH_POA_BATCHED)
            tuned[ring] = (time.perf_counter() - t0) * 1e3
            log(f"tune ring_rows={ring or 'auto'}: {tuned[ring]:.1f} ms")
        best_ring = min(tuned, key=tuned.get)
        batch.set_tuning(ring_rows=best_ring)
    else:
        best_ring = 0
    for i in range(args.warmup):
        batch.run(racon_amd.RH_POA_BATCHED)
        log(f"warmup {i + 1}/{args.warmup} done")
    step_ms, kernel_ms = [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        batch.run(racon_amd.RH_POA_BATCHED)
        step_ms.append((time.perf_counter() - t0) * 1e3)
        st = batch.stats()
        kernel_ms.append(st.kernel_ms)
        log(f"step {i + 1}/{args.steps}: {step_ms[-1]:.1f} ms (kernel {st.kernel_ms:.1f} ms, "
            f"slots {st.slots}, retried {st.retried})")
    st = batch.stats()
    phases = batch.phase_seconds()
    total_phase = sum(phases.values()) or 1.0
    log("phase share of wave time: " + ", ".join(
        f"{k} {100 * v / total_phase:.1f}%" for k, v in phases.items()))
    # busy wave time over resident-slot time: below 1 when waves idle at the
    # end of the queue (load imbalance) or between windows
    busy_frac = total_phase / (st.slots * (st.kernel_ms / 1e3)) if st.slots else 0.0
    log(f"wave busy fraction {busy_frac:.3f} ({st.slots} resident waves)")
    sched = queue_report(windows, batch, st.slots, st.kernel_ms)
    gpu_out, gpu_status = batch.download()
    n_windows = len(windows)
    cells = st.cells
    ms_per_step = sum(step_ms) / len(step_ms)
    kernel_avg = sum(kernel_ms) / len(kernel_ms)
    windows_per_s = n_windows / (ms_per_step / 1e3)
    poa_gcups = cells / (kernel_avg / 1e3) / 1e9
    batch.close()

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

    # ---- CPU baseline: the oracle, single core, bounded sample ----
    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        done, t0 = 0, time.perf_counter()
        for w in windows:
            cons, status, _, _ = oracle_py.window_consensus(
                w.backbone, w.backbone_quality,
                [(l.sequence, l.quality, l.begin, l.end) for l in w.layers], w.tgs, True,
                m, x, g)
            if cons != gpu_out[done] or status != gpu_status[done]:
                raise RuntimeError(f"window {done}: GPU output differs from the oracle")
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(done / cpu_s, 3), "unit": "windows/s", "cores": 1,
               "kind": "port",
               "sample": f"first {done} windows of the same workload, oracle POA "
                         f"(scalar C++ restatement of spoa), {cpu_s:.1f} s; every sampled "
                         f"window byte-identical to the GPU output"}
        log(f"cpu: {done} windows in {cpu_s:.1f} s")

    algo_bytes = cells * POA_BYTES_PER_CELL + st.input_bytes
    achieved_gbs = algo_bytes / (kernel_avg / 1e3) / 1e9
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
                   "window_length": 500, "scores": [m, x, g], "seed": CONFIGS[args.config]["seed"]},
        "poa_gcups": round(poa_gcups, 3),
        "poa_cells": int(cells),
        "kernel_ms": round(kernel_avg, 3),
        "plain_windows_per_s": round(plain_wps, 3),
        "batched_over_plain": round(windows_per_s / plain_wps, 2),
        "align_gcups_full_matrix": round(align_gcups, 2),
        "align_kernel_ms": round(ast.kernel_ms, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
                     "traffic": args.traffic_bytes or None},
        "roofline_valu": {"bound": "valu-int32",
                          "achieved": round(cells * POA_OPS_PER_CELL / (kernel_avg / 1e3) / 1e12, 4),
                          "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                          "frac": round(cells * POA_OPS_PER_CELL / (kernel_avg / 1e3) / 1e12
                                        / VALU_PEAK_TOPS, 5)},
        "cpu_baseline": cpu,
        "poa_phase_share": {k: round(v / total_phase, 4) for k, v in phases.items()},
        "poa_wave_busy_frac": round(busy_frac, 4),
        "poa_queue": sched,
        "poa_resident_waves": st.slots,
        "autotune_ms": {str(k): round(v, 1) for k, v in tuned.items()},
        "ring_rows": best_ring,
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
