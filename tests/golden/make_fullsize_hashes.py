#!/usr/bin/env python3
"""Writes tests/golden/fullsize_sha256.json: the CPU oracle CLI's polished
FASTA hash for every full-size case of tests/fullsize_cases.py.

Run on the CPU (no GPU needed) after building oracle/ and racon_amd/csrc:

    python tests/golden/make_fullsize_hashes.py --workdir /tmp/racon_full -t 8 [case ...]

The inputs are regenerated from their seeds; only the hashes, sizes and
timings are committed. Cases already in the JSON are kept unless named.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import fullsize_cases  # noqa: E402

OUT = os.path.join(HERE, "fullsize_sha256.json")
SIMGEN = os.path.join(ROOT, "racon_amd", "bin", "simgen")
ORACLE_CLI = os.path.join(ROOT, "oracle", "_build", "racon_cpu")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workdir", default="/tmp/racon_full")
    ap.add_argument("-t", "--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("cases", nargs="*", default=list(fullsize_cases.CASES))
    args = ap.parse_args()

    table = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            table = json.load(f)
    for name in args.cases:
        files = fullsize_cases.inputs(name, args.workdir, SIMGEN)
        cmd = [ORACLE_CLI, "-t", str(args.threads)] + fullsize_cases.CASES[name]["args"] + files
        t0 = time.time()
        p = subprocess.run(cmd, capture_output=True, check=True)
        secs = time.time() - t0
        table[name] = {
            "sha256": hashlib.sha256(p.stdout).hexdigest(),
            "bytes": len(p.stdout),
            "sequences": p.stdout.count(b">"),
            "oracle_seconds": round(secs, 1),
            "oracle_threads": args.threads,
        }
        print(f"{name}: {table[name]}", flush=True)
        with open(OUT, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
            f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
