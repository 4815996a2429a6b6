"""racon --serve: whole polishing jobs (reads + overlaps + contigs) served
together on one device. Each job's output file must hold exactly the bytes
the product CLI prints for that job alone (which the other GPU tests compare
with the CPU oracle), whatever else shares the device.
"""
import json
import os
import subprocess

import pytest

from racon_amd import _lib

pytestmark = pytest.mark.gpu

BIN = os.path.join(_lib.PKG_DIR, "bin")
OPTS = ["-t", "8", "-w", "500", "-m", "3", "-x", "-5", "-g", "-4", "-q", "10", "-e", "0.3"]

JOBS = [  # (seed, contig bp, depth)
    (1061, 20000, 20),
    (1062, 150000, 50),
    (1063, 600000, 90),
]


def make_job(workdir, seed, contig, depth):
    prefix = os.path.join(workdir, f"job{seed}")
    subprocess.run([os.path.join(BIN, "simgen"), "contig", str(seed), prefix, str(contig),
                    str(depth), "8000", "4000", "0.12", "0.4:0.2:0.4"], check=True,
                   stdout=subprocess.DEVNULL)
    return [prefix + ".reads.fastq", prefix + ".paf", prefix + ".contigs.fasta"]


@pytest.mark.timeout(600)
def test_served_jobs_match_the_cli_alone(tmp_path):
    work = str(tmp_path)
    inputs = [make_job(work, *j) for j in JOBS]
    alone = []
    for files in inputs:
        out = subprocess.run([os.path.join(BIN, "racon")] + OPTS + files, check=True,
                             stdout=subprocess.PIPE, stderr=subprocess.DEVNULL).stdout
        assert out.startswith(b">")
        alone.append(out)
    print("alone:", [len(a) for a in alone], flush=True)
    # the big job first, the two small ones while it runs, then the middle
    # one again, arriving later
    lines, expect = [], {}
    for k, (files, arrival) in enumerate(zip(inputs + [inputs[1]], [0, 0.05, 0.05, 0.3])):
        path = os.path.join(work, f"served{k}.fasta")
        lines.append(" ".join([path] + files + [str(arrival)]))
        expect[path] = alone[k if k < 3 else 1]
    order = [2, 0, 1, 3]
    jobs = os.path.join(work, "jobs.txt")
    with open(jobs, "w") as f:
        f.write("# output sequences overlaps targets arrival\n")
        f.write("\n".join(lines[i] for i in order) + "\n")
    res = subprocess.run([os.path.join(BIN, "racon")] + OPTS +
                         ["--serve", jobs, "--serve-budget", "4", "--serve-executors", "3",
                          "--serve-jobs", "4"],
                         check=True, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    report = [json.loads(l) for l in res.stdout.decode().splitlines() if l.startswith("{")]
    for r in report:
        print(r, flush=True)
    per_job = [r for r in report if "job" in r]
    assert len(per_job) == 4
    for r in per_job:
        assert r["end_s"] >= r["start_s"] >= r["arrival_s"] - 1e-3
        assert r["windows"] > 0
    for path, want in expect.items():
        with open(path, "rb") as f:
            got = f.read()
        assert got == want, path
