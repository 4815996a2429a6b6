"""The benchmark workloads at full size, as parity cases.

Each case is a seeded simgen command (racon_amd/bin/simgen: splitmix64, the
same bytes on any machine) and racon's options. The CPU oracle CLI's output
hash for every case is committed in tests/golden/fullsize_sha256.json
(written by tests/golden/make_fullsize_hashes.py in a CPU run); the GPU test
(tests/test_gpu_fullsize.py) regenerates the inputs and compares the product
CLI's hash with it.

- cfg2: 5 Mbp contig, 50x ONT-like 12% reads (lognormal 8 +- 4 kbp), w500;
- cfg3: contigs of 200 kbp, 1 Mbp and 4 Mbp, gamma-distributed depth per
  10 kbp region (mean 40x, clipped to 5-200x), half ONT-like 15% and half
  CLR-like 13% insertion-heavy reads of 1-40 kbp, windows of 500 and 1000;
- cfg4: fragment correction (-f) of 20k reads (lognormal 10 +- 5 kbp, 12%)
  from a 10 Mbp random genome, all-vs-all dual PAF, at most 40 overlaps per
  read.
"""
import os

SCORES = ["-m", "3", "-x", "-5", "-g", "-4", "-q", "10", "-e", "0.3"]

CASES = {
    "cfg2": dict(
        sim=["contig", "1002", "{prefix}", "5000000", "50", "8000", "4000", "0.12",
             "0.4:0.2:0.4"],
        args=["-w", "500"] + SCORES,
        files=("reads.fastq", "paf", "contigs.fasta")),
    "cfg3_w500": dict(
        sim=["contig", "1003", "{prefix}", "200000,1000000,4000000", "40", "8000", "6000",
             "0.15", "0.4:0.2:0.4", "--depth-gamma", "--min-len", "1000", "--max-len", "40000",
             "--second", "0.5", "0.13", "0.2:0.5:0.3"],
        args=["-w", "500"] + SCORES,
        files=("reads.fastq", "paf", "contigs.fasta")),
    "cfg3_w1000": dict(
        sim="cfg3_w500",  # same inputs
        args=["-w", "1000"] + SCORES,
        files=("reads.fastq", "paf", "contigs.fasta")),
    "cfg4": dict(
        sim=["frag", "1004", "{prefix}", "10000000", "20000", "10000", "5000", "0.12",
             "--max-overlaps", "40"],
        args=["-f", "-w", "500"] + SCORES,
        files=("reads.fastq", "ava.paf", "reads.fastq")),
}


def inputs(name, workdir, simgen):
    """Generates (once) the inputs of case `name` under workdir and returns
    racon's three positional file arguments."""
    import subprocess

    case = CASES[name]
    source = name
    while isinstance(CASES[source]["sim"], str):
        source = CASES[source]["sim"]
    prefix = os.path.join(workdir, source)
    files = [f"{prefix}.{ext}" for ext in case["files"]]
    if not all(os.path.exists(f) for f in files):
        os.makedirs(workdir, exist_ok=True)
        args = [a.format(prefix=prefix) for a in CASES[source]["sim"]]
        subprocess.run([simgen] + args, check=True)
    return files
