"""The benchmark workloads as parity cases: the GPU racon CLI vs the CPU oracle CLI,
byte for byte, on seeded synthetic inputs (racon_amd/bin/simgen).

- cfg1 at full size (50 kbp contig, 30x ONT 12%, w500, -t 1 semantics);
- cfg3 scaled to a few hundred kbp: 3 contigs, gamma-distributed depth
  (5-200x), ONT 15% mixed with CLR-like 13% insertion-heavy reads, 1-40 kbp,
  windows of 500 and 1000;
- cfg4 scaled: fragment correction (-f) of reads from a random genome with
  all-vs-all dual overlaps.
Full-size cfg2 is checked in bench.py (sampled windows vs the oracle).
"""
import os
import subprocess

import pytest

from racon_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIMGEN = os.path.join(ROOT, "racon_amd", "bin", "simgen")
ORACLE_CLI = os.path.join(ROOT, "oracle", "_build", "racon_cpu")


def run(binary, args, threads=16):
    p = subprocess.run([binary, "-t", str(threads)] + args, capture_output=True, timeout=900)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    return p.stdout


def compare(args, files):
    gpu = run(_lib.RACON_BIN, args + files)
    cpu = run(ORACLE_CLI, args + files)
    n = gpu.count(b">")
    print(f"{' '.join(args)}: {n} sequences, gpu {len(gpu)} B, oracle {len(cpu)} B", flush=True)
    assert n > 0
    assert gpu == cpu


def test_cfg1_full(tmp_path):
    prefix = str(tmp_path / "cfg1")
    subprocess.run([SIMGEN, "contig", "1001", prefix, "50000", "30", "8000", "4000", "0.12",
                    "0.4:0.2:0.4"], check=True)
    compare(["-w", "500", "-m", "3", "-x", "-5", "-g", "-4", "-q", "10", "-e", "0.3"],
            [prefix + ".reads.fastq", prefix + ".paf", prefix + ".contigs.fasta"])


@pytest.mark.parametrize("window", [500, 1000])
def test_cfg3_uneven_scaled(tmp_path, window):
    prefix = str(tmp_path / "cfg3")
    subprocess.run([SIMGEN, "contig", "1003", prefix, "20000,60000,160000", "40", "8000", "6000",
                    "0.15", "0.4:0.2:0.4", "--depth-gamma", "--min-len", "1000", "--max-len",
                    "40000", "--second", "0.5", "0.13", "0.2:0.5:0.3"], check=True)
    compare(["-w", str(window), "-m", "3", "-x", "-5", "-g", "-4", "-q", "10", "-e", "0.3"],
            [prefix + ".reads.fastq", prefix + ".paf", prefix + ".contigs.fasta"])


def test_cfg4_fragment_scaled(tmp_path):
    prefix = str(tmp_path / "cfg4")
    subprocess.run([SIMGEN, "frag", "1004", prefix, "400000", "600", "10000", "5000", "0.12",
                    "--max-overlaps", "40"], check=True)
    reads = prefix + ".reads.fastq"
    compare(["-f", "-w", "500", "-m", "3", "-x", "-5", "-g", "-4", "-q", "10", "-e", "0.3"],
            [reads, prefix + ".ava.paf", reads])
