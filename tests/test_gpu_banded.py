"""racon -b (--cuda-banded-alignment): the banded POA approximation.

racon pins its device path only as "close to the CPU path" (the CUDA cases
of test/racon_test.cpp:292-496 carry their own values beside the CPU ones,
e.g. 1385 vs 1312). The banded rows here are this build's own design (a
256-column band along the diagonal, rerunning exactly any window whose path
leaves it), so the checks are: the polished contigs stay within a few
percent of the CPU path's edit distance to the reference, fragment
correction keeps racon's sequence count and total length within 1%, and on
synthetic windows most consensus strings equal the exact ones.
"""
import gzip
import os
import subprocess
import sys

import pytest

from racon_amd import _lib
from synth import make_windows

pytestmark = pytest.mark.gpu

D = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "racon_data")

# (case, reads, overlaps, window, scores, CPU pin, CUDA pin of racon_test.cpp)
CONTIG = [
    ("ConsensusWithQualities", "sample_reads.fastq.gz", "sample_overlaps.paf.gz", 500,
     (5, -4, -8), 1312, 1385),
    ("ConsensusWithoutQualities", "sample_reads.fasta.gz", "sample_overlaps.paf.gz", 500,
     (5, -4, -8), 1566, 1607),
    ("ConsensusWithQualitiesLargerWindow", "sample_reads.fastq.gz", "sample_overlaps.paf.gz",
     1000, (5, -4, -8), 1289, 4168),
    ("ConsensusWithQualitiesEditDistance", "sample_reads.fastq.gz", "sample_overlaps.paf.gz",
     500, (1, -1, -1), 1321, 1361),
]


def run_cli(args):
    p = subprocess.run([_lib.RACON_BIN, "-t", "8"] + args, capture_output=True, timeout=600)
    assert p.returncode == 0, p.stderr.decode()[-1000:]
    lines = p.stdout.decode().split("\n")
    return [(lines[i][1:], lines[i + 1]) for i in range(0, len(lines) - 1, 2)]


def reverse_complement(s):
    return s.translate(str.maketrans("ACGT", "TGCA"))[::-1]


@pytest.fixture(scope="module")
def reference():
    with gzip.open(os.path.join(D, "sample_reference.fasta.gz"), "rt") as f:
        return "".join(l.strip() for l in f if not l.startswith(">"))


@pytest.mark.parametrize("name,reads,overlaps,w,scores,cpu,cuda", CONTIG,
                         ids=[c[0] for c in CONTIG])
def test_banded_contig_close_to_cpu(name, reads, overlaps, w, scores, cpu, cuda, reference):
    sys.path.insert(0, os.path.join(_lib.PKG_DIR, "..", "oracle"))
    import oracle_py

    m, x, g = scores
    args = ["-w", str(w), "-q", "10", "-e", "0.3", "-m", str(m), "-x", str(x), "-g", str(g),
            os.path.join(D, reads), os.path.join(D, overlaps),
            os.path.join(D, "sample_layout.fasta.gz")]
    exact = run_cli(args)
    banded = run_cli(["-b"] + args)
    assert len(banded) == 1
    d_exact = oracle_py.edit_distance(reverse_complement(exact[0][1]).encode(),
                                      reference.encode())
    d_band = oracle_py.edit_distance(reverse_complement(banded[0][1]).encode(),
                                     reference.encode())
    print(f"{name}: exact {d_exact} (racon CPU {cpu}), banded {d_band} (racon CUDA {cuda})",
          flush=True)
    assert d_exact == cpu
    assert d_band <= 1.05 * cpu + 10


def test_banded_fragment_correction_close_to_cpu():
    out = run_cli(["-b", "-f", "-u", "-w", "500", "-q", "10", "-e", "0.3", "-m", "1", "-x", "-1",
                   "-g", "-1", os.path.join(D, "sample_reads.fastq.gz"),
                   os.path.join(D, "sample_ava_overlaps.paf.gz"),
                   os.path.join(D, "sample_reads.fastq.gz")])
    total = sum(len(s) for _, s in out)
    print(f"banded fragment correction: {len(out)} sequences, {total} bp (CPU 236, 1658216)",
          flush=True)
    assert len(out) == 236
    assert abs(total - 1658216) <= 0.01 * 1658216


@pytest.mark.parametrize("length", [500, 1000])
def test_banded_windows_mostly_equal_exact(length):
    ws = make_windows(900 + length, 48, length=length, depth=30)
    out = {}
    for band in (0, 256):
        b = _lib.PoaBatch(3, -5, -4, trim=True)
        b.set_band(band)
        b.upload(ws)
        b.run()
        out[band] = b.download()
        st = b.stats()
        print(f"length {length} band {band}: kernel {st.kernel_ms:.2f} ms, cells {st.cells}, "
              f"band reruns {st.band_reruns}, retried {st.retried}", flush=True)
        b.close()
    same = sum(a == c for a, c in zip(out[0][0], out[256][0]))
    print(f"length {length}: {same}/{len(ws)} banded consensus strings equal the exact ones",
          flush=True)
    assert all(s in (_lib.RH_WINDOW_POLISHED, _lib.RH_WINDOW_POLISHED_UNTRIMMED)
               for s in out[256][1])
    assert same >= 0.75 * len(ws)
    # every banded consensus stays close to the exact one
    for a, c in zip(out[0][0], out[256][0]):
        assert abs(len(a) - len(c)) <= 0.05 * len(a) + 5
