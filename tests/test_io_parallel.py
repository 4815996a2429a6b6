"""The overlap parser's parallel ranges (files of 1 MiB or more are cut at line
starts and parsed on the pool, then appended in file order): the oracle CLI's
windows (--dump-windows: every window's layers in racon's order) are the same
with 1 and with 8 threads on PAF and MHAP inputs above that size. The inputs
are racon's fixtures with the overlap lines repeated; racon's filter keeps one
overlap per read among the copies, so the result depends on their order."""
import gzip
import os
import subprocess

import oracle_py

D = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "racon_data")


def _repeat(src, dst, copies):
    with gzip.open(os.path.join(D, src), "rb") as f:
        text = f.read()
    with open(dst, "wb") as f:
        for _ in range(copies):
            f.write(text)
    return os.path.getsize(dst)


def _windows(threads, reads, overlaps, targets, out):
    p = subprocess.run([oracle_py.RACON_CPU, "-t", str(threads), "-w", "500",
                        "--dump-windows", out, os.path.join(D, reads), overlaps,
                        os.path.join(D, targets)], capture_output=True, timeout=600)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    with open(out, "rb") as f:
        return f.read()


def test_paf_ranges_keep_file_order(tmp_path):
    paf = str(tmp_path / "overlaps.paf")
    assert _repeat("sample_ava_overlaps.paf.gz", paf, 3) >= 1 << 20
    one = _windows(1, "sample_reads.fastq.gz", paf, "sample_reads.fastq.gz",
                   str(tmp_path / "t1.windows"))
    eight = _windows(8, "sample_reads.fastq.gz", paf, "sample_reads.fastq.gz",
                     str(tmp_path / "t8.windows"))
    assert len(one) > 1000
    assert one == eight


def test_mhap_ranges_keep_file_order(tmp_path):
    mhap = str(tmp_path / "overlaps.mhap")
    assert _repeat("sample_ava_overlaps.mhap.gz", mhap, 3) >= 1 << 20
    one = _windows(1, "sample_reads.fastq.gz", mhap, "sample_reads.fastq.gz",
                   str(tmp_path / "t1.windows"))
    eight = _windows(8, "sample_reads.fastq.gz", mhap, "sample_reads.fastq.gz",
                     str(tmp_path / "t8.windows"))
    assert len(one) > 1000
    assert one == eight
