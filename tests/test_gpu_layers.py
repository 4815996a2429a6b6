"""Layer extraction on the device (racon src/polisher.cpp:403-457): rh_layers_*
over the device breaking points against racon's loop restated here over the
same points (the points themselves are checked against the oracle walk in
test_breaking_points.py), and the CLI's windows (--dump-windows, layers in
racon's order) against the oracle CLI's, which runs racon's host loop."""
import os
import random
import subprocess

import pytest

from racon_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_CLI = os.path.join(ROOT, "oracle", "_build", "racon_cpu")


def racon_layers(points, sources, qualities, w, threshold):
    """polisher.cpp:403-457 over the overlaps in order: the span filter, the
    mean-quality filter (double sum of uint32(char) - 33, char signed), the
    window id and add_layer's begin/end; layers listed window by window."""
    per_window = {}
    for pts, (ov, qlen, fw, qoff, mirrored) in zip(points, sources):
        for j in range(0, len(pts), 2):
            (t0, q0), (t1, q1) = pts[j], pts[j + 1]
            span = (q1 - q0) & 0xFFFFFFFF
            if span < 0.02 * w:
                continue
            if qoff is not None:
                average = 0.0
                for k in range(q0, q1):
                    c = qualities[qoff + (qlen - 1 - k if mirrored else k)]
                    signed = c - 256 if c >= 128 else c
                    average += float(((signed & 0xFFFFFFFF) - 33) & 0xFFFFFFFF)
                average /= span
                if average < threshold:
                    continue
            start = (t0 // w) * w
            per_window.setdefault(fw + t0 // w, []).append(
                (fw + t0 // w, ov, q0, q1, t0 - start, t1 - start - 1))
    return [r for k in sorted(per_window) for r in per_window[k]]


def _mutate(rng, seq, rate):
    out = bytearray()
    for c in seq:
        r = rng.random()
        if r < rate / 3:
            continue
        if r < 2 * rate / 3:
            out.append(rng.choice(b"ACGT"))
        elif r < rate:
            out.append(c)
            out.append(rng.choice(b"ACGT"))
        else:
            out.append(c)
    return bytes(out) or b"A"


def _quality(rng, n, kind):
    if kind == "random":
        return bytes(33 + rng.randrange(0, 25) for _ in range(n))
    if kind == "at_threshold":  # mean exactly 10: kept (racon drops below only)
        return b"+" * n
    if kind == "below":  # mean 9: dropped
        return b"*" * n
    if kind == "high_bytes":  # sign-extended bytes: huge uint32 terms
        return bytes(rng.choice([33, 200, 255]) for _ in range(n))
    return bytes(rng.choice([33, 60]) for _ in range(n))  # mixed around 10


@pytest.mark.parametrize("w", [500, 37])
def test_device_layers_match_racon_loop(w):
    rng = random.Random(29 + w)
    threshold = 10.0
    n_targets, target_length = 4, 30000
    first_window = [k * ((target_length + w - 1) // w) for k in range(n_targets + 1)]
    n_windows = first_window[-1]
    kinds = ["random", "at_threshold", "below", "high_bytes", "mixed", None]
    batches = []
    qualities = bytearray()
    overlap = 0
    for size in (70, 45):
        pairs, spans, sources = [], [], []
        for _ in range(size):
            t = bytes(rng.choice(b"ACGT") for _ in range(rng.randrange(1, 2500)))
            q = _mutate(rng, t, 0.15)
            tid = rng.randrange(n_targets)
            t_begin = rng.choice([0, 499, 500, 501, rng.randrange(0, target_length - len(t))])
            t_begin = min(t_begin, target_length - len(t))
            spans.append((t_begin, t_begin + len(t), rng.randrange(0, 3)))
            pairs.append((q, t))
            # the read: the query plus a few bases either side (q_first), its
            # quality stored forward or mirrored
            qlen = len(q) + spans[-1][2] + rng.randrange(0, 3)
            kind = rng.choice(kinds)
            if kind is None:
                qoff = None
            else:
                qoff = len(qualities)
                qualities += _quality(rng, qlen, kind)
            sources.append((overlap, qlen, first_window[tid], qoff, rng.random() < 0.5))
            overlap += 1
        batches.append((pairs, spans, sources))

    lib = _lib.load_library()
    capacity = 0
    for _, spans, _ in batches:
        arr = (_lib.RhBreakSpan * len(spans))()
        for i, (tb, te, qf) in enumerate(spans):
            arr[i].t_begin, arr[i].t_end, arr[i].q_first = tb, te, qf
        capacity += lib.rh_breaking_points_bound(w, arr, len(spans)) // 2
    layers = _lib.LayerSet(w, threshold, n_windows, capacity, bytes(qualities))
    b = _lib.AlignBatch()  # one batch, reused: the set keeps every batch's layers
    all_points, all_sources = [], []
    for pairs, spans, sources in batches:
        b.upload(pairs)
        b.run()
        pts = b.breaking_points(w, spans)
        layers.add(b, sources)
        all_points += pts
        all_sources += sources
    got = layers.finish()
    layers.close()
    b.close()
    want = racon_layers(all_points, all_sources, qualities, w, threshold)
    print(f"w {w}: {len(got)} layers of {sum(len(p) // 2 for p in all_points)} pairs", flush=True)
    assert len(want) > 50
    assert got == want


def test_layer_set_rejects_bad_sources():
    b = _lib.AlignBatch()
    b.upload([(b"ACGTACGT", b"ACGTACGA")])
    b.run()
    s = _lib.LayerSet(37, 10.0, 4, 8, b"+" * 8)
    with pytest.raises(RuntimeError):
        s.add(b, [(0, 8, 0, None, False)])  # no breaking points computed yet
    # one window (t 0..7): first match (0, 0), after the last (8, 8); span 8
    # >= 0.02 * 37, mean quality exactly 10
    assert b.breaking_points(37, [(0, 8, 0)]) == [[(0, 0), (8, 8)]]
    with pytest.raises(RuntimeError):
        s.add(b, [(0, 9, 0, 0, False)])  # quality runs past the bytes
    with pytest.raises(RuntimeError):
        s.add(b, [(0, 8, 0, 0, False), (1, 8, 0, 0, False)])  # not the batch's pairs
    s.add(b, [(0, 8, 0, 0, False)])
    assert s.finish() == [(0, 0, 0, 8, 0, 7)]
    s.close()
    b.close()


DUMP_CASES = [
    ("with_qualities", "sample_reads.fastq.gz", "sample_overlaps.paf.gz",
     "sample_layout.fasta.gz", ["-m", "5", "-x", "-4", "-g", "-8", "-w", "500"]),
    ("without_qualities", "sample_reads.fasta.gz", "sample_overlaps.paf.gz",
     "sample_layout.fasta.gz", ["-m", "5", "-x", "-4", "-g", "-8", "-w", "500"]),
    ("fragment_kf_q20", "sample_reads.fastq.gz", "sample_ava_overlaps.paf.gz",
     "sample_reads.fastq.gz", ["-w", "500", "-f", "-q", "20"]),
    ("larger_window_q15", "sample_reads.fastq.gz", "sample_overlaps.paf.gz",
     "sample_layout.fasta.gz", ["-w", "1000", "-q", "15"]),
]


@pytest.mark.parametrize("name,reads,overlaps,targets,args", DUMP_CASES,
                         ids=[c[0] for c in DUMP_CASES])
def test_cli_windows_match_oracle(name, reads, overlaps, targets, args, data_dir, tmp_path):
    files = [os.path.join(data_dir, f) for f in (reads, overlaps, targets)]
    dumps = []
    for binary, tag in ((_lib.RACON_BIN, "gpu"), (ORACLE_CLI, "oracle")):
        path = str(tmp_path / f"{tag}.windows")
        p = subprocess.run([binary, "-t", "8"] + args + ["--dump-windows", path] + files,
                           capture_output=True, timeout=600)
        assert p.returncode == 0, p.stderr.decode()[-2000:]
        with open(path, "rb") as f:
            dumps.append(f.read())
    print(f"{name}: windows dump {len(dumps[0])} bytes", flush=True)
    assert dumps[0] == dumps[1]
