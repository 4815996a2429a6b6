"""Window consensus on the GPU vs the CPU oracle, byte for byte."""
import sys

import numpy as np
import pytest

from racon_amd import _lib
from synth import make_windows, oracle_args

sys.path.insert(0, _lib.PKG_DIR + "/../oracle")
import oracle_py  # noqa: E402

pytestmark = pytest.mark.gpu

SCORES = [(3, -5, -4), (5, -4, -8), (1, -1, -1)]


def run_gpu(windows, scores, trim=True, mode=_lib.RH_POA_BATCHED):
    b = _lib.PoaBatch(*scores, trim=trim)
    b.upload(windows)
    b.run(mode)
    out, status = b.download()
    st = b.stats()
    b.close()
    return out, status, st


def check(windows, scores, trim=True, mode=_lib.RH_POA_BATCHED):
    out, status, st = run_gpu(windows, scores, trim, mode)
    cells = 0
    for i, w in enumerate(windows):
        bb, bq, layers, tgs = oracle_args(w)
        cons, ost, ocells, _ = oracle_py.window_consensus(bb, bq, layers, tgs, trim, *scores)
        print(f"window {i}: gpu {len(out[i])} bp status {status[i]} / oracle {len(cons)} bp "
              f"status {ost}", flush=True)
        assert status[i] == ost, f"window {i} status"
        assert out[i] == cons, f"window {i} consensus differs"
        cells += ocells
    assert st.cells == cells
    return st


@pytest.mark.parametrize("scores", SCORES)
def test_small_windows(scores):
    wins = make_windows(11, 6, length=120, depth=8)
    check(wins, scores)


def test_quality_mix_and_shallow():
    wins = make_windows(12, 4, length=300, depth=12, quality=False)
    wins += make_windows(13, 4, length=300, depth=12, backbone_quality=True)
    wins += make_windows(14, 2, length=200, depth=1)  # < 3 sequences
    wins += make_windows(15, 2, length=200, depth=0)
    check(wins, (3, -5, -4))


def test_no_trim_and_ngs():
    wins = make_windows(16, 4, length=250, depth=10, tgs=False)
    check(wins, (3, -5, -4))
    wins = make_windows(17, 4, length=250, depth=10)
    check(wins, (3, -5, -4), trim=False)


def test_other_bytes():
    # N and lower-case letters in the backbone and the layers: bytes are
    # compared as they are (a different byte is a mismatch, as in spoa)
    rng = np.random.default_rng(25)

    def spice(seq):
        out = bytearray(seq)
        for i in rng.choice(len(out), size=max(1, len(out) // 20), replace=False):
            out[i] = b"Nacgt"[rng.integers(0, 5)]
        return bytes(out)

    wins = []
    for w in make_windows(24, 4, length=300, depth=12):
        layers = [_lib.Layer(spice(l.sequence), l.quality, l.begin, l.end) for l in w.layers]
        wins.append(_lib.Window(spice(w.backbone), w.backbone_quality, layers, w.tgs))
    check(wins, (3, -5, -4))


def test_racon_sized_windows():
    wins = make_windows(18, 24, length=500, depth=30)
    check(wins, (3, -5, -4))


def test_plain_matches_batched():
    wins = make_windows(19, 5, length=400, depth=15)
    out_b, st_b, _ = run_gpu(wins, (3, -5, -4), mode=_lib.RH_POA_BATCHED)
    out_p, st_p, _ = run_gpu(wins, (3, -5, -4), mode=_lib.RH_POA_PLAIN)
    assert out_b == out_p and st_b == st_p


def test_static_grid_matches_oracle():
    # one workgroup per window in upload order (the comparator of the
    # persistent kernel), shallow windows mixed in
    wins = make_windows(25, 40, length=500, depth=20)
    wins[3:3] = make_windows(26, 2, length=200, depth=1)
    st = check(wins, (3, -5, -4), mode=_lib.RH_POA_STATIC)
    assert st.launches >= 1


def test_long_layers_two_tiles():
    # layers longer than one 512-column tile
    wins = make_windows(20, 3, length=1000, depth=12)
    check(wins, (5, -4, -8))


@pytest.mark.parametrize("ring_rows", [2, 4, 16])
def test_lds_ring_depths(ring_rows):
    # a shallow ring sends most predecessor rows through the slot's global H
    wins = make_windows(21, 6, length=400, depth=14)
    b = _lib.PoaBatch(3, -5, -4, trim=True)
    b.set_tuning(ring_rows=ring_rows)
    b.upload(wins)
    b.run()
    out, status = b.download()
    b.close()
    for i, w in enumerate(wins):
        bb, bq, layers, tgs = oracle_args(w)
        cons, ost, _, _ = oracle_py.window_consensus(bb, bq, layers, tgs, True, 3, -5, -4)
        assert (out[i], status[i]) == (cons, ost), f"window {i}"


def test_wide_scores_use_32bit_rows():
    # max(|m|,|x|,|g|) * (V + L) beyond 16 bits: the 32-bit ring path
    wins = make_windows(22, 4, length=500, depth=20)
    check(wins, (30, -40, -40))


def test_small_dp_row_capacity_reruns():
    # a DP row cap below the subgraph size forces the rerun in large slots
    wins = make_windows(23, 5, length=300, depth=10)
    b = _lib.PoaBatch(3, -5, -4, trim=True)
    b.set_tuning(rows=64)
    b.upload(wins)
    b.run()
    out, status = b.download()
    st = b.stats()
    b.close()
    assert st.retried == len(wins)
    for i, w in enumerate(wins):
        bb, bq, layers, tgs = oracle_args(w)
        cons, ost, _, _ = oracle_py.window_consensus(bb, bq, layers, tgs, True, 3, -5, -4)
        assert (out[i], status[i]) == (cons, ost), f"window {i}"


def test_submit_wait_two_batches_in_flight():
    # rh_poa_submit queues a batch's kernels on its own stream; two batches
    # in flight at once, waited for in the other order, give the oracle's
    # consensus; a running batch refuses another upload
    wa = make_windows(70, 12, length=500, depth=20)
    wb = make_windows(71, 9, length=400, depth=15, quality=False)
    ba = _lib.PoaBatch(3, -5, -4, trim=True)
    bb = _lib.PoaBatch(3, -5, -4, trim=True)
    ba.upload(wa)
    bb.upload(wb)
    ba.submit()
    bb.submit()
    with pytest.raises(RuntimeError):
        ba.upload(wa)
    bb.wait()
    ba.wait()
    for batch, wins in ((ba, wa), (bb, wb)):
        out, status = batch.download()
        assert batch.device_bytes() > 0
        batch.close()
        for i, w in enumerate(wins):
            cons, ost, _, _ = oracle_py.window_consensus(*oracle_args(w), True, 3, -5, -4)
            assert (out[i], status[i]) == (cons, ost), f"window {i}"


def test_reused_batch_stages_uploads_in_pinned_memory():
    # a batch object's second and later uploads pack into pinned staging and
    # copy the arena to the device chunk by chunk while packing; the results
    # must not depend on it (shallow windows return their backbone from the
    # staged bytes)
    big = make_windows(80, 400, length=500, depth=30)  # ~12 MB arena: several chunks
    small = make_windows(81, 5, length=300, depth=8) + make_windows(82, 3, length=200, depth=1)
    b = _lib.PoaBatch(3, -5, -4, trim=True)
    try:
        b.upload(big)  # first upload: pageable arena
        b.run()
        first = b.download()
        b.upload(small)
        b.run()
        out, status = b.download()
        for i, w in enumerate(small):
            cons, ost, _, _ = oracle_py.window_consensus(*oracle_args(w), True, 3, -5, -4)
            assert (out[i], status[i]) == (cons, ost), f"window {i}"
        b.upload(big)
        b.run()
        again = b.download()
    finally:
        b.close()
    assert again == first


def test_graph_beyond_the_size_estimate_reruns_exactly():
    # main passes size graph arrays for backbone + max(2 x longest layer,
    # layer bytes / 4) nodes; eight random layers six times as long as the
    # backbone grow the graph to about 1.3 times that (the oracle: 800-850
    # nodes against 650), so the window ends kPoaNodeOverflow and reruns in
    # exact slots
    from synth import random_seq
    rng = np.random.default_rng(90)
    wins = []
    for _ in range(3):
        backbone = random_seq(rng, 50)
        layers = [_lib.Layer(random_seq(rng, 300), None, 0, 49) for _ in range(8)]
        wins.append(_lib.Window(backbone, b"!" * 50, layers, True))
    out, status, st = run_gpu(wins, (3, -5, -4))
    assert st.retried >= 1
    for i, w in enumerate(wins):
        cons, ost, _, _ = oracle_py.window_consensus(*oracle_args(w), True, 3, -5, -4)
        assert (out[i], status[i]) == (cons, ost), f"window {i}"


def test_deep_windows_through_the_score_row_pool():
    # with a 2-row LDS ring nearly every DP row that a later row reads takes
    # an entry of the slot's pool of global score rows, and deep windows reuse
    # each entry many times per layer (an exhausted pool ends
    # kPoaRowOverflow and reruns exactly, like a row cap that is too small)
    wins = make_windows(92, 6, length=500, depth=60)
    b = _lib.PoaBatch(3, -5, -4, trim=True)
    b.set_tuning(ring_rows=2)
    b.upload(wins)
    b.run()
    out, status = b.download()
    b.close()
    for i, w in enumerate(wins):
        cons, ost, _, _ = oracle_py.window_consensus(*oracle_args(w), True, 3, -5, -4)
        assert (out[i], status[i]) == (cons, ost), f"window {i}"
