"""Pins the CPU restatement of spoa's POA (the oracle for the POA kernel).

- DP score: on a chain graph it is textbook linear-gap Needleman-Wunsch; on
  small DAGs (chains plus an inserted layer) it is the best NW score over
  every source-to-sink path (brute-force enumeration);
- alignment validity: the aligned nodes form a source-to-sink path, bases
  and positions line up, the pair score equals the optimum;
- consensus: identical layers give the backbone; a majority of layers
  carrying the same correction wins; trimming keeps well-covered ends.
What only spoa's own source could pin (tie-breaking among co-optimal
alignments and heaviest bundles, DFS rank order) is pinned end to end by
racon's test values in test_racon_pins.py.
"""
import numpy as np

import oracle_py
from synth import mutate, random_seq

SCORES = [(3, -5, -4), (5, -4, -8), (1, -1, -1)]


def nw_linear(a: bytes, b: bytes, m: int, x: int, g: int) -> int:
    H = np.zeros((len(a) + 1, len(b) + 1), dtype=np.int64)
    H[0, :] = np.arange(len(b) + 1) * g
    H[:, 0] = np.arange(len(a) + 1) * g
    for i in range(1, len(a) + 1):
        for j in range(1, len(b) + 1):
            s = m if a[i - 1] == b[j - 1] else x
            H[i, j] = max(H[i - 1, j - 1] + s, H[i - 1, j] + g, H[i, j - 1] + g)
    return int(H[-1, -1])


def pair_score(pairs, bases, seq, m, x, g):
    total = 0
    for node, pos in pairs:
        if node == -1 or pos == -1:
            total += g
        else:
            total += m if bases[node] == seq[pos] else x
    return total


def all_paths(n_nodes, edges):
    succ = [[] for _ in range(n_nodes)]
    indeg = [0] * n_nodes
    for t, h in edges:
        succ[t].append(h)
        indeg[h] += 1
    out = []

    def walk(v, path):
        path.append(v)
        if not succ[v]:
            out.append(list(path))
        for w in succ[v]:
            walk(w, path)
        path.pop()

    for v in range(n_nodes):
        if indeg[v] == 0:
            walk(v, [])
    return out, succ, indeg


def test_chain_is_textbook_nw():
    rng = np.random.default_rng(200)
    for m, x, g in SCORES:
        for _ in range(20):
            bb = random_seq(rng, int(rng.integers(1, 25)))
            seq = mutate(rng, bb, 0.3)
            pairs, bases, _ = oracle_py.poa_align([bb], seq, m, x, g)
            assert pair_score(pairs, bases, seq, m, x, g) == nw_linear(bb, seq, m, x, g)
            assert [p for _, p in pairs if p != -1] == list(range(len(seq)))


def test_small_dags_brute_force():
    rng = np.random.default_rng(201)
    for m, x, g in SCORES:
        for _ in range(25):
            a = random_seq(rng, int(rng.integers(2, 9)))
            chains = [a]
            if rng.random() < 0.5:
                chains.append(mutate(rng, a, 0.4))
            layer = mutate(rng, a, 0.35)
            seq = mutate(rng, a, 0.3)
            pairs, bases, edges = oracle_py.poa_align(chains, seq, m, x, g, layer=layer)
            paths, succ, indeg = all_paths(len(bases), edges)
            best = max(nw_linear(bytes(bases[v] for v in p), seq, m, x, g) for p in paths)
            assert pair_score(pairs, bases, seq, m, x, g) == best
            nodes = [n for n, _ in pairs if n != -1]
            assert indeg[nodes[0]] == 0 and not succ[nodes[-1]]
            for u, v in zip(nodes, nodes[1:]):
                assert v in succ[u]
            assert [p for _, p in pairs if p != -1] == list(range(len(seq)))


def q(n, v=30):
    return bytes([33 + v]) * n


def test_identical_layers_give_backbone():
    rng = np.random.default_rng(202)
    bb = random_seq(rng, 200)
    layers = [(bb, q(200), 0, 199) for _ in range(6)]
    for scores in SCORES:
        cons, st, _, _ = oracle_py.window_consensus(bb, b"!" * 200, layers, True, True, *scores)
        assert (cons, st) == (bb, 1)


def test_majority_correction_wins():
    rng = np.random.default_rng(203)
    truth = random_seq(rng, 150)
    bb = truth[:70] + (b"A" if truth[70:71] != b"A" else b"C") + truth[71:]
    layers = [(truth, q(150), 0, 149) for _ in range(7)]
    cons, st, _, _ = oracle_py.window_consensus(bb, b"!" * 150, layers, True, False, 3, -5, -4)
    assert cons == truth and st == 1


def test_shallow_window_is_backbone():
    bb = b"ACGTACGTAC"
    cons, st, cells, _ = oracle_py.window_consensus(bb, b"!" * 10, [(bb, None, 0, 9)], True,
                                                    True, 3, -5, -4)
    assert (cons, st, cells) == (bb, 0, 0)


def test_trim_cuts_low_coverage_ends():
    rng = np.random.default_rng(204)
    bb = random_seq(rng, 300)
    # four layers cover only 100..199; the ends are backbone-only
    layers = [(bb[100:200], q(100), 100, 199) for _ in range(4)]
    cons, st, _, _ = oracle_py.window_consensus(bb, b"!" * 300, layers, True, True, 3, -5, -4)
    assert st == 1 and cons == bb[100:200]
    cons, st, _, _ = oracle_py.window_consensus(bb, b"!" * 300, layers, True, False, 3, -5, -4)
    assert cons == bb
