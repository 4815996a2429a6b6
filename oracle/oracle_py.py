"""ORACLE (test infrastructure only): ctypes access to oracle/_build/liboracle.so.

Imported from tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
only, always as the checker, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
RACON_CPU = os.path.join(ORACLE_DIR, "_build", "racon_cpu")

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing at {LIB_PATH}: run make -C oracle")
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_edlib_align.restype = ctypes.c_int32
    lib.oracle_edlib_align.argtypes = [
        ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
        ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
        ctypes.POINTER(ctypes.c_int32)]
    lib.oracle_edit_distance.restype = ctypes.c_int32
    lib.oracle_edit_distance.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p,
                                         ctypes.c_int32]
    lib.oracle_window_consensus.restype = ctypes.c_int32
    lib.oracle_window_consensus.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32,
        ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
        ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
        ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
        ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p, ctypes.c_uint64,
        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
        ctypes.POINTER(ctypes.c_uint64)]
    lib.oracle_poa_align.restype = ctypes.c_int32
    lib.oracle_poa_align.argtypes = [
        ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
        ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_int32,
        ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
        ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32, ctypes.c_char_p,
        ctypes.POINTER(ctypes.c_uint32)]
    lib.oracle_poa_graph_edges.restype = ctypes.c_int32
    lib.oracle_poa_graph_edges.argtypes = [
        ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
        ctypes.c_char_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32]
    lib.oracle_sorted_layer_order.restype = None
    lib.oracle_sorted_layer_order.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(ctypes.c_uint32)]
    _lib = lib
    return lib


def edlib_align(query: bytes, target: bytes, exact: bool = False) -> Tuple[int, bytes, int]:
    """(edit distance, ops, status) of the edlib NW path restatement."""
    lib = load()
    cap = len(query) + len(target) + 1
    ops = ctypes.create_string_buffer(cap)
    n = ctypes.c_uint64()
    st = ctypes.c_int32()
    d = lib.oracle_edlib_align(query, len(query), target, len(target), int(exact), ops, cap,
                               ctypes.byref(n), ctypes.byref(st))
    return d, ops.raw[:n.value], st.value


def edit_distance(query: bytes, target: bytes) -> int:
    return load().oracle_edit_distance(query, len(query), target, len(target))


def layer_order(begins: Sequence[int]) -> List[int]:
    """racon's layer order (std::sort by begin) as 0-based layer indices."""
    lib = load()
    n = len(begins)
    arr = (ctypes.c_uint32 * max(1, n))(*begins)
    out = (ctypes.c_uint32 * max(1, n))()
    lib.oracle_sorted_layer_order(n, arr, out)
    return [out[i] for i in range(n)]


def window_consensus(backbone: bytes, backbone_quality: bytes, layers, tgs: bool,
                     trim: bool, match: int, mismatch: int, gap: int,
                     presorted: bool = False):
    """(consensus, status, cells, graph nodes) of one window.

    layers: sequence of (sequence, quality or None, begin, end) in insertion
    order; they are put in racon's order first unless presorted.
    """
    lib = load()
    if not presorted:
        order = layer_order([l[2] for l in layers])
        layers = [layers[i] for i in order]
    n = len(layers)
    seqs = (ctypes.c_char_p * max(1, n))(*[l[0] for l in layers])
    quals = (ctypes.c_char_p * max(1, n))(*[l[1] for l in layers])
    lens = (ctypes.c_uint32 * max(1, n))(*[len(l[0]) for l in layers])
    begins = (ctypes.c_uint32 * max(1, n))(*[l[2] for l in layers])
    ends = (ctypes.c_uint32 * max(1, n))(*[l[3] for l in layers])
    cap = len(backbone) + sum(len(l[0]) for l in layers) + 16
    out = ctypes.create_string_buffer(cap)
    out_len = ctypes.c_uint64()
    cells = ctypes.c_uint64()
    nodes = ctypes.c_uint64()
    st = lib.oracle_window_consensus(backbone, backbone_quality, len(backbone), n, seqs, quals,
                                     lens, begins, ends, int(tgs), int(trim), match, mismatch,
                                     gap, out, cap, ctypes.byref(out_len), ctypes.byref(cells),
                                     ctypes.byref(nodes))
    if st < 0:
        raise RuntimeError("oracle window output overflow")
    return out.raw[:out_len.value], st, cells.value, nodes.value


def poa_align(chains, sequence: bytes, match: int, mismatch: int, gap: int,
              layer: bytes = None):
    """Aligns sequence to a graph of parallel chains (+ an inserted layer).

    Returns (pairs, node_bases, edges): the alignment as (node, position)
    pairs, the base of every node and the (tail, head) edge list.
    """
    lib = load()
    n = len(chains)
    arr = (ctypes.c_char_p * max(1, n))(*chains)
    lens = (ctypes.c_uint32 * max(1, n))(*[len(c) for c in chains])
    cap = sum(len(c) for c in chains) + len(sequence) + (len(layer) if layer else 0) + 8
    nodes = (ctypes.c_int32 * cap)()
    pos = (ctypes.c_int32 * cap)()
    bases = ctypes.create_string_buffer(cap)
    nn = ctypes.c_uint32()
    k = lib.oracle_poa_align(arr, lens, n, layer, len(layer) if layer else 0, sequence,
                             len(sequence), match, mismatch, gap, nodes, pos, cap, bases,
                             ctypes.byref(nn))
    if k < 0:
        raise RuntimeError("oracle_poa_align capacity")
    ecap = 4 * cap
    tails = (ctypes.c_int32 * ecap)()
    heads = (ctypes.c_int32 * ecap)()
    ne = lib.oracle_poa_graph_edges(arr, lens, n, layer, len(layer) if layer else 0, match,
                                    mismatch, gap, tails, heads, ecap)
    return ([(nodes[i], pos[i]) for i in range(k)], bases.raw[:nn.value],
            [(tails[i], heads[i]) for i in range(ne)])


def breaking_points(ops: bytes, t_begin: int, t_end: int, q_first: int,
                    window_length: int) -> List[Tuple[int, int]]:
    """racon's breaking-point walk (src/overlap.cpp:226-292) over edit
    operations (0 match, 1 insertion, 2 deletion, 3 mismatch; the CIGAR of
    src/overlap.cpp:205-216 run by run): for each window of window_length
    target bases the alignment crosses, the (target, query) position of its
    first match and the position after its last match."""
    ends = [i - 1 for i in range(0, t_end, window_length) if i > t_begin]
    ends.append(t_end - 1)
    out: List[Tuple[int, int]] = []
    w = 0
    found = False
    first = last = (0, 0)
    q = q_first - 1
    t = t_begin - 1

    def end_reached():
        nonlocal w, found
        if w < len(ends) and t == ends[w]:
            if found:
                out.append(first)
                out.append(last)
            found = False
            w += 1

    for o in ops:
        o &= 3
        if o == 1:
            q += 1
        elif o == 2:
            t += 1
            end_reached()
        else:
            q += 1
            t += 1
            if not found:
                found = True
                first = (t, q)
            last = (t + 1, q + 1)
            end_reached()
    return out
