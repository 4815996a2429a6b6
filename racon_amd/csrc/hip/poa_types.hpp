// Shared host/device layouts of the POA batch (windows, layers, per-slot
// scratch). Plain structs only; no HIP or STL types.
#pragma once

#include <cstdint>

namespace rh {

// One window as packed in device memory (layers in racon's alignment order).
struct PoaWindow {
  uint64_t backbone_off;   // arena offset of the backbone bytes
  uint64_t backbone_qual;  // arena offset of the backbone quality bytes
  uint64_t out_off;        // output arena offset
  uint32_t backbone_len;
  uint32_t n_layers;
  uint32_t layer_first;    // index of the first layer in the layer table
  uint32_t offset;         // (uint32)(0.01 * backbone_len), window.cpp:87
  uint32_t flags;          // bit 0: TGS window, bit 1: trim
  uint32_t out_cap;        // bytes reserved for the consensus
  uint32_t node_bound;     // backbone_len + sum of layer lengths
  uint32_t edge_bound;     // upper bound on edges
  uint32_t max_layer_len;
  uint32_t pad;
};

struct PoaLayer {
  uint64_t seq_off;   // arena offset of the layer bytes
  uint64_t qual_off;  // arena offset of the quality bytes, or kNoQuality
  uint32_t len;
  uint32_t begin;
  uint32_t end;
  uint32_t pad;
};

constexpr uint64_t kNoQuality = ~uint64_t(0);

// Phases of the per-window work, timed with the device's constant clock.
enum : int {
  kPhaseInclude = 0,   // subgraph node set
  kPhaseSubSort,       // subgraph topological order
  kPhaseRows,          // DP row map and predecessor lists
  kPhaseColumn0,       // DP column 0
  kPhaseDp,            // DP rows
  kPhaseTraceback,     // end cell and traceback
  kPhaseInsert,        // graph insertion
  kPhaseSort,          // topological order of the whole graph
  kPhaseConsensus,     // heaviest bundle, coverage, trimming
  kPhaseSetup,         // backbone chain
  kPoaPhases
};

// Per-window result written by the kernel.
struct PoaResult {
  uint32_t length;     // consensus bytes written at out_off
  int32_t status;      // kPoa* below
  uint64_t cells;      // DP cells computed (sum |V_aligned| * L)
  uint32_t nodes;      // final graph size
  uint32_t max_rows;   // largest DP row count used
  uint64_t phase[kPoaPhases];  // device clock ticks spent per phase
};

enum : int32_t {
  kPoaPolished = 1,
  kPoaPolishedUntrimmed = 2,   // trimming would empty it (racon warns)
  kPoaNodeOverflow = -1,       // graph scratch too small: rerun with bigger slot
  kPoaRowOverflow = -2,        // DP matrix scratch too small
  kPoaOutOverflow = -3,        // consensus longer than out_cap
  kPoaLoopLimit = -4,          // heaviest-bundle completion did not end
  kPoaBadInput = -5,           // layer end outside the graph
  kPoaInDegree = -6,           // a node with more than 8192 in-edges
  kPoaBandMiss = -7,           // banded DP: the path leaves the band (rerun unbanded)
  kPoaAlignedFull = -8,        // an aligned set is full: rerun in the wide kernel
  kPoaNotRun = -100,
  kPoaRejectedOnHost = -101,   // outside the device limits, never launched
};

// Capacities of one scratch slot.
struct PoaSlotCaps {
  uint32_t nodes;   // graph nodes
  uint32_t edges;   // graph edges
  uint32_t rows;    // DP rows (subgraph nodes + 1)
  uint32_t cols;    // DP row pitch in elements (>= max layer length + 1, padded)
  uint32_t stack;   // DFS stack entries
  uint32_t aligned; // aligned-set entries per node (the kernel variant's)
  uint32_t hrows;   // pooled global score rows (0: one per DP row)
};

// Pooled global score rows of a main pass: a DP row that finds no LDS slot
// takes one of these (a window that needs more reruns in exact slots, which
// keep one global row per DP row).
constexpr uint32_t kPoaHPoolRows = 128;

// Aligned-set entries kept per node: a column of the graph holds at most
// this many + 1 distinct bytes. The default kernel keeps kMaxAligned; the
// wide variant (poa_kernel_wide.hip) reruns windows that need more (A, C,
// G, T and the 11 IUPAC ambiguity codes, N included, are 15 bytes).
constexpr uint32_t kMaxAligned = 8;
constexpr uint32_t kMaxAlignedWide = 16;

// DP row pitch >= layer length + kPoaColPad: strips that end past column L
// (up to 12 columns, stored from column + 3) stay inside their row.
constexpr uint32_t kPoaColPad = 24;

// Byte offsets of every array inside one slot.
struct PoaSlotLayout {
  uint64_t base, in_head, out_head, in_tail, out_tail, n_aligned, aligned, cov,
      last_label, e_tail, e_head, e_weight, e_next_in, e_next_out, rank_to_node,
      node_rank, mark, lead, stack, in_sub, row_of, sub_order, pred_off, pred_rows,
      row_meta, H, code, scores, preds, path, seq, qual, anchor, lpath,
      sink_score, tie_mark, row_pred4, total;
  uint64_t hpool;  // caps.hrows
};

constexpr uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

inline PoaSlotLayout poa_slot_layout(const PoaSlotCaps& c) {
  PoaSlotLayout l{};
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    uint64_t at = o;
    o = align_up(o + bytes, 256);
    return at;
  };
  const uint64_t n = c.nodes, e = c.edges;
  l.base = take(n);
  l.in_head = take(4 * n);
  l.out_head = take(4 * n);
  l.in_tail = take(4 * n);
  l.out_tail = take(4 * n);
  l.n_aligned = take(n);
  l.aligned = take(4 * n * (c.aligned != 0 ? c.aligned : kMaxAligned));
  l.cov = take(4 * n);
  l.last_label = take(4 * n);
  l.e_tail = take(4 * e);
  l.e_head = take(4 * e);
  l.e_weight = take(4 * e);
  l.e_next_in = take(4 * e);
  l.e_next_out = take(4 * e);
  l.rank_to_node = take(4 * n);
  l.node_rank = take(4 * n);
  l.mark = take(n);
  l.lead = take(n);
  l.stack = take(4 * uint64_t(c.stack));
  l.in_sub = take(n);
  l.row_of = take(4 * n);
  l.sub_order = take(4 * n);
  l.pred_off = take(4 * (n + 1));
  l.pred_rows = take(4 * (e + n));
  l.row_meta = take(4 * n);
  l.H = take(4 * uint64_t(c.hrows != 0 ? c.hrows : c.rows) * c.cols);
  l.code = take(2 * uint64_t(c.rows) * c.cols);
  l.scores = take(8 * n);
  l.preds = take(4 * n);
  l.path = take(4 * n);
  l.seq = take(c.cols);
  l.qual = take(4 * uint64_t(c.cols));
  l.anchor = take(4 * uint64_t(c.cols));
  l.lpath = take(4 * uint64_t(c.cols));
  l.sink_score = take(4 * (n + 1));
  l.tie_mark = take(4 * n);
  l.row_pred4 = take(8 * uint64_t(c.rows));
  l.total = o;
  l.hpool = c.hrows;
  return l;
}

struct PoaParams {
  const PoaWindow* windows;
  const PoaLayer* layers;
  const uint8_t* arena;
  uint8_t* out;
  PoaResult* results;
  const uint32_t* order;     // processing order of window indices
  uint32_t n_order;          // number of entries in order
  uint32_t* queue_head;      // work counter (persistent mode)
  uint8_t* scratch;          // slots
  uint64_t slot_bytes;
  PoaSlotCaps caps;
  PoaSlotLayout layout;
  int32_t match;
  int32_t mismatch;
  int32_t gap;
  uint32_t n_slots;
  uint32_t ring_rows;  // DP rows of 16-bit scores kept in LDS (int32 mode: half)
  uint32_t lds_bytes;  // dynamic LDS per wave: ring + layer bytes
  uint32_t n_priority; // queue positions below this run at raised wave priority
  uint32_t band;       // 0: exact DP; else band rows for long layers (racon -b)
};

// Dynamic LDS of one wave: the DP row ring (ring_rows x cols x 2 bytes)
// followed by the current layer's bytes (cols).
inline uint32_t poa_lds_bytes(uint32_t cols, uint32_t ring_rows) {
  return static_cast<uint32_t>(align_up(uint64_t(ring_rows) * cols * 2, 16) + align_up(cols, 16));
}

}  // namespace rh
