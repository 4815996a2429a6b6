// ORACLE (test infrastructure only): a small C ABI over the CPU
// restatements so pytest can compare single items against the device
// library. Loaded with ctypes from tests/ only.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <cstring>
#include <string>
#include <vector>

#include "edlib_cpu.hpp"
#include "poa_cpu.hpp"

extern "C" {

// Returns the edit distance (or -1); writes min(n_ops, capacity) ops.
int32_t oracle_edlib_align(const char* query, int32_t query_length, const char* target,
                           int32_t target_length, int32_t exact, uint8_t* ops,
                           uint64_t capacity, uint64_t* n_ops, int32_t* status) {
  oracle::EdlibResult r =
      exact ? oracle::edlib_align_nw_exact(query, query_length, target, target_length)
            : oracle::edlib_align_nw(query, query_length, target, target_length);
  *n_ops = r.ops.size();
  *status = r.status;
  if (ops != nullptr) {
    memcpy(ops, r.ops.data(), std::min<uint64_t>(capacity, r.ops.size()));
  }
  return r.edit_distance;
}

int32_t oracle_edit_distance(const char* query, int32_t query_length, const char* target,
                             int32_t target_length) {
  return oracle::edit_distance(query, query_length, target, target_length);
}

// One window: layers given in alignment order. Writes the consensus
// (capacity bytes at most) and returns the status (0/1/2), or -1 when the
// output does not fit.
int32_t oracle_window_consensus(const char* backbone, const char* backbone_quality,
                                uint32_t backbone_length, uint32_t n_layers,
                                const char* const* sequences, const char* const* qualities,
                                const uint32_t* lengths, const uint32_t* begins,
                                const uint32_t* ends, int32_t tgs, int32_t trim,
                                int32_t match, int32_t mismatch, int32_t gap, char* out,
                                uint64_t capacity, uint64_t* out_length, uint64_t* cells,
                                uint64_t* graph_nodes) {
  oracle::WindowInput in;
  in.backbone = backbone;
  in.backbone_quality = backbone_quality;
  in.backbone_length = backbone_length;
  in.tgs = tgs != 0;
  for (uint32_t i = 0; i < n_layers; ++i) {
    in.layers.push_back({sequences[i], qualities[i], lengths[i], begins[i], ends[i]});
  }
  std::string consensus;
  oracle::PoaStats stats;
  uint8_t status = oracle::window_consensus(in, match, mismatch, gap, trim != 0,
                                            consensus, &stats);
  *out_length = consensus.size();
  if (cells != nullptr) *cells = stats.cells;
  if (graph_nodes != nullptr) *graph_nodes = stats.graph_nodes;
  if (consensus.size() > capacity) {
    return -1;
  }
  memcpy(out, consensus.data(), consensus.size());
  return status;
}

// Aligns `sequence` to a graph built from the given chains (each added with
// an empty alignment, i.e. as a new path) plus, when `layer` is given, that
// layer aligned to the chain and inserted: used by tests to check the DP on
// small graphs. Writes (node, position) pairs, returns their count.
int32_t oracle_poa_align(const char* const* chains, const uint32_t* chain_lengths,
                         uint32_t n_chains, const char* layer, uint32_t layer_length,
                         const char* sequence, uint32_t length, int32_t match,
                         int32_t mismatch, int32_t gap, int32_t* out_nodes,
                         int32_t* out_positions, uint32_t capacity, char* node_bases,
                         uint32_t* n_nodes) {
  oracle::PoaGraph g;
  for (uint32_t c = 0; c < n_chains; ++c) {
    std::vector<uint32_t> w(chain_lengths[c], 1);
    g.add_alignment(oracle::Alignment(), chains[c], chain_lengths[c], w);
  }
  if (layer != nullptr && layer_length > 0) {
    auto a = oracle::poa_align_nw(layer, layer_length, g, match, mismatch, gap, nullptr);
    std::vector<uint32_t> w(layer_length, 1);
    g.add_alignment(a, layer, layer_length, w);
  }
  auto a = oracle::poa_align_nw(sequence, length, g, match, mismatch, gap, nullptr);
  *n_nodes = static_cast<uint32_t>(g.nodes().size());
  if (node_bases != nullptr) {
    for (uint32_t i = 0; i < g.nodes().size() && i < capacity; ++i) node_bases[i] = g.nodes()[i].base;
  }
  if (a.size() > capacity) return -1;
  for (size_t i = 0; i < a.size(); ++i) {
    out_nodes[i] = a[i].first;
    out_positions[i] = a[i].second;
  }
  return static_cast<int32_t>(a.size());
}

// Edges of the graph used by oracle_poa_align (same construction), for the
// brute-force path enumeration of the tests: writes tail/head pairs.
int32_t oracle_poa_graph_edges(const char* const* chains, const uint32_t* chain_lengths,
                               uint32_t n_chains, const char* layer, uint32_t layer_length,
                               int32_t match, int32_t mismatch, int32_t gap, int32_t* tails,
                               int32_t* heads, uint32_t capacity) {
  oracle::PoaGraph g;
  for (uint32_t c = 0; c < n_chains; ++c) {
    std::vector<uint32_t> w(chain_lengths[c], 1);
    g.add_alignment(oracle::Alignment(), chains[c], chain_lengths[c], w);
  }
  if (layer != nullptr && layer_length > 0) {
    auto a = oracle::poa_align_nw(layer, layer_length, g, match, mismatch, gap, nullptr);
    std::vector<uint32_t> w(layer_length, 1);
    g.add_alignment(a, layer, layer_length, w);
  }
  if (g.edges().size() > capacity) return -1;
  for (size_t i = 0; i < g.edges().size(); ++i) {
    tails[i] = static_cast<int32_t>(g.edges()[i].tail);
    heads[i] = static_cast<int32_t>(g.edges()[i].head);
  }
  return static_cast<int32_t>(g.edges().size());
}

// racon's layer order (std::sort of indices 1..n by begin, window.cpp:78-85),
// returned as 0-based layer indices.
void oracle_sorted_layer_order(uint32_t n, const uint32_t* begins, uint32_t* out) {
  std::vector<uint32_t> rank(n + 1);
  std::iota(rank.begin(), rank.end(), 0u);
  std::sort(rank.begin() + 1, rank.end(),
            [&](uint32_t l, uint32_t r) { return begins[l - 1] < begins[r - 1]; });
  for (uint32_t i = 0; i < n; ++i) out[i] = rank[i + 1] - 1;
}

}  // extern "C"
