// ORACLE (test infrastructure only; never linked into the product).
//
// CPU restatement of the partial-order-alignment step that racon takes from
// spoa (rvaser/spoa, v3-era API: createAlignmentEngine(kNW, m, x, g),
// createGraph, Graph::add_alignment / subgraph / update_alignment /
// generate_consensus; call sites racon src/window.cpp:65-142 and
// src/polisher.cpp:181-185). spoa is an un-vendored submodule
// (vendor/spoa is empty, no commit pinned), so this file restates its
// published algorithm (Lee, Grasso & Sharlow 2002; heaviest bundle, Lee 2003)
// with spoa's tie-breaking:
//   * linear-gap NW of a sequence against the graph in topological order,
//     first-sink-in-rank-order end cell, traceback preferring diagonal over
//     predecessors in in-edge order, then vertical, then horizontal;
//   * graph insertion with aligned-node sets and edge weights w[i-1] + w[i];
//   * DFS topological sort that emits aligned nodes right after their leader;
//   * subgraph = nodes reached backwards from `end` (in-edges and aligned
//     links) without passing below node id `begin`;
//   * heaviest-bundle consensus with branch completion and per-node coverage.
// What pins each part is listed in tests/test_oracle_poa.py and the top-level
// design notes; the end-to-end pins are racon's test/racon_test.cpp:88-290.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace oracle {

// (graph node id or -1, sequence position or -1) pairs, in alignment order.
using Alignment = std::vector<std::pair<int32_t, int32_t>>;

class PoaGraph {
 public:
  struct Edge {
    uint32_t tail;
    uint32_t head;
    int64_t weight;
    std::vector<uint32_t> labels;
  };
  struct Node {
    char base;
    std::vector<uint32_t> in;       // edge indices, insertion order
    std::vector<uint32_t> out;      // edge indices, insertion order
    std::vector<uint32_t> aligned;  // node ids, insertion order
  };

  // Adds a sequence along `alignment` (empty alignment: a new chain) with
  // per-base weights.
  void add_alignment(const Alignment& alignment, const char* sequence,
                     uint32_t length, const std::vector<uint32_t>& weights);

  // spoa's DFS topological order (rank -> node id).
  void topological_sort();

  // Induced subgraph for a layer spanning backbone ids [begin, end];
  // mapping[subgraph id] = graph id.
  std::unique_ptr<PoaGraph> subgraph(uint32_t begin, uint32_t end,
                                     std::vector<int32_t>& mapping) const;

  // Heaviest-bundle consensus and per-position coverage.
  std::string consensus(std::vector<uint32_t>& coverage);

  const std::vector<Node>& nodes() const { return nodes_; }
  const std::vector<Edge>& edges() const { return edges_; }
  const std::vector<uint32_t>& rank_to_node() const { return rank_to_node_; }
  uint32_t num_sequences() const { return num_sequences_; }

  // Number of distinct sequence labels on the node's in- and out-edges.
  uint32_t node_coverage(uint32_t node) const;

 private:
  uint32_t add_node(char base);
  void add_edge(uint32_t tail, uint32_t head, uint32_t weight);
  int32_t add_chain(const char* sequence, const std::vector<uint32_t>& weights,
                    uint32_t begin, uint32_t end);
  void traverse_heaviest_bundle(std::vector<uint32_t>& path) const;
  uint32_t branch_completion(std::vector<int64_t>& scores,
                             std::vector<int32_t>& predecessors,
                             uint32_t rank) const;

  std::vector<Node> nodes_;
  std::vector<Edge> edges_;
  std::vector<uint32_t> rank_to_node_;
  uint32_t num_sequences_ = 0;
};

// Linear-gap global alignment of `sequence` to `graph` (spoa kNW, gap open =
// gap extend = g). Adds |V| * length to *cells when cells is not null.
Alignment poa_align_nw(const char* sequence, uint32_t length,
                       const PoaGraph& graph, int32_t match, int32_t mismatch,
                       int32_t gap, uint64_t* cells);

// One window, as racon's Window::generate_consensus sees it. Layers are
// already in racon's alignment order (std::sort by begin).
struct WindowInput {
  const char* backbone;
  const char* backbone_quality;  // never null (racon passes '!' x w)
  uint32_t backbone_length;
  struct Layer {
    const char* sequence;
    const char* quality;  // may be null
    uint32_t length;
    uint32_t begin;
    uint32_t end;
  };
  std::vector<Layer> layers;
  bool tgs;
};

struct PoaStats {
  uint64_t cells = 0;        // sum over aligned layers of |V_aligned| * L
  uint64_t graph_nodes = 0;  // final graph size
};

// window.cpp:65-142: returns 0 unpolished (< 3 sequences), 1 polished,
// 2 polished but left untrimmed (racon's "might be chimeric" warning).
uint8_t window_consensus(const WindowInput& window, int32_t match,
                         int32_t mismatch, int32_t gap, bool trim,
                         std::string& consensus, PoaStats* stats);

}  // namespace oracle
