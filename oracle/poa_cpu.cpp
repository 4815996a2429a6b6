// ORACLE (test infrastructure only). See poa_cpu.hpp for scope and sources.
#include "poa_cpu.hpp"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <unordered_set>

namespace oracle {

namespace {
constexpr int32_t kNegativeInfinity = INT32_MIN + 1024;
}

uint32_t PoaGraph::add_node(char base) {
  nodes_.push_back(Node{base, {}, {}, {}});
  return static_cast<uint32_t>(nodes_.size() - 1);
}

void PoaGraph::add_edge(uint32_t tail, uint32_t head, uint32_t weight) {
  // an existing tail->head edge gains the label and the weight
  for (uint32_t e : nodes_[tail].out) {
    if (edges_[e].head == head) {
      edges_[e].labels.push_back(num_sequences_);
      edges_[e].weight += weight;
      return;
    }
  }
  edges_.push_back(Edge{tail, head, static_cast<int64_t>(weight), {num_sequences_}});
  const uint32_t e = static_cast<uint32_t>(edges_.size() - 1);
  nodes_[tail].out.push_back(e);
  nodes_[head].in.push_back(e);
}

int32_t PoaGraph::add_chain(const char* sequence, const std::vector<uint32_t>& weights,
                            uint32_t begin, uint32_t end) {
  if (begin == end) {
    return -1;
  }
  int32_t first = static_cast<int32_t>(add_node(sequence[begin]));
  for (uint32_t i = begin + 1; i < end; ++i) {
    uint32_t node = add_node(sequence[i]);
    add_edge(node - 1, node, weights[i - 1] + weights[i]);
  }
  return first;
}

void PoaGraph::add_alignment(const Alignment& alignment, const char* sequence,
                             uint32_t length, const std::vector<uint32_t>& weights) {
  if (length == 0) {
    return;
  }
  if (weights.size() != length) {
    fprintf(stderr, "[oracle::PoaGraph] error: unequal weights size!\n");
    exit(1);
  }
  if (alignment.empty()) {
    add_chain(sequence, weights, 0, length);
    ++num_sequences_;
    topological_sort();
    return;
  }

  std::vector<uint32_t> valid;
  for (const auto& it : alignment) {
    if (it.second != -1) {
      valid.push_back(static_cast<uint32_t>(it.second));
    }
  }

  const uint32_t before = static_cast<uint32_t>(nodes_.size());
  add_chain(sequence, weights, 0, valid.front());
  int32_t head = before == nodes_.size() ? -1 : static_cast<int32_t>(nodes_.size() - 1);
  int32_t tail = add_chain(sequence, weights, valid.back() + 1, length);

  uint32_t prev_weight = head == -1 ? 0 : weights[valid.front() - 1];
  for (const auto& it : alignment) {
    if (it.second == -1) {
      continue;
    }
    const char letter = sequence[it.second];
    int32_t node;
    if (it.first == -1) {
      node = static_cast<int32_t>(add_node(letter));
    } else if (nodes_[it.first].base == letter) {
      node = it.first;
    } else {
      int32_t same = -1;
      for (uint32_t a : nodes_[it.first].aligned) {
        if (nodes_[a].base == letter) {
          same = static_cast<int32_t>(a);
          break;
        }
      }
      if (same == -1) {
        node = static_cast<int32_t>(add_node(letter));
        const std::vector<uint32_t> group = nodes_[it.first].aligned;
        for (uint32_t a : group) {
          nodes_[node].aligned.push_back(a);
          nodes_[a].aligned.push_back(static_cast<uint32_t>(node));
        }
        nodes_[node].aligned.push_back(static_cast<uint32_t>(it.first));
        nodes_[it.first].aligned.push_back(static_cast<uint32_t>(node));
      } else {
        node = same;
      }
    }
    if (head != -1) {
      add_edge(static_cast<uint32_t>(head), static_cast<uint32_t>(node),
               prev_weight + weights[it.second]);
    }
    head = node;
    prev_weight = weights[it.second];
  }
  if (tail != -1) {
    add_edge(static_cast<uint32_t>(head), static_cast<uint32_t>(tail),
             prev_weight + weights[valid.back() + 1]);
  }
  ++num_sequences_;
  topological_sort();
}

void PoaGraph::topological_sort() {
  rank_to_node_.clear();
  const uint32_t n = static_cast<uint32_t>(nodes_.size());
  // 0 unvisited, 1 on the stack, 2 done
  std::vector<uint8_t> mark(n, 0);
  std::vector<bool> lead(n, true);  // emits its aligned set when done
  std::vector<uint32_t> stack;
  for (uint32_t root = 0; root < n; ++root) {
    if (mark[root] != 0) {
      continue;
    }
    stack.push_back(root);
    while (!stack.empty()) {
      const uint32_t v = stack.back();
      bool ready = true;
      if (mark[v] != 2) {
        for (uint32_t e : nodes_[v].in) {
          const uint32_t u = edges_[e].tail;
          if (mark[u] != 2) {
            stack.push_back(u);
            ready = false;
          }
        }
        if (lead[v]) {
          for (uint32_t a : nodes_[v].aligned) {
            if (mark[a] != 2) {
              stack.push_back(a);
              lead[a] = false;
              ready = false;
            }
          }
        }
        if (ready) {
          mark[v] = 2;
          if (lead[v]) {
            rank_to_node_.push_back(v);
            for (uint32_t a : nodes_[v].aligned) {
              rank_to_node_.push_back(a);
            }
          }
        } else {
          mark[v] = 1;
        }
      }
      if (ready) {
        stack.pop_back();
      }
    }
  }
  if (rank_to_node_.size() != n) {
    fprintf(stderr, "[oracle::PoaGraph] error: topological sort lost nodes!\n");
    exit(1);
  }
}

std::unique_ptr<PoaGraph> PoaGraph::subgraph(uint32_t begin, uint32_t end,
                                             std::vector<int32_t>& mapping) const {
  const uint32_t n = static_cast<uint32_t>(nodes_.size());
  std::vector<bool> keep(n, false);
  std::vector<uint32_t> stack{end};
  while (!stack.empty()) {
    const uint32_t v = stack.back();
    stack.pop_back();
    if (!keep[v] && v >= begin) {
      for (uint32_t e : nodes_[v].in) {
        stack.push_back(edges_[e].tail);
      }
      for (uint32_t a : nodes_[v].aligned) {
        stack.push_back(a);
      }
      keep[v] = true;
    }
  }

  std::unique_ptr<PoaGraph> sub(new PoaGraph());
  sub->num_sequences_ = num_sequences_;
  mapping.assign(n, -1);
  std::vector<int32_t> to_sub(n, -1);
  for (uint32_t v = 0; v < n; ++v) {
    if (!keep[v]) {
      continue;
    }
    const uint32_t s = sub->add_node(nodes_[v].base);
    to_sub[v] = static_cast<int32_t>(s);
    mapping[s] = static_cast<int32_t>(v);
  }
  for (uint32_t v = 0; v < n; ++v) {
    if (!keep[v]) {
      continue;
    }
    const uint32_t s = static_cast<uint32_t>(to_sub[v]);
    for (uint32_t e : nodes_[v].in) {
      const int32_t t = to_sub[edges_[e].tail];
      if (t == -1) {
        continue;
      }
      sub->add_edge(static_cast<uint32_t>(t), s, static_cast<uint32_t>(edges_[e].weight));
    }
    for (uint32_t a : nodes_[v].aligned) {
      if (to_sub[a] == -1) {
        continue;
      }
      sub->nodes_[s].aligned.push_back(static_cast<uint32_t>(to_sub[a]));
    }
  }
  sub->topological_sort();
  return sub;
}

uint32_t PoaGraph::node_coverage(uint32_t node) const {
  std::unordered_set<uint32_t> labels;
  for (uint32_t e : nodes_[node].in) {
    labels.insert(edges_[e].labels.begin(), edges_[e].labels.end());
  }
  for (uint32_t e : nodes_[node].out) {
    labels.insert(edges_[e].labels.begin(), edges_[e].labels.end());
  }
  return static_cast<uint32_t>(labels.size());
}

uint32_t PoaGraph::branch_completion(std::vector<int64_t>& scores,
                                     std::vector<int32_t>& predecessors,
                                     uint32_t rank) const {
  const uint32_t v = rank_to_node_[rank];
  for (uint32_t e : nodes_[v].out) {
    for (uint32_t f : nodes_[edges_[e].head].in) {
      if (edges_[f].tail != v) {
        scores[edges_[f].tail] = -1;
      }
    }
  }
  int64_t max_score = 0;
  uint32_t max_node = 0;
  for (uint32_t r = rank + 1; r < rank_to_node_.size(); ++r) {
    const uint32_t u = rank_to_node_[r];
    scores[u] = -1;
    predecessors[u] = -1;
    for (uint32_t e : nodes_[u].in) {
      const uint32_t t = edges_[e].tail;
      if (scores[t] == -1) {
        continue;
      }
      if (scores[u] < edges_[e].weight ||
          (scores[u] == edges_[e].weight && scores[predecessors[u]] <= scores[t])) {
        scores[u] = edges_[e].weight;
        predecessors[u] = static_cast<int32_t>(t);
      }
    }
    if (predecessors[u] != -1) {
      scores[u] += scores[predecessors[u]];
    }
    if (max_score < scores[u]) {
      max_score = scores[u];
      max_node = u;
    }
  }
  return max_node;
}

void PoaGraph::traverse_heaviest_bundle(std::vector<uint32_t>& path) const {
  const uint32_t n = static_cast<uint32_t>(nodes_.size());
  std::vector<int32_t> predecessors(n, -1);
  std::vector<int64_t> scores(n, -1);
  uint32_t max_node = 0;
  for (uint32_t v : rank_to_node_) {
    for (uint32_t e : nodes_[v].in) {
      const uint32_t t = edges_[e].tail;
      if (scores[v] < edges_[e].weight ||
          (scores[v] == edges_[e].weight && scores[predecessors[v]] <= scores[t])) {
        scores[v] = edges_[e].weight;
        predecessors[v] = static_cast<int32_t>(t);
      }
    }
    if (predecessors[v] != -1) {
      scores[v] += scores[predecessors[v]];
    }
    if (scores[max_node] < scores[v]) {
      max_node = v;
    }
  }
  if (!nodes_[max_node].out.empty()) {
    std::vector<uint32_t> node_to_rank(n, 0);
    for (uint32_t r = 0; r < n; ++r) {
      node_to_rank[rank_to_node_[r]] = r;
    }
    while (!nodes_[max_node].out.empty()) {
      max_node = branch_completion(scores, predecessors, node_to_rank[max_node]);
    }
  }
  path.clear();
  while (predecessors[max_node] != -1) {
    path.push_back(max_node);
    max_node = static_cast<uint32_t>(predecessors[max_node]);
  }
  path.push_back(max_node);
  std::reverse(path.begin(), path.end());
}

std::string PoaGraph::consensus(std::vector<uint32_t>& coverage) {
  std::vector<uint32_t> path;
  traverse_heaviest_bundle(path);
  std::string out;
  coverage.clear();
  for (uint32_t v : path) {
    out += nodes_[v].base;
    uint32_t total = node_coverage(v);
    for (uint32_t a : nodes_[v].aligned) {
      total += node_coverage(a);
    }
    coverage.push_back(total);
  }
  return out;
}

Alignment poa_align_nw(const char* sequence, uint32_t length, const PoaGraph& graph,
                       int32_t match, int32_t mismatch, int32_t gap, uint64_t* cells) {
  const auto& nodes = graph.nodes();
  const auto& edges = graph.edges();
  const auto& order = graph.rank_to_node();
  if (nodes.empty() || length == 0) {
    return Alignment();
  }
  const uint32_t V = static_cast<uint32_t>(nodes.size());
  const uint32_t W = length + 1;
  if (cells != nullptr) {
    *cells += static_cast<uint64_t>(V) * length;
  }
  std::vector<uint32_t> row_of(V);  // node id -> DP row (rank + 1)
  for (uint32_t r = 0; r < V; ++r) {
    row_of[order[r]] = r + 1;
  }
  std::vector<int32_t> H(static_cast<size_t>(V + 1) * W);
  auto at = [&](uint32_t i, uint32_t j) -> int32_t& { return H[static_cast<size_t>(i) * W + j]; };
  auto score = [&](uint32_t v, uint32_t j) -> int32_t {
    return nodes[v].base == sequence[j - 1] ? match : mismatch;
  };
  auto pred_row = [&](uint32_t v, uint32_t k) -> uint32_t {
    return row_of[edges[nodes[v].in[k]].tail];
  };

  at(0, 0) = 0;
  for (uint32_t j = 1; j < W; ++j) {
    at(0, j) = static_cast<int32_t>(j) * gap;
  }
  for (uint32_t r = 0; r < V; ++r) {
    const uint32_t v = order[r];
    if (nodes[v].in.empty()) {
      at(r + 1, 0) = gap;
    } else {
      int32_t best = kNegativeInfinity;
      for (uint32_t k = 0; k < nodes[v].in.size(); ++k) {
        best = std::max(best, at(pred_row(v, k), 0));
      }
      at(r + 1, 0) = best + gap;
    }
  }

  int32_t max_score = kNegativeInfinity;
  int32_t max_i = -1, max_j = -1;
  for (uint32_t r = 0; r < V; ++r) {
    const uint32_t v = order[r];
    const uint32_t i = r + 1;
    const uint32_t p0 = nodes[v].in.empty() ? 0 : pred_row(v, 0);
    for (uint32_t j = 1; j < W; ++j) {
      at(i, j) = std::max(at(p0, j - 1) + score(v, j), at(p0, j) + gap);
    }
    for (uint32_t k = 1; k < nodes[v].in.size(); ++k) {
      const uint32_t p = pred_row(v, k);
      for (uint32_t j = 1; j < W; ++j) {
        at(i, j) = std::max(at(p, j - 1) + score(v, j), std::max(at(i, j), at(p, j) + gap));
      }
    }
    for (uint32_t j = 1; j < W; ++j) {
      at(i, j) = std::max(at(i, j - 1) + gap, at(i, j));
    }
    if (nodes[v].out.empty() && max_score < at(i, W - 1)) {
      max_score = at(i, W - 1);
      max_i = static_cast<int32_t>(i);
      max_j = static_cast<int32_t>(W - 1);
    }
  }
  if (max_i == -1 && max_j == -1) {
    return Alignment();
  }

  Alignment alignment;
  uint32_t i = static_cast<uint32_t>(max_i), j = static_cast<uint32_t>(max_j);
  while (!(i == 0 && j == 0)) {
    const int32_t h = at(i, j);
    bool found = false;
    uint32_t prev_i = 0, prev_j = 0;
    if (i != 0 && j != 0) {
      const uint32_t v = order[i - 1];
      const int32_t s = score(v, j);
      const uint32_t n_in = static_cast<uint32_t>(nodes[v].in.size());
      for (uint32_t k = 0; k < std::max(n_in, 1u); ++k) {
        const uint32_t p = n_in == 0 ? 0 : pred_row(v, k);
        if (h == at(p, j - 1) + s) {
          prev_i = p;
          prev_j = j - 1;
          found = true;
          break;
        }
      }
    }
    if (!found && i != 0) {
      const uint32_t v = order[i - 1];
      const uint32_t n_in = static_cast<uint32_t>(nodes[v].in.size());
      for (uint32_t k = 0; k < std::max(n_in, 1u); ++k) {
        const uint32_t p = n_in == 0 ? 0 : pred_row(v, k);
        if (h == at(p, j) + gap) {
          prev_i = p;
          prev_j = j;
          found = true;
          break;
        }
      }
    }
    if (!found && j != 0 && h == at(i, j - 1) + gap) {
      prev_i = i;
      prev_j = j - 1;
      found = true;
    }
    if (!found) {
      fprintf(stderr, "[oracle::poa_align_nw] error: traceback lost its path!\n");
      exit(1);
    }
    alignment.emplace_back(i == prev_i ? -1 : static_cast<int32_t>(order[i - 1]),
                           j == prev_j ? -1 : static_cast<int32_t>(j - 1));
    i = prev_i;
    j = prev_j;
  }
  std::reverse(alignment.begin(), alignment.end());
  return alignment;
}

uint8_t window_consensus(const WindowInput& window, int32_t match, int32_t mismatch,
                         int32_t gap, bool trim, std::string& consensus,
                         PoaStats* stats) {
  const uint32_t n_seqs = static_cast<uint32_t>(window.layers.size()) + 1;
  if (n_seqs < 3) {
    consensus.assign(window.backbone, window.backbone_length);
    return 0;
  }
  PoaGraph graph;
  {
    std::vector<uint32_t> w(window.backbone_length);
    for (uint32_t i = 0; i < window.backbone_length; ++i) {
      w[i] = static_cast<uint32_t>(window.backbone_quality[i] - 33);
    }
    graph.add_alignment(Alignment(), window.backbone, window.backbone_length, w);
  }
  const uint32_t offset = static_cast<uint32_t>(0.01 * window.backbone_length);
  std::vector<uint32_t> w;
  for (const auto& layer : window.layers) {
    Alignment alignment;
    if (layer.begin < offset && layer.end > window.backbone_length - offset) {
      alignment = poa_align_nw(layer.sequence, layer.length, graph, match, mismatch,
                               gap, stats ? &stats->cells : nullptr);
    } else {
      std::vector<int32_t> mapping;
      auto sub = graph.subgraph(layer.begin, layer.end, mapping);
      alignment = poa_align_nw(layer.sequence, layer.length, *sub, match, mismatch,
                               gap, stats ? &stats->cells : nullptr);
      for (auto& it : alignment) {
        if (it.first != -1) {
          it.first = mapping[it.first];
        }
      }
    }
    w.assign(layer.length, 1);
    if (layer.quality != nullptr) {
      for (uint32_t i = 0; i < layer.length; ++i) {
        w[i] = static_cast<uint32_t>(layer.quality[i] - 33);
      }
    }
    graph.add_alignment(alignment, layer.sequence, layer.length, w);
  }

  std::vector<uint32_t> coverages;
  consensus = graph.consensus(coverages);
  if (stats != nullptr) {
    stats->graph_nodes = graph.nodes().size();
  }

  uint8_t status = 1;
  if (window.tgs && trim) {
    // window.cpp:118-139
    const uint32_t average_coverage = (n_seqs - 1) / 2;
    int32_t begin = 0, end = static_cast<int32_t>(consensus.size()) - 1;
    for (; begin < static_cast<int32_t>(consensus.size()); ++begin) {
      if (coverages[begin] >= average_coverage) {
        break;
      }
    }
    for (; end >= 0; --end) {
      if (coverages[end] >= average_coverage) {
        break;
      }
    }
    if (begin >= end) {
      status = 2;
    } else {
      consensus = consensus.substr(begin, end - begin + 1);
    }
  }
  return status;
}

}  // namespace oracle
