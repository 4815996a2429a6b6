#include "window.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace racon {

std::shared_ptr<Window> createWindow(uint64_t id, uint32_t rank, WindowType type,
                                     const char* backbone, uint32_t backbone_length,
                                     const char* quality, uint32_t quality_length,
                                     uint32_t layers) {
  // window.cpp:15-27
  if (backbone_length == 0 || backbone_length != quality_length) {
    fprintf(stderr, "[racon::createWindow] error: "
                    "empty backbone sequence/unequal quality length!\n");
    exit(1);
  }
  return std::make_shared<Window>(id, rank, type, backbone, backbone_length,
                                  quality, quality_length, layers);
}

Window::Window(uint64_t id, uint32_t rank, WindowType type, const char* backbone,
               uint32_t backbone_length, const char* quality,
               uint32_t quality_length, uint32_t layers)
    : id_(id), rank_(rank), type_(type) {
  if (layers > 0) reserve_layers(layers + 1);
  sequences_.emplace_back(backbone, backbone_length);
  qualities_.emplace_back(quality, quality_length);
  positions_.emplace_back(0, 0);
}

void Window::add_layer(const char* sequence, uint32_t sequence_length,
                       const char* quality, uint32_t quality_length,
                       uint32_t begin, uint32_t end) {
  // window.cpp:42-63
  if (sequence_length == 0 || begin == end) {
    return;
  }
  if (quality != nullptr && sequence_length != quality_length) {
    fprintf(stderr, "[racon::Window::add_layer] error: unequal quality size!\n");
    exit(1);
  }
  if (begin >= end || begin > sequences_.front().second ||
      end > sequences_.front().second) {
    fprintf(stderr, "[racon::Window::add_layer] error: "
                    "layer begin and end positions are invalid!\n");
    exit(1);
  }
  sequences_.emplace_back(sequence, sequence_length);
  qualities_.emplace_back(quality, quality_length);
  positions_.emplace_back(begin, end);
}

std::vector<uint32_t> Window::layer_order() const {
  std::vector<uint32_t> rank;
  rank.reserve(sequences_.size());
  for (uint32_t i = 0; i < sequences_.size(); ++i) {
    rank.emplace_back(i);
  }
  std::sort(rank.begin() + 1, rank.end(), [&](uint32_t lhs, uint32_t rhs) {
    return positions_[lhs].first < positions_[rhs].first;
  });
  return std::vector<uint32_t>(rank.begin() + 1, rank.end());
}

void Window::warn_chimeric() const {
  fprintf(stderr, "[racon::Window::generate_consensus] warning: "
                  "contig %lu might be chimeric in window %u!\n",
          static_cast<unsigned long>(id_), rank_);
}

}  // namespace racon
