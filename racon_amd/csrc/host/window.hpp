// Window of a target: backbone plus the read layers that cover it.
//
// Construction and add_layer follow racon src/window.cpp:15-63. The
// consensus itself (racon's Window::generate_consensus, window.cpp:65-142:
// POA alignment of each layer, graph insertion, heaviest-bundle consensus and
// trimming) is produced by the polishing engine for whole batches of windows;
// the window keeps racon's layer ordering (std::sort on the layer begin,
// window.cpp:78-85) so every engine sees the layers in the same order.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace racon {

enum class WindowType {
  kNGS,  // Next Generation Sequencing
  kTGS   // Third Generation Sequencing
};

class Window;
// layers: room for that many add_layer calls (0: grow as they come).
std::shared_ptr<Window> createWindow(uint64_t id, uint32_t rank, WindowType type,
                                     const char* backbone, uint32_t backbone_length,
                                     const char* quality, uint32_t quality_length,
                                     uint32_t layers = 0);

class Window {
 public:
  Window(uint64_t id, uint32_t rank, WindowType type, const char* backbone,
         uint32_t backbone_length, const char* quality, uint32_t quality_length,
         uint32_t layers = 0);
  Window(const Window&) = delete;
  Window& operator=(const Window&) = delete;

  uint64_t id() const { return id_; }
  uint32_t rank() const { return rank_; }
  WindowType type() const { return type_; }
  const std::string& consensus() const { return consensus_; }
  void set_consensus(std::string consensus) { consensus_ = std::move(consensus); }

  void add_layer(const char* sequence, uint32_t sequence_length,
                 const char* quality, uint32_t quality_length, uint32_t begin,
                 uint32_t end);
  // Room for n more layers (no reallocation while they are added).
  void reserve_layers(uint32_t n) {
    sequences_.reserve(sequences_.size() + n);
    qualities_.reserve(qualities_.size() + n);
    positions_.reserve(positions_.size() + n);
  }

  // Number of sequences (backbone + layers).
  uint32_t size() const { return static_cast<uint32_t>(sequences_.size()); }
  const std::pair<const char*, uint32_t>& sequence(uint32_t i) const {
    return sequences_[i];
  }
  const std::pair<const char*, uint32_t>& quality(uint32_t i) const {
    return qualities_[i];
  }
  const std::pair<uint32_t, uint32_t>& position(uint32_t i) const {
    return positions_[i];
  }

  // Indices 1..size()-1 in racon's alignment order (window.cpp:78-85):
  // std::sort of the layer indices by begin position (not stable).
  std::vector<uint32_t> layer_order() const;

  // The short-circuit of window.cpp:68-71: fewer than 3 sequences means the
  // consensus is the backbone and the window counts as unpolished.
  bool too_shallow() const { return sequences_.size() < 3; }

  // Warning racon prints when trimming would empty the consensus
  // (window.cpp:133-136).
  void warn_chimeric() const;

 private:
  uint64_t id_;
  uint32_t rank_;
  WindowType type_;
  std::string consensus_;
  std::vector<std::pair<const char*, uint32_t>> sequences_;
  std::vector<std::pair<const char*, uint32_t>> qualities_;
  std::vector<std::pair<uint32_t, uint32_t>> positions_;
};

}  // namespace racon
