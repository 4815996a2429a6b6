// Polisher: loads targets, reads and overlaps, builds windows and stitches
// the polished targets (racon src/polisher.hpp:36-98, src/polisher.cpp).
//
// The hot path (overlap alignment and window consensus) is delegated to a
// PolishEngine, which is the drop-in boundary: the product binary links the
// HIP engine, the oracle binary the CPU restatement, and both share this
// driver so the output format and the host-side filters are identical.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"
#include "io.hpp"

namespace racon {

class Sequence;
class Overlap;
class Window;
class Logger;
class ThreadPool;

enum class PolisherType {
  kC,  // Contig polishing
  kF   // Fragment error correction
};

class Polisher;

// racon's factory (polisher.hpp:42-48, polisher.cpp:55-160). The device
// switches keep racon's names; which engine runs is fixed by the executable
// (see createEngine in engine.hpp).
std::unique_ptr<Polisher> createPolisher(
    const std::string& sequences_path, const std::string& overlaps_path,
    const std::string& target_path, PolisherType type, uint32_t window_length,
    double quality_threshold, double error_threshold, bool trim, int8_t match,
    int8_t mismatch, int8_t gap, uint32_t num_threads, uint32_t cuda_batches = 0,
    bool cuda_banded_alignment = false, uint32_t cudaaligner_batches = 0,
    uint32_t cudaaligner_band_width = 0, void* poa_server = nullptr);

// Phase timings of the last run (seconds), for reports.
struct PolisherTimings {
  double load = 0;
  double align = 0;      // overlap alignment + breaking points
  double windows = 0;    // window/layer construction
  double consensus = 0;  // window consensus (engine)
  double stitch = 0;
  uint64_t num_windows = 0;
  uint64_t num_overlaps_aligned = 0;
};

class Polisher {
 public:
  Polisher(const std::string& sequences_path, SequenceFormat sequences_format,
           const std::string& overlaps_path, OverlapFormat overlaps_format,
           const std::string& target_path, SequenceFormat target_format,
           PolisherType type, uint32_t window_length, double quality_threshold,
           double error_threshold, bool trim, uint32_t num_threads,
           std::unique_ptr<PolishEngine> engine);
  virtual ~Polisher();

  Polisher(const Polisher&) = delete;
  Polisher& operator=(const Polisher&) = delete;

  virtual void initialize();
  virtual void polish(std::vector<std::unique_ptr<Sequence>>& dst,
                      bool drop_unpolished_sequences);
  // polish() with every finished target handed to sink as soon as its
  // windows are done (racon's stitching, src/polisher.cpp:506-538, streamed
  // per target): the same sequences in the same order.
  void polish_streamed(bool drop_unpolished_sequences,
                       const std::function<void(std::unique_ptr<Sequence>)>& sink);

  const PolisherTimings& timings() const { return timings_; }

  // Writes the windows (after initialize) or the overlap alignment inputs
  // (during initialize) to binary files for the benchmark driver.
  void set_dump_paths(const std::string& windows_path, const std::string& overlaps_path) {
    dump_windows_path_ = windows_path;
    dump_overlaps_path_ = overlaps_path;
  }
  void dump_windows() const;
  PolishEngine& engine() { return *engine_; }

 protected:
  virtual void find_overlap_breaking_points(
      std::vector<std::unique_ptr<Overlap>>& overlaps);
  // The alignment inputs of the overlaps that need one (and the
  // --dump-overlaps file).
  void collect_alignment_tasks(std::vector<std::unique_ptr<Overlap>>& overlaps,
                               std::vector<AlignTask>* tasks, std::vector<uint64_t>* task_overlap);
  // Aligns the overlaps and extracts the window layers on the engine's
  // device (polisher.cpp:403-457) when the engine can and every overlap needs
  // an alignment; false (nothing done) otherwise.
  bool align_layers_on_device(std::vector<std::unique_ptr<Overlap>>& overlaps,
                              uint64_t targets_size, std::vector<LayerRecord>* layers);

  std::string sequences_path_;
  SequenceFormat sequences_format_;
  std::string overlaps_path_;
  OverlapFormat overlaps_format_;
  std::string target_path_;
  SequenceFormat target_format_;

  PolisherType type_;
  double quality_threshold_;
  double error_threshold_;
  bool trim_;

  std::vector<std::unique_ptr<Sequence>> sequences_;
  std::vector<uint32_t> targets_coverages_;
  std::string dummy_quality_;

  uint32_t window_length_;
  std::vector<std::shared_ptr<Window>> windows_;

  std::unique_ptr<ThreadPool> thread_pool_;
  std::unique_ptr<Logger> logger_;
  std::unique_ptr<PolishEngine> engine_;
  PolisherTimings timings_;
  std::string dump_windows_path_;
  std::string dump_overlaps_path_;
};

}  // namespace racon
