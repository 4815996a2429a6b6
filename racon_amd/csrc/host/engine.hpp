// The hot-path interface of the polisher.
//
// racon runs its hot path item by item on a thread pool: one edlib NW
// alignment per overlap (Polisher::find_overlap_breaking_points,
// polisher.cpp:462-484) and one spoa consensus per window
// (Polisher::polish -> Window::generate_consensus, polisher.cpp:486-504,
// window.cpp:65-142). A PolishEngine takes the same work as whole batches so
// that a device engine can keep thousands of items in flight.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace racon {

class Window;

// One pairwise global alignment (edlib's query/target orientation).
struct AlignTask {
  const char* query;
  uint32_t query_length;
  const char* target;
  uint32_t target_length;
  // where the alignment sits, for its breaking points (overlap.cpp:228-240):
  // target span [t_begin, t_end) on the contig and the query position the
  // walk starts after
  uint32_t t_begin = 0;
  uint32_t t_end = 0;
  uint32_t q_first = 0;
};

// Where the layers of one aligned overlap go (polisher.cpp:403-457): the
// overlap, the read's length, the first window of the overlap's target, and
// the read's quality bytes in the job's quality buffer (kNoQuality: none),
// mirrored when those bytes run opposite to the aligned query (the forward
// quality of a '-' overlap or the reversed one of a '+' overlap).
constexpr uint64_t kNoQuality = ~0ull;
struct LayerSource {
  uint32_t overlap = 0;
  uint32_t query_length = 0;
  uint64_t first_window = 0;
  uint64_t quality = kNoQuality;
  uint32_t mirrored = 0;
  uint32_t reserved = 0;
};

// One window layer: query bytes [q_begin, q_end) of the overlap (in its
// aligned orientation) over backbone positions begin..end (inclusive) of
// window `window` (the arguments of Window::add_layer, polisher.cpp:431-453).
struct LayerRecord {
  uint32_t window;
  uint32_t overlap;
  uint32_t q_begin;
  uint32_t q_end;
  uint32_t begin;
  uint32_t end;
};

// Scores and switches shared by every engine (racon CLI defaults:
// main.cpp:51-66).
struct EngineConfig {
  int8_t match = 3;
  int8_t mismatch = -5;
  int8_t gap = -4;
  uint32_t num_threads = 1;
  uint32_t window_length = 500;
  // device engine knobs (racon's -c/--cudapoa-batches and
  // --cudaaligner-batches select the device path in the reference)
  uint32_t poa_batches = 0;
  uint32_t aligner_batches = 0;
  bool banded = false;
  uint32_t aligner_band_width = 0;
  // a device serving loop shared by several polishers (racon --serve): the
  // windows go to it as requests instead of to batches of their own
  void* poa_server = nullptr;
};

// Per-window outcome of generate_consensus.
enum : uint8_t {
  kWindowUnpolished = 0,  // < 3 sequences: consensus is the backbone
  kWindowPolished = 1,
  kWindowPolishedChimeric = 2,  // polished, trimming skipped (racon warns)
};

class PolishEngine {
 public:
  virtual ~PolishEngine() = default;

  virtual std::string name() const = 0;

  // Aligns every task (edlib NW, path) and calls done(i, ops, n_ops) once per
  // task with the edit operations in alignment order. done() may be called
  // from several threads, but never twice for the same i.
  virtual void align_overlaps(
      const std::vector<AlignTask>& tasks,
      const std::function<void(uint64_t, const uint8_t*, uint64_t)>& done) = 0;

  // Aligns every task and calls done(i, points, n_points) once per task
  // with its breaking points as (target, query) pairs, two per window of
  // window_length target bases with a match (overlap.cpp:226-292), computed
  // by the engine itself. Returns false, calling nothing, when the engine
  // leaves that walk to the caller (over align_overlaps' operations).
  virtual bool align_breaking_points(
      const std::vector<AlignTask>& tasks, uint32_t window_length,
      const std::function<void(uint64_t, const uint32_t*, uint64_t)>& done) {
    (void)tasks;
    (void)window_length;
    (void)done;
    return false;
  }

  // True when align_layers does the work (checked before any is done).
  virtual bool device_layers() const { return false; }

  // Aligns every task (sources[i] belongs to tasks[i]) and turns the
  // breaking points into window layers with racon's filters (query span at
  // least 2% of window_length, mean quality at least quality_threshold when
  // the read has qualities): layers come back grouped by window, in task
  // order within a window, windows [0, n_windows). gather_qualities fills
  // the quality bytes the sources point into; the engine may call it on
  // another thread while the first alignments run. Only called when
  // device_layers() is true.
  virtual void align_layers(const std::vector<AlignTask>& tasks,
                            const std::vector<LayerSource>& sources, uint32_t window_length,
                            double quality_threshold, uint64_t n_windows,
                            const std::function<void(std::vector<char>*)>& gather_qualities,
                            std::vector<LayerRecord>* layers) {
    (void)tasks;
    (void)sources;
    (void)window_length;
    (void)quality_threshold;
    (void)n_windows;
    (void)gather_qualities;
    (void)layers;
  }

  // Generates the consensus of every window (set through
  // Window::set_consensus) and fills status[i] with kWindow* codes. ready(n)
  // is called (one call at a time, n increasing, the last call with
  // windows.size(), all before generate_consensus returns; possibly from a
  // helper thread) once windows [0, n) and their status are final, so the
  // caller can emit finished targets while later windows still run.
  virtual void generate_consensus(std::vector<std::shared_ptr<Window>>& windows,
                                  bool trim, std::vector<uint8_t>& status,
                                  const std::function<void(uint64_t)>& ready) = 0;
};

// Defined once per executable: the product links the HIP engine, the oracle
// build links the CPU restatement.
std::unique_ptr<PolishEngine> createEngine(const EngineConfig& config);

}  // namespace racon
