// The product's PolishEngine: racon's hot path on the GPU through the C ABI
// of libracon_hip (include/racon_hip.h). Plays the role of racon's
// CUDAPolisher batch objects (src/cuda/cudapolisher.cpp:74-214, 216-418),
// except that every window and overlap runs on the device: there is no CPU
// fallback, and a device failure ends the run with an error.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "engine.hpp"
#include "racon_hip.h"
#include "window.hpp"

namespace racon {

namespace {

[[noreturn]] void die(const char* what, int32_t rc) {
  fprintf(stderr, "[racon::HipEngine] error: %s failed (%d): %s\n", what, rc, rh_last_error());
  exit(1);
}

class HipEngine : public PolishEngine {
 public:
  explicit HipEngine(const EngineConfig& config) : config_(config) {
    if (rh_device_count() < 1) {
      fprintf(stderr, "[racon::HipEngine] error: no HIP device available!\n");
      exit(1);
    }
  }

  std::string name() const override { return "hip-gfx950"; }

  void align_overlaps(
      const std::vector<AlignTask>& tasks,
      const std::function<void(uint64_t, const uint8_t*, uint64_t)>& done) override {
    rh_align_batch* batch = nullptr;
    int32_t rc = rh_align_create(0, &batch);
    if (rc != RH_OK) die("rh_align_create", rc);
    // chunks of about 1 GiB of sequence
    const uint64_t kChunkBytes = 1ull << 30;
    uint64_t first = 0;
    std::vector<rh_pair> pairs;
    std::vector<uint8_t> ops;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> n_ops;
    std::vector<int32_t> distance;
    while (first < tasks.size()) {
      uint64_t last = first, bytes = 0;
      pairs.clear();
      while (last < tasks.size() && (last == first || bytes < kChunkBytes)) {
        const AlignTask& t = tasks[last];
        pairs.push_back(rh_pair{t.query, t.query_length, t.target, t.target_length});
        bytes += t.query_length + t.target_length;
        ++last;
      }
      rc = rh_align_upload(batch, static_cast<uint32_t>(pairs.size()), pairs.data());
      if (rc != RH_OK) die("rh_align_upload", rc);
      rc = rh_align_run(batch);
      if (rc != RH_OK) die("rh_align_run", rc);
      ops.resize(rh_align_output_bound(batch));
      offsets.resize(pairs.size());
      n_ops.resize(pairs.size());
      distance.resize(pairs.size());
      rc = rh_align_download(batch, ops.data(), ops.size(), offsets.data(), n_ops.data(),
                             distance.data());
      if (rc != RH_OK) die("rh_align_download", rc);
      for (uint64_t i = 0; i < pairs.size(); ++i) {
        done(first + i, ops.data() + offsets[i], n_ops[i]);
      }
      first = last;
    }
    rh_align_destroy(batch);
  }

  // The breaking-point walk runs on the device over the operations the
  // aligner left there; only the points come back.
  bool align_breaking_points(
      const std::vector<AlignTask>& tasks, uint32_t window_length,
      const std::function<void(uint64_t, const uint32_t*, uint64_t)>& done) override {
    rh_align_batch* batch = nullptr;
    int32_t rc = rh_align_create(0, &batch);
    if (rc != RH_OK) die("rh_align_create", rc);
    const uint64_t kChunkBytes = 1ull << 30;
    uint64_t first = 0;
    std::vector<rh_pair> pairs;
    std::vector<rh_break_span> spans;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> counts;
    std::vector<uint32_t> points;
    while (first < tasks.size()) {
      uint64_t last = first, bytes = 0;
      pairs.clear();
      spans.clear();
      while (last < tasks.size() && (last == first || bytes < kChunkBytes)) {
        const AlignTask& t = tasks[last];
        pairs.push_back(rh_pair{t.query, t.query_length, t.target, t.target_length});
        spans.push_back(rh_break_span{t.t_begin, t.t_end, t.q_first});
        bytes += t.query_length + t.target_length;
        ++last;
      }
      const uint32_t n = static_cast<uint32_t>(pairs.size());
      rc = rh_align_upload(batch, n, pairs.data());
      if (rc != RH_OK) die("rh_align_upload", rc);
      rc = rh_align_run(batch);
      if (rc != RH_OK) die("rh_align_run", rc);
      points.resize(2 * rh_breaking_points_bound(window_length, spans.data(), n));
      offsets.resize(n);
      counts.resize(n);
      rc = rh_align_breaking_points(batch, window_length, spans.data(), offsets.data(),
                                    counts.data(), points.data(), points.size() / 2);
      if (rc != RH_OK) die("rh_align_breaking_points", rc);
      for (uint32_t i = 0; i < n; ++i) {
        done(first + i, points.data() + 2 * offsets[i], counts[i]);
      }
      first = last;
    }
    rh_align_destroy(batch);
    return true;
  }

  void generate_consensus(std::vector<std::shared_ptr<Window>>& windows, bool trim,
                          std::vector<uint8_t>& status) override {
    status.assign(windows.size(), kWindowUnpolished);
    rh_poa_batch* batch = nullptr;
    int32_t rc = rh_poa_create(0, config_.match, config_.mismatch, config_.gap, trim ? 1 : 0,
                               &batch);
    if (rc != RH_OK) die("rh_poa_create", rc);
    const uint64_t kChunkBytes = 1ull << 30;
    std::vector<rh_window> batch_windows;
    std::vector<std::vector<rh_layer>> layers;
    std::vector<char> out;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> lengths;
    std::vector<uint8_t> codes;
    uint64_t first = 0;
    while (first < windows.size()) {
      uint64_t last = first, bytes = 0;
      batch_windows.clear();
      layers.clear();
      while (last < windows.size() && (last == first || bytes < kChunkBytes)) {
        const Window& w = *windows[last];
        std::vector<rh_layer> ls;
        ls.reserve(w.size() - 1);
        for (uint32_t k = 1; k < w.size(); ++k) {
          ls.push_back(rh_layer{w.sequence(k).first, w.quality(k).first, w.sequence(k).second,
                                w.position(k).first, w.position(k).second});
          bytes += 2ull * w.sequence(k).second;
        }
        bytes += 2ull * w.sequence(0).second;
        layers.push_back(std::move(ls));
        ++last;
      }
      for (uint64_t i = first; i < last; ++i) {
        const Window& w = *windows[i];
        const auto& ls = layers[i - first];
        batch_windows.push_back(rh_window{w.sequence(0).first, w.quality(0).first,
                                          w.sequence(0).second,
                                          static_cast<uint32_t>(ls.size()), ls.data(),
                                          w.type() == WindowType::kTGS ? 1u : 0u});
      }
      rc = rh_poa_upload(batch, static_cast<uint32_t>(batch_windows.size()),
                         batch_windows.data());
      if (rc != RH_OK) die("rh_poa_upload", rc);
      rc = rh_poa_run(batch, RH_POA_BATCHED);
      if (rc != RH_OK) die("rh_poa_run", rc);
      out.resize(rh_poa_output_bound(batch));
      offsets.resize(batch_windows.size());
      lengths.resize(batch_windows.size());
      codes.resize(batch_windows.size());
      rc = rh_poa_download(batch, out.data(), out.size(), offsets.data(), lengths.data(),
                           codes.data());
      if (rc != RH_OK) die("rh_poa_download", rc);
      for (uint64_t i = first; i < last; ++i) {
        const uint64_t k = i - first;
        windows[i]->set_consensus(std::string(out.data() + offsets[k], lengths[k]));
        status[i] = codes[k] == RH_WINDOW_UNPOLISHED
                        ? kWindowUnpolished
                        : (codes[k] == RH_WINDOW_POLISHED_UNTRIMMED ? kWindowPolishedChimeric
                                                                   : kWindowPolished);
      }
      first = last;
    }
    rh_poa_destroy(batch);
  }

 private:
  EngineConfig config_;
};

}  // namespace

std::unique_ptr<PolishEngine> createEngine(const EngineConfig& config) {
  return std::unique_ptr<PolishEngine>(new HipEngine(config));
}

}  // namespace racon
