// The product's PolishEngine: racon's hot path on the GPU through the C ABI
// of libracon_hip (include/racon_hip.h). Plays the role of racon's
// CUDAPolisher batch objects (src/cuda/cudapolisher.cpp:74-214, 216-418),
// except that every window and overlap runs on the device: there is no CPU
// fallback, and a device failure ends the run with an error.
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
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
  explicit HipEngine(const EngineConfig& config)
      : config_(config), chunk_log_(getenv("RACON_AMD_CHUNK_LOG") != nullptr) {
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
    rc = rh_align_set_band(batch, config_.aligner_band_width);
    if (rc != RH_OK) die("rh_align_set_band", rc);
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
        if (distance[i] < 0) {
          fprintf(stderr,
                  "[racon::HipEngine] error: overlap %lu (%u x %u) could not be aligned on the "
                  "device; no CPU path in this build\n",
                  static_cast<unsigned long>(first + i), pairs[i].query_length,
                  pairs[i].target_length);
          exit(1);
        }
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
    align_chunks(tasks, window_length, &done, {}, nullptr);
    return true;
  }

  bool device_layers() const override { return true; }

  // The breaking points stay on the device: each chunk's go through the
  // layer filters into one set, which groups the layers by window at the end.
  void align_layers(const std::vector<AlignTask>& tasks, const std::vector<LayerSource>& sources,
                    uint32_t window_length, double quality_threshold, uint64_t n_windows,
                    const std::function<void(std::vector<char>*)>& gather_qualities,
                    std::vector<LayerRecord>* layers) override {
    static_assert(sizeof(LayerSource) == sizeof(rh_layer_source) &&
                      offsetof(LayerSource, first_window) == offsetof(rh_layer_source, first_window) &&
                      offsetof(LayerSource, quality) == offsetof(rh_layer_source, quality) &&
                      offsetof(LayerSource, mirrored) == offsetof(rh_layer_source, mirrored),
                  "LayerSource mirrors rh_layer_source");
    static_assert(sizeof(LayerRecord) == sizeof(rh_layer_record) &&
                      offsetof(LayerRecord, end) == offsetof(rh_layer_record, end),
                  "LayerRecord mirrors rh_layer_record");
    std::vector<rh_break_span> spans(tasks.size());
    for (uint64_t i = 0; i < tasks.size(); ++i) {
      spans[i] = rh_break_span{tasks[i].t_begin, tasks[i].t_end, tasks[i].q_first};
    }
    uint64_t room = 0;  // breaking-point pairs the chunks can give
    for (uint64_t i = 0; i < tasks.size(); i += 1u << 20) {
      const uint32_t n = static_cast<uint32_t>(std::min<uint64_t>(tasks.size() - i, 1u << 20));
      room += rh_breaking_points_bound(window_length, spans.data() + i, n) / 2;
    }
    // the qualities are gathered and the set is made (its device buffers,
    // the quality upload) beside the first chunk's alignment
    rh_layer_set* set = nullptr;
    int32_t set_rc = RH_OK;
    std::thread maker([&] {
      std::vector<char> qualities;
      gather_qualities(&qualities);
      set_rc = rh_layers_create(0, window_length, quality_threshold, n_windows, room,
                                qualities.data(), qualities.size(), &set);
    });
    auto layer_set = [&]() -> rh_layer_set* {
      if (maker.joinable()) maker.join();
      if (set_rc != RH_OK) die("rh_layers_create", set_rc);
      return set;
    };
    align_chunks(tasks, window_length, nullptr, layer_set,
                 reinterpret_cast<const rh_layer_source*>(sources.data()));
    layer_set();  // (no chunks: made here)
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t n_layers = 0;
    int32_t rc = rh_layers_finish(set, &n_layers);
    if (rc != RH_OK) die("rh_layers_finish", rc);
    layers->resize(n_layers);
    rc = rh_layers_download(set, reinterpret_cast<rh_layer_record*>(layers->data()), n_layers);
    if (rc != RH_OK) die("rh_layers_download", rc);
    rh_layers_destroy(set);
    fprintf(stderr,
            "[racon_amd::layers] %lu layers of %lu windows from %lu breaking-point pairs: "
            "grouped and copied in %.1f ms\n",
            static_cast<unsigned long>(n_layers), static_cast<unsigned long>(n_windows),
            static_cast<unsigned long>(room),
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count());
  }

 private:
  // Chunks of the tasks through two aligner batches. Each chunk's breaking
  // points go to done() (per task) or, with a layer set, through
  // rh_layers_add with sources[i] for task i.
  void align_chunks(const std::vector<AlignTask>& tasks, uint32_t window_length,
                    const std::function<void(uint64_t, const uint32_t*, uint64_t)>* done,
                    const std::function<rh_layer_set*()>& layer_set,
                    const rh_layer_source* sources) {
    // chunks of about 1 GiB of sequence on two batch objects: chunk c + 1 is
    // packed and copied while chunk c's kernel runs
    const uint64_t kChunkBytes = 1ull << 30;
    std::vector<std::pair<uint64_t, uint64_t>> chunks;
    for (uint64_t first = 0; first < tasks.size();) {
      uint64_t last = first, bytes = 0;
      while (last < tasks.size() && (last == first || bytes < kChunkBytes)) {
        bytes += tasks[last].query_length + tasks[last].target_length;
        ++last;
      }
      chunks.emplace_back(first, last);
      first = last;
    }
    struct Slot {
      rh_align_batch* batch = nullptr;
      std::vector<rh_pair> pairs;
      std::vector<rh_break_span> spans;
    };
    Slot slots[2];
    const uint32_t n_batches = chunks.size() > 1 ? 2 : 1;
    for (uint32_t k = 0; k < n_batches; ++k) {
      int32_t rc = rh_align_create(0, &slots[k].batch);
      if (rc != RH_OK) die("rh_align_create", rc);
      rc = rh_align_set_band(slots[k].batch, config_.aligner_band_width);
      if (rc != RH_OK) die("rh_align_set_band", rc);
    }
    auto upload = [&](uint64_t c) {
      Slot& sl = slots[c % n_batches];
      sl.pairs.clear();
      sl.spans.clear();
      for (uint64_t i = chunks[c].first; i < chunks[c].second; ++i) {
        const AlignTask& t = tasks[i];
        sl.pairs.push_back(rh_pair{t.query, t.query_length, t.target, t.target_length});
        sl.spans.push_back(rh_break_span{t.t_begin, t.t_end, t.q_first});
      }
      return rh_align_upload(sl.batch, static_cast<uint32_t>(sl.pairs.size()), sl.pairs.data());
    };
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> counts;
    std::vector<uint32_t> points;
    rh_align_stats total{};  // summed over the chunks
    double points_ms = 0;
    int32_t rc = chunks.empty() ? RH_OK : upload(0);
    if (rc != RH_OK) die("rh_align_upload", rc);
    // the kernel of a chunk runs on a helper thread (rh_align_run blocks):
    // chunk c + 1 is uploaded while chunk c's kernel runs; chunk c's
    // breaking points (a few ms of device time) come before chunk c + 1's
    // kernel, which would hold every compute unit until it ends, and chunk
    // c's callbacks and chunk c + 2's upload run beside that kernel
    int32_t run_rc = RH_OK;
    std::thread runner;
    auto start = [&](uint64_t c) {
      rh_align_batch* b = slots[c % n_batches].batch;
      runner = std::thread([&run_rc, b] { run_rc = rh_align_run(b); });
    };
    if (!chunks.empty()) start(0);
    for (uint64_t c = 0; c < chunks.size(); ++c) {
      Slot& sl = slots[c % n_batches];
      int32_t up_rc = RH_OK;
      if (c + 1 < chunks.size()) up_rc = upload(c + 1);
      runner.join();
      if (run_rc != RH_OK) die("rh_align_run", run_rc);
      if (up_rc != RH_OK) die("rh_align_upload", up_rc);
      rh_align_stats st{};
      if (rh_align_get_stats(sl.batch, &st) == RH_OK) {
        if (chunk_log_) {
          fprintf(stderr,
                  "[racon_amd::align] chunk %lu: %lu pairs, pack %.1f ms, H2D %.1f ms, kernels "
                  "%.1f ms\n",
                  static_cast<unsigned long>(c), static_cast<unsigned long>(st.pairs), st.pack_ms,
                  st.h2d_ms, st.kernel_ms);
        }
        total.pack_ms += st.pack_ms;
        total.h2d_ms += st.h2d_ms;
        total.kernel_ms += st.kernel_ms;
        total.pairs += st.pairs;
      }
      const auto t_points = std::chrono::steady_clock::now();
      const uint32_t n = static_cast<uint32_t>(sl.pairs.size());
      points.resize(2 * rh_breaking_points_bound(window_length, sl.spans.data(), n));
      offsets.resize(n);
      counts.resize(n);
      rc = rh_align_breaking_points(sl.batch, window_length, sl.spans.data(), offsets.data(),
                                    counts.data(), points.data(), points.size() / 2);
      if (rc != RH_OK) die("rh_align_breaking_points", rc);
      const uint64_t first = chunks[c].first;
      for (uint32_t i = 0; i < n; ++i) {
        if (counts[i] == RH_PAIR_REJECTED) {
          fprintf(stderr,
                  "[racon::HipEngine] error: overlap %lu (%u x %u) could not be aligned on the "
                  "device; no CPU path in this build\n",
                  static_cast<unsigned long>(first + i), sl.pairs[i].query_length,
                  sl.pairs[i].target_length);
          exit(1);
        }
      }
      if (layer_set) {
        rc = rh_layers_add(layer_set(), sl.batch, sources + first, n);
        if (rc != RH_OK) die("rh_layers_add", rc);
      }
      points_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                             t_points)
                       .count();
      if (c + 1 < chunks.size()) start(c + 1);
      if (done != nullptr) {
        for (uint32_t i = 0; i < n; ++i) {
          (*done)(first + i, points.data() + 2 * offsets[i], counts[i]);
        }
      }
    }
    for (uint32_t k = 0; k < n_batches; ++k) rh_align_destroy(slots[k].batch);
    fprintf(stderr,
            "[racon_amd::align] %lu overlaps in %zu chunks: pack %.1f ms, H2D %.1f ms, kernels "
            "%.1f ms, breaking points%s (kernels + copy) %.1f ms\n",
            static_cast<unsigned long>(total.pairs), chunks.size(), total.pack_ms, total.h2d_ms,
            total.kernel_ms, layer_set ? " and layer filters" : "", points_ms);
  }

 public:

  // Windows go to the device in chunks of about 1 GiB of sequence, on two
  // batch objects: while one chunk's kernel runs, the next chunk is packed
  // and copied on the other batch's stream. Results are set in window order
  // and reported through ready() chunk by chunk.
  void generate_consensus(std::vector<std::shared_ptr<Window>>& windows, bool trim,
                          std::vector<uint8_t>& status,
                          const std::function<void(uint64_t)>& ready) override {
    status.assign(windows.size(), kWindowUnpolished);
    if (config_.poa_server != nullptr) {
      serve_consensus(static_cast<rh_server*>(config_.poa_server), windows, status, ready);
      return;
    }
    const uint64_t kChunkBytes = 1ull << 30;
    std::vector<std::pair<uint64_t, uint64_t>> chunks;
    for (uint64_t first = 0; first < windows.size();) {
      uint64_t last = first, bytes = 0;
      while (last < windows.size() && (last == first || bytes < kChunkBytes)) {
        const Window& w = *windows[last];
        for (uint32_t k = 0; k < w.size(); ++k) bytes += 2ull * w.sequence(k).second;
        ++last;
      }
      chunks.emplace_back(first, last);
      first = last;
    }
    struct Slot {
      rh_poa_batch* batch = nullptr;
      std::vector<rh_window> windows;
      std::vector<rh_layer> layers;
    };
    Slot slots[2];
    const uint32_t n_batches = chunks.size() > 1 ? 2 : 1;
    const auto t_create = std::chrono::steady_clock::now();
    for (uint32_t k = 0; k < n_batches; ++k) {
      int32_t rc = rh_poa_create(0, config_.match, config_.mismatch, config_.gap, trim ? 1 : 0,
                                 &slots[k].batch);
      if (rc != RH_OK) die("rh_poa_create", rc);
      // racon -b: the banded approximation (GenomeWorks' band of 256)
      rc = rh_poa_set_band(slots[k].batch, config_.banded ? 256 : 0);
      if (rc != RH_OK) die("rh_poa_set_band", rc);
    }
    const double create_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_create)
            .count();
    rh_poa_stats total{};  // summed over the chunks
    uint64_t device_bytes = 0;
    auto upload = [&](uint64_t c) {
      Slot& sl = slots[c % n_batches];
      const uint64_t first = chunks[c].first, last = chunks[c].second;
      sl.windows.clear();
      sl.layers.clear();
      for (uint64_t i = first; i < last; ++i) {
        const Window& w = *windows[i];
        for (uint32_t k = 1; k < w.size(); ++k) {
          sl.layers.push_back(rh_layer{w.sequence(k).first, w.quality(k).first,
                                       w.sequence(k).second, w.position(k).first,
                                       w.position(k).second});
        }
      }
      uint64_t at = 0;
      for (uint64_t i = first; i < last; ++i) {
        const Window& w = *windows[i];
        sl.windows.push_back(rh_window{w.sequence(0).first, w.quality(0).first,
                                       w.sequence(0).second, w.size() - 1,
                                       sl.layers.data() + at,
                                       w.type() == WindowType::kTGS ? 1u : 0u});
        at += w.size() - 1;
      }
      return rh_poa_upload(sl.batch, static_cast<uint32_t>(sl.windows.size()),
                           sl.windows.data());
    };
    // results of two chunks: chunk c's are handed out (take_result,
    // ready) on a helper thread while the main thread uploads chunk c + 2
    struct Results {
      std::vector<char> out;
      std::vector<uint64_t> offsets;
      std::vector<uint32_t> lengths;
      std::vector<uint8_t> codes;
    };
    Results results[2];
    std::thread taker;
    int32_t rc = chunks.empty() ? RH_OK : upload(0);
    if (rc != RH_OK) die("rh_poa_upload", rc);
    if (!chunks.empty()) {
      rc = rh_poa_submit(slots[0].batch, RH_POA_BATCHED);
      if (rc != RH_OK) die("rh_poa_submit", rc);
    }
    for (uint64_t c = 0; c < chunks.size(); ++c) {
      rh_poa_batch* batch = slots[c % n_batches].batch;
      // chunk c's kernels run while the next chunk is packed and copied
      // here; chunk c + 1's kernels are queued before chunk c's results are
      // downloaded, and those are handed out on a helper thread while the
      // next upload runs here, so the device does not wait for the host
      if (c + 1 < chunks.size()) {
        rc = upload(c + 1);
        if (rc != RH_OK) die("rh_poa_upload", rc);
      }
      rc = rh_poa_wait(batch);
      if (rc != RH_OK) die("rh_poa_wait", rc);
      if (c + 1 < chunks.size()) {
        rc = rh_poa_submit(slots[(c + 1) % n_batches].batch, RH_POA_BATCHED);
        if (rc != RH_OK) die("rh_poa_submit", rc);
      }
      const uint64_t first = chunks[c].first, last = chunks[c].second;
      const uint64_t n = last - first;
      if (taker.joinable()) taker.join();  // chunk c - 1 handed out
      Results& r = results[c % 2];
      r.out.resize(rh_poa_output_bound(batch));
      r.offsets.resize(n);
      r.lengths.resize(n);
      r.codes.resize(n);
      rc = rh_poa_download(batch, r.out.data(), r.out.size(), r.offsets.data(),
                           r.lengths.data(), r.codes.data());
      if (rc != RH_OK) die("rh_poa_download", rc);
      device_bytes = std::max<uint64_t>(device_bytes, rh_poa_device_bytes(batch));
      rh_poa_stats st{};
      if (rh_poa_get_stats(batch, &st) == RH_OK) {
        if (chunk_log_) {
          fprintf(stderr,
                  "[racon_amd::consensus] chunk %lu: %lu windows, %.0f MB, pack %.1f ms, H2D "
                  "%.1f ms, kernels %.1f ms, D2H %.1f ms\n",
                  static_cast<unsigned long>(c), static_cast<unsigned long>(n),
                  st.input_bytes / 1e6, st.pack_ms, st.h2d_ms, st.kernel_ms, st.d2h_ms);
        }
        total.pack_ms += st.pack_ms;
        total.h2d_ms += st.h2d_ms;
        total.kernel_ms += st.kernel_ms;
        total.d2h_ms += st.d2h_ms;
        total.windows += st.windows;
        total.retried += st.retried;
      }
      const Results* rp = &r;
      taker = std::thread([&windows, &status, &ready, rp, first, last] {
        for (uint64_t i = first; i < last; ++i) {
          const uint64_t k = i - first;
          take_result(windows, i, rp->out.data() + rp->offsets[k], rp->lengths[k],
                      rp->codes[k], status);
        }
        if (ready) ready(last);
      });
    }
    if (taker.joinable()) taker.join();
    if (ready && chunks.empty()) ready(0);
    for (uint32_t k = 0; k < n_batches; ++k) rh_poa_destroy(slots[k].batch);
    // where the consensus phase went (the next chunk's upload overlaps the
    // kernels of the current one, so the parts can sum to more than the phase)
    fprintf(stderr,
            "[racon_amd::consensus] %lu windows in %zu chunks: batch setup %.1f ms, pack %.1f ms, "
            "H2D %.1f ms, kernels %.1f ms, D2H %.1f ms, %lu reruns, %.1f GB of device memory per "
            "batch\n",
            static_cast<unsigned long>(total.windows), chunks.size(), create_ms, total.pack_ms,
            total.h2d_ms, total.kernel_ms, total.d2h_ms, static_cast<unsigned long>(total.retried),
            device_bytes / 1e9);
  }

 private:
  // Sets window i's consensus and status from a device result, or stops on
  // a window the device refused (this binary has no CPU path).
  static void take_result(std::vector<std::shared_ptr<Window>>& windows, uint64_t i,
                          const char* data, uint32_t length, uint8_t code,
                          std::vector<uint8_t>& status) {
    if (code == RH_WINDOW_REJECTED) {
      // racon's CUDA path would polish it on the CPU; this binary has only
      // the device path, so it stops rather than print the backbone
      fprintf(stderr,
              "[racon::HipEngine] error: window %lu of target %lu (rank %u, %u layers) is "
              "outside the device limits; no CPU path in this build\n",
              static_cast<unsigned long>(i), static_cast<unsigned long>(windows[i]->id()),
              windows[i]->rank(), windows[i]->size() - 1);
      exit(1);
    }
    windows[i]->set_consensus(std::string(data, length));
    status[i] = code == RH_WINDOW_UNPOLISHED
                    ? kWindowUnpolished
                    : (code == RH_WINDOW_POLISHED_UNTRIMMED ? kWindowPolishedChimeric
                                                            : kWindowPolished);
  }

  // The windows as requests of a shared serving loop: chunks of at most a
  // third of its pool, all submitted at once (the server admits them as its
  // pool frees), then collected in order.
  void serve_consensus(rh_server* server, std::vector<std::shared_ptr<Window>>& windows,
                       std::vector<uint8_t>& status, const std::function<void(uint64_t)>& ready) {
    rh_server_stats st{};
    int32_t rc = rh_server_get_stats(server, &st);
    if (rc != RH_OK) die("rh_server_get_stats", rc);
    const uint64_t chunk_bytes = std::max<uint64_t>(1, st.pool_bytes / 3);
    struct Chunk {
      uint64_t first, last, request;
    };
    std::vector<Chunk> chunks;
    std::vector<rh_window> ws;
    std::vector<rh_layer> ls;
    for (uint64_t first = 0; first < windows.size();) {
      uint64_t last = first, bytes = 0;
      while (last < windows.size() && (last == first || bytes < chunk_bytes)) {
        const Window& w = *windows[last];
        for (uint32_t k = 0; k < w.size(); ++k) bytes += 3ull * w.sequence(k).second + 64;
        ++last;
      }
      ws.clear();
      ls.clear();
      for (uint64_t i = first; i < last; ++i) {
        const Window& w = *windows[i];
        for (uint32_t k = 1; k < w.size(); ++k) {
          ls.push_back(rh_layer{w.sequence(k).first, w.quality(k).first, w.sequence(k).second,
                                w.position(k).first, w.position(k).second});
        }
      }
      uint64_t at = 0;
      for (uint64_t i = first; i < last; ++i) {
        const Window& w = *windows[i];
        ws.push_back(rh_window{w.sequence(0).first, w.quality(0).first, w.sequence(0).second,
                               w.size() - 1, ls.data() + at,
                               w.type() == WindowType::kTGS ? 1u : 0u});
        at += w.size() - 1;
      }
      Chunk c{first, last, 0};
      rc = rh_server_submit(server, static_cast<uint32_t>(ws.size()), ws.data(), &c.request);
      if (rc != RH_OK) die("rh_server_submit", rc);
      chunks.push_back(c);
      first = last;
    }
    std::vector<char> out;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> lengths;
    std::vector<uint8_t> codes;
    for (const Chunk& c : chunks) {
      rc = rh_server_wait(server, c.request, -1);
      if (rc != RH_OK) die("rh_server_wait", rc);
      rh_request_info info{};
      rc = rh_server_request_info(server, c.request, &info);
      if (rc != RH_OK) die("rh_server_request_info", rc);
      const uint64_t n = c.last - c.first;
      out.resize(std::max<uint64_t>(1, info.output_bound));
      offsets.resize(n);
      lengths.resize(n);
      codes.resize(n);
      rc = rh_server_fetch(server, c.request, out.data(), out.size(), offsets.data(),
                           lengths.data(), codes.data());
      if (rc != RH_OK) die("rh_server_fetch", rc);
      for (uint64_t k = 0; k < n; ++k) {
        take_result(windows, c.first + k, out.data() + offsets[k], lengths[k], codes[k], status);
      }
      if (ready) ready(c.last);
    }
    if (ready && chunks.empty()) ready(0);
  }

  EngineConfig config_;
  // RACON_AMD_CHUNK_LOG set: one stderr line per chunk (upload parts and
  // kernel time), to see which step holds a pipeline up
  bool chunk_log_ = false;
};

}  // namespace

std::unique_ptr<PolishEngine> createEngine(const EngineConfig& config) {
  return std::unique_ptr<PolishEngine>(new HipEngine(config));
}

}  // namespace racon
