// ORACLE (test infrastructure only; never linked into the product).
//
// The CPU hot path behind racon's host driver: one edlib restatement call
// per overlap and one spoa restatement call per window on a thread pool, the
// way racon fans out (polisher.cpp:462-484 and 486-504). Linked with the
// shared host sources into oracle/_build/racon_cpu, the baseline that the
// device build is compared against byte for byte.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <future>

#include "edlib_cpu.hpp"
#include "engine.hpp"
#include "serve.hpp"
#include "poa_cpu.hpp"
#include "thread_pool.hpp"
#include "window.hpp"

namespace racon {

namespace {

class CpuEngine : public PolishEngine {
 public:
  explicit CpuEngine(const EngineConfig& config)
      : config_(config), pool_(config.num_threads) {}

  std::string name() const override { return "cpu-oracle"; }

  void align_overlaps(
      const std::vector<AlignTask>& tasks,
      const std::function<void(uint64_t, const uint8_t*, uint64_t)>& done) override {
    std::vector<std::future<void>> futures;
    futures.reserve(tasks.size());
    for (uint64_t i = 0; i < tasks.size(); ++i) {
      futures.emplace_back(pool_.submit(
          [&](uint64_t j) -> void {
            const AlignTask& t = tasks[j];
            oracle::EdlibResult r = oracle::edlib_align_nw(
                t.query, static_cast<int32_t>(t.query_length), t.target,
                static_cast<int32_t>(t.target_length));
            if (r.status != 0 || r.edit_distance < 0) {
              fprintf(stderr, "[racon::Overlap::find_breaking_points] error: "
                              "edlib unable to align pair (%u x %u)!\n",
                      t.query_length, t.target_length);
              exit(1);
            }
            done(j, r.ops.data(), r.ops.size());
          },
          i));
    }
    for (const auto& f : futures) {
      f.wait();
    }
  }

  void generate_consensus(std::vector<std::shared_ptr<Window>>& windows, bool trim,
                          std::vector<uint8_t>& status,
                          const std::function<void(uint64_t)>& ready) override {
    status.assign(windows.size(), kWindowUnpolished);
    std::vector<std::future<void>> futures;
    futures.reserve(windows.size());
    for (uint64_t i = 0; i < windows.size(); ++i) {
      futures.emplace_back(pool_.submit(
          [&](uint64_t j) -> void {
            Window& w = *windows[j];
            oracle::WindowInput in;
            in.backbone = w.sequence(0).first;
            in.backbone_length = w.sequence(0).second;
            in.backbone_quality = w.quality(0).first;
            in.tgs = w.type() == WindowType::kTGS;
            if (!w.too_shallow()) {
              for (uint32_t k : w.layer_order()) {
                in.layers.push_back({w.sequence(k).first, w.quality(k).first,
                                     w.sequence(k).second, w.position(k).first,
                                     w.position(k).second});
              }
            } else {
              for (uint32_t k = 1; k < w.size(); ++k) {
                in.layers.push_back({w.sequence(k).first, w.quality(k).first,
                                     w.sequence(k).second, w.position(k).first,
                                     w.position(k).second});
              }
            }
            std::string consensus;
            status[j] = oracle::window_consensus(in, config_.match, config_.mismatch,
                                                 config_.gap, trim, consensus, nullptr);
            w.set_consensus(std::move(consensus));
          },
          i));
    }
    // windows in order, reported every 256
    for (uint64_t i = 0; i < futures.size(); ++i) {
      futures[i].wait();
      if (ready && ((i + 1) % 256 == 0 || i + 1 == futures.size())) ready(i + 1);
    }
    if (ready && futures.empty()) ready(0);
  }

 private:
  EngineConfig config_;
  ThreadPool pool_;
};

}  // namespace

std::unique_ptr<PolishEngine> createEngine(const EngineConfig& config) {
  return std::unique_ptr<PolishEngine>(new CpuEngine(config));
}

// --serve runs jobs on the device serving loop; the oracle build has none.
std::vector<ServeJob> read_jobs(const std::string&) {
  fprintf(stderr, "[racon::serve] error: --serve needs the device build!\n");
  exit(1);
}

int serve_jobs(const std::vector<ServeJob>&, const ServeOptions&) { return 1; }

}  // namespace racon
