// racon --serve (serve.hpp): polishing jobs on one device through the
// serving loop of libracon_hip (rh_server_*).
#include "serve.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <thread>

#include "racon_hip.h"
#include "sequence.hpp"

namespace racon {

std::vector<ServeJob> read_jobs(const std::string& path) {
  std::ifstream in(path);
  if (!in) {
    fprintf(stderr, "[racon::serve] error: unable to open job list %s!\n", path.c_str());
    exit(1);
  }
  std::vector<ServeJob> jobs;
  std::string line;
  uint64_t number = 0;
  while (std::getline(in, line)) {
    ++number;
    std::istringstream fields(line);
    ServeJob job;
    if (!(fields >> job.output) || job.output[0] == '#') continue;
    if (!(fields >> job.sequences >> job.overlaps >> job.targets)) {
      fprintf(stderr,
              "[racon::serve] error: job list line %lu needs <output> <sequences> <overlaps> "
              "<targets> [arrival seconds]!\n",
              static_cast<unsigned long>(number));
      exit(1);
    }
    fields >> job.arrival;
    jobs.push_back(job);
  }
  return jobs;
}

int serve_jobs(const std::vector<ServeJob>& jobs, const ServeOptions& o) {
  using Clock = std::chrono::steady_clock;
  rh_server* server = nullptr;
  int32_t rc = rh_server_create(0, o.match, o.mismatch, o.gap, o.trim ? 1 : 0, o.budget_bytes,
                                o.executors, 2048, &server);
  if (rc != RH_OK) {
    fprintf(stderr, "[racon::serve] error: rh_server_create failed (%d): %s\n", rc,
            rh_last_error());
    return 1;
  }
  struct Record {
    double start = 0, end = 0;
    uint64_t windows = 0, sequences = 0;
  };
  std::vector<Record> records(jobs.size());
  const Clock::time_point t0 = Clock::now();
  auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  std::atomic<size_t> next{0};
  auto worker = [&] {
    for (size_t i = next++; i < jobs.size(); i = next++) {
      const ServeJob& job = jobs[i];
      const double wait = job.arrival - now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      Record& r = records[i];
      r.start = now();
      FILE* out = fopen(job.output.c_str(), "w");
      if (out == nullptr) {
        fprintf(stderr, "[racon::serve] error: unable to write %s!\n", job.output.c_str());
        exit(1);
      }
      {
        auto polisher = createPolisher(job.sequences, job.overlaps, job.targets, o.type,
                                       o.window_length, o.quality_threshold, o.error_threshold,
                                       o.trim, o.match, o.mismatch, o.gap, o.num_threads, 0,
                                       false, 0, o.aligner_band_width, server);
        polisher->initialize();
        // racon's output loop (main.cpp:164-166), into the job's file
        polisher->polish_streamed(o.drop_unpolished, [&](std::unique_ptr<Sequence> s) {
          fprintf(out, ">%s\n%s\n", s->name().c_str(), s->data().c_str());
          ++r.sequences;
        });
        r.windows = polisher->timings().num_windows;
      }
      fclose(out);
      r.end = now();
    }
  };
  std::vector<std::thread> pool;
  const size_t n_workers = std::max<size_t>(1, std::min<size_t>(o.concurrent_jobs, jobs.size()));
  for (size_t k = 0; k < n_workers; ++k) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  rh_server_stats st{};
  rh_server_get_stats(server, &st);
  rh_server_destroy(server);
  for (size_t i = 0; i < jobs.size(); ++i) {
    const Record& r = records[i];
    printf("{\"job\": %lu, \"output\": \"%s\", \"arrival_s\": %.6f, \"start_s\": %.6f, "
           "\"end_s\": %.6f, \"latency_s\": %.6f, \"windows\": %lu, \"sequences\": %lu}\n",
           static_cast<unsigned long>(i), jobs[i].output.c_str(), jobs[i].arrival, r.start, r.end,
           r.end - jobs[i].arrival, static_cast<unsigned long>(r.windows),
           static_cast<unsigned long>(r.sequences));
  }
  printf("{\"server\": {\"launches\": %lu, \"shared_launches\": %lu, \"windows_launched\": %lu, "
         "\"kernel_ms\": %.3f, \"pool_bytes\": %lu, \"scratch_bytes\": %lu}}\n",
         static_cast<unsigned long>(st.launches), static_cast<unsigned long>(st.shared_launches),
         static_cast<unsigned long>(st.windows_launched), st.kernel_ms,
         static_cast<unsigned long>(st.pool_bytes), static_cast<unsigned long>(st.scratch_bytes));
  fflush(stdout);
  return 0;
}

}  // namespace racon
