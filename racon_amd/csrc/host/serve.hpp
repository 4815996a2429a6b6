// racon --serve: many polishing jobs (reads + overlaps + targets each) in
// one process on one device. No counterpart in racon, which polishes one job
// per process. Every job runs racon's own host path (Polisher::initialize,
// the overlap alignment, windows, stitching) on threads of its own, and the
// windows of all jobs go to one device serving loop (rh_server_* of
// racon_hip.h) as requests, so that jobs of different sizes share the GPU.
// A job's output file holds the bytes `racon` prints for it alone.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "polisher.hpp"

namespace racon {

struct ServeJob {
  std::string output;     // FASTA written here
  std::string sequences;  // racon's three positional inputs
  std::string overlaps;
  std::string targets;
  double arrival = 0;     // seconds after the start: the job begins no earlier
};

struct ServeOptions {
  PolisherType type = PolisherType::kC;
  uint32_t window_length = 500;
  double quality_threshold = 10.0;
  double error_threshold = 0.3;
  bool trim = true;
  int8_t match = 3, mismatch = -5, gap = -4;
  uint32_t num_threads = 1;  // per job
  bool drop_unpolished = true;
  uint32_t aligner_band_width = 0;
  uint64_t budget_bytes = 8ull << 30;  // device pool of the serving loop
  uint32_t executors = 4;
  uint32_t concurrent_jobs = 4;
};

// Job lines: "<output> <sequences> <overlaps> <targets> [arrival seconds]";
// empty lines and lines starting with '#' are skipped. Exits on a bad line.
std::vector<ServeJob> read_jobs(const std::string& path);

// Runs the jobs and prints one JSON line per job to stdout (arrival, start
// and end on the serve clock, windows, sequences written). Returns the exit
// status.
int serve_jobs(const std::vector<ServeJob>& jobs, const ServeOptions& options);

}  // namespace racon
