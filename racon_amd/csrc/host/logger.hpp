// Phase timer and progress bar on stderr (racon src/logger.hpp:17-54,
// src/logger.cpp:20-54). The phase lines are also the CPU/GPU timing source.
#pragma once

#include <chrono>
#include <string>

namespace racon {

class Logger {
 public:
  Logger() = default;

  // Starts a new phase (adds the previous phase to the total).
  void log();
  // Prints msg with the time since the phase started.
  void log(const std::string& msg) const;
  // Advances a 20-step progress bar.
  void bar(const std::string& msg);
  // Prints msg with the total time.
  void total(const std::string& msg) const;

  // Seconds since the phase started.
  double elapsed() const;

 private:
  double time_ = 0.;
  uint32_t bar_ = 0;
  std::chrono::time_point<std::chrono::steady_clock> time_point_;
};

}  // namespace racon
