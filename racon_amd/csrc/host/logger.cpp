#include "logger.hpp"

#include <iostream>

namespace racon {

using Seconds = std::chrono::duration<double>;

void Logger::log() {
  auto now = std::chrono::steady_clock::now();
  if (time_point_ != std::chrono::time_point<std::chrono::steady_clock>()) {
    time_ += std::chrono::duration_cast<Seconds>(now - time_point_).count();
  }
  time_point_ = now;
}

double Logger::elapsed() const {
  return std::chrono::duration_cast<Seconds>(std::chrono::steady_clock::now() -
                                             time_point_).count();
}

void Logger::log(const std::string& msg) const {
  std::cerr << msg << " " << std::fixed << elapsed() << " s" << std::endl;
}

void Logger::bar(const std::string& msg) {
  ++bar_;
  std::string progress = "[" + std::string(bar_, '=') +
                         (bar_ == 20 ? "" : ">" + std::string(19 - bar_, ' ')) + "]";
  std::cerr << msg << " " << progress << " " << std::fixed << elapsed() << " s";
  bar_ %= 20;
  if (bar_ == 0) {
    std::cerr << std::endl;
  } else {
    std::cerr << "\r";
  }
}

void Logger::total(const std::string& msg) const {
  std::cerr << msg << " " << std::fixed << time_ + elapsed() << " s" << std::endl;
}

}  // namespace racon
