// Host helpers shared by the C ABI translation units: per-thread error
// text, HIP status checks and event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <string>

#include "racon_hip.h"

namespace rh {

void set_error(const std::string& message);
void clear_error();

#define RH_HIP_CHECK(expr, code)                                                   \
  do {                                                                             \
    hipError_t rh_status_ = (expr);                                                \
    if (rh_status_ != hipSuccess) {                                                \
      ::rh::set_error(std::string(#expr) + ": " + hipGetErrorString(rh_status_)); \
      return (code);                                                               \
    }                                                                              \
  } while (0)

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
      .count();
}

// Device memory buffer that grows on demand (never shrinks).
struct DeviceBuffer {
  void* ptr = nullptr;
  uint64_t bytes = 0;
  hipError_t reserve(uint64_t n) {
    if (n <= bytes) {
      return hipSuccess;
    }
    if (ptr != nullptr) {
      hipError_t e = hipFree(ptr);
      ptr = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipMalloc(&ptr, n == 0 ? 1 : n);
    if (e == hipSuccess) {
      bytes = n;
    }
    return e;
  }
  void release() {
    if (ptr != nullptr) {
      (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

}  // namespace rh
