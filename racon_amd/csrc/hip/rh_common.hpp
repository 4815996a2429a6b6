// Host helpers shared by the C ABI translation units: per-thread error
// text, HIP status checks and event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "racon_hip.h"

namespace rh {

void set_error(const std::string& message);
void clear_error();
const char* last_error();

#define RH_HIP_CHECK(expr, code)                                                   \
  do {                                                                             \
    hipError_t rh_status_ = (expr);                                                \
    if (rh_status_ != hipSuccess) {                                                \
      ::rh::set_error(std::string(#expr) + ": " + hipGetErrorString(rh_status_)); \
      return (code);                                                               \
    }                                                                              \
  } while (0)

// The breaking points an aligner batch holds on the device after
// rh_align_breaking_points: spans[i] (room and offset of pair i), counts[i]
// and the points, on the batch's stream.
struct BreakSpan;
struct BreakView {
  const BreakSpan* spans;
  const uint32_t* counts;
  const uint32_t* points;
  uint32_t n;
  uint32_t window_length;
  uint64_t points_total;
  hipStream_t stream;
  int32_t device;
};
// false when the batch holds no breaking points (not computed since its last
// upload)
bool break_view(const rh_align_batch* b, BreakView* v);

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
      .count();
}

// Allocator whose resize leaves new elements uninitialised (host arenas
// whose bytes are all overwritten by the packing copies).
template <typename T>
struct DefaultInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = DefaultInit<U>;
  };
  using std::allocator<T>::allocator;
  template <typename U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <typename U, typename... Args>
  void construct(U* p, Args&&... args) {
    ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};

using HostArena = std::vector<uint8_t, DefaultInit<uint8_t>>;

// One host-to-arena copy of a packing pass.
struct ArenaCopy {
  const void* src;
  uint64_t dst;
  uint32_t len;
};

// Runs the copies into `base`, on up to 16 host threads when the batch is
// large (the copies do not overlap).
inline void run_copies(uint8_t* base, const std::vector<ArenaCopy>& copies, uint64_t bytes) {
  const size_t nc = copies.size();
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const unsigned workers = bytes >= (16u << 20) ? hw : 1u;
  auto run = [&](size_t begin, size_t end) {
    for (size_t c = begin; c < end; ++c) {
      if (copies[c].len != 0) std::memcpy(base + copies[c].dst, copies[c].src, copies[c].len);
    }
  };
  if (workers <= 1 || nc < 2) {
    run(0, nc);
    return;
  }
  std::vector<std::thread> pool;
  const size_t chunk = (nc + workers - 1) / workers;
  for (size_t c = 0; c < nc; c += chunk) {
    pool.emplace_back(run, c, std::min(nc, c + chunk));
  }
  for (auto& t : pool) t.join();
}

// Pinned host memory that grows on demand (never shrinks): staging for
// uploads that a batch object repeats, so that the copy engine reads it
// directly instead of through the runtime's bounce buffers.
struct PinnedBuffer {
  uint8_t* ptr = nullptr;
  uint64_t bytes = 0;
  hipError_t reserve(uint64_t n) {
    if (n <= bytes) return hipSuccess;
    release();
    const uint64_t want = n + n / 4;  // headroom for the next, slightly larger batch
    void* p = nullptr;
    const hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    ptr = static_cast<uint8_t*>(p);
    bytes = want;
    return hipSuccess;
  }
  void release() {
    if (ptr != nullptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

// Runs the copies (laid out back to back in arena order, covering
// [0, bytes)) into the pinned `base` on up to 16 host threads, and sends each
// finished chunk of about 8 MiB to `dev` on `stream` while later chunks are
// still being copied. The copies must all be finished (and the stream
// synchronised) before `base` is reused.
inline hipError_t copy_and_upload(uint8_t* base, const std::vector<ArenaCopy>& copies,
                                  uint64_t bytes, uint8_t* dev, hipStream_t stream) {
  const size_t nc = copies.size();
  if (nc == 0 || bytes == 0) return hipSuccess;
  constexpr uint64_t kChunk = 8ull << 20;
  constexpr size_t kClaim = 256;  // copies per claim of a worker
  // chunk k: copies [cb[k], cb[k + 1]), arena bytes from copies[cb[k]].dst
  std::vector<size_t> cb{0};
  uint64_t start = 0;
  for (size_t c = 0; c < nc; ++c) {
    const uint64_t end = copies[c].dst + copies[c].len;
    if (end - start >= kChunk) {
      cb.push_back(c + 1);
      start = end;
    }
  }
  if (cb.back() != nc) cb.push_back(nc);
  const size_t nk = cb.size() - 1;
  std::unique_ptr<std::atomic<size_t>[]> left(new std::atomic<size_t>[nk]);
  for (size_t k = 0; k < nk; ++k) left[k].store(cb[k + 1] - cb[k], std::memory_order_relaxed);
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      const size_t c0 = next.fetch_add(kClaim, std::memory_order_relaxed);
      if (c0 >= nc) return;
      const size_t c1 = std::min(nc, c0 + kClaim);
      for (size_t c = c0; c < c1; ++c) {
        if (copies[c].len != 0) std::memcpy(base + copies[c].dst, copies[c].src, copies[c].len);
      }
      // credit every chunk the claimed range touches
      size_t k = static_cast<size_t>(std::upper_bound(cb.begin(), cb.end(), c0) - cb.begin()) - 1;
      for (size_t c = c0; c < c1; ++k) {
        const size_t e = std::min(c1, cb[k + 1]);
        left[k].fetch_sub(e - c, std::memory_order_release);
        c = e;
      }
    }
  };
  const unsigned workers = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < workers; ++t) pool.emplace_back(work);
  hipError_t err = hipSuccess;
  for (size_t k = 0; k < nk; ++k) {
    while (left[k].load(std::memory_order_acquire) != 0) std::this_thread::yield();
    const uint64_t b0 = copies[cb[k]].dst;
    const uint64_t b1 = k + 1 < nk ? copies[cb[k + 1]].dst : bytes;
    if (err == hipSuccess && b1 > b0) {
      err = hipMemcpyAsync(dev + b0, base + b0, b1 - b0, hipMemcpyHostToDevice, stream);
    }
  }
  for (auto& t : pool) t.join();
  return err;
}

// A copy on `stream`, waited for. Unlike hipMemcpy (the null stream, which
// waits for every other stream of the device) it does not wait for kernels
// that other batch objects run at the same time.
inline hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                            hipStream_t stream) {
  if (bytes == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  return e;
}

// Device memory buffer that grows on demand (never shrinks).
struct DeviceBuffer {
  void* ptr = nullptr;
  uint64_t bytes = 0;
  // Room for n bytes. A buffer that must grow is freed and allocated again,
  // and hipFree waits for the whole device: a buffer refilled per chunk
  // while another stream's kernels run takes a quarter of headroom
  // (grow_headroom), so that chunks of about one size reuse it.
  hipError_t reserve(uint64_t n, bool grow_headroom = false) {
    if (n <= bytes) {
      return hipSuccess;
    }
    if (ptr != nullptr) {
      hipError_t e = hipFree(ptr);
      ptr = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    const uint64_t want = grow_headroom ? n + n / 4 : n;
    hipError_t e = hipMalloc(&ptr, want == 0 ? 1 : want);
    if (e == hipSuccess) {
      bytes = want;
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
