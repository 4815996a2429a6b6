// C ABI: overlap alignment batches (racon_hip.h, rh_align_*).
//
// Packs query/target spans of many overlaps into one arena, runs the
// persistent alignment kernel over a work queue ordered by cost (largest
// first) and returns edlib-coded operations per pair.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "align_types.hpp"
#include "rh_common.hpp"

namespace rh {
hipError_t launch_align_batch(const AlignParams& params, uint32_t grid, hipStream_t stream);
hipError_t align_occupancy(int* blocks_per_cu);
hipError_t launch_breaking_points(const AlignPair* pairs, const AlignResult* results,
                                  const uint8_t* ops, const BreakSpan* spans, uint32_t n,
                                  uint32_t window_length, uint32_t* points, uint32_t* counts,
                                  hipStream_t stream);

// Windows of window_length target bases that the span [t_begin, t_end)
// crosses: one per multiple of W inside (t_begin, t_end), plus the last.
uint64_t span_windows(uint32_t window_length, const rh_break_span& s) {
  if (s.t_end <= s.t_begin) return 1;
  const uint64_t W = window_length;
  const uint64_t first = s.t_begin / W + 1;   // multiples k * W > t_begin
  const uint64_t last = (s.t_end - 1ull) / W;  // multiples k * W < t_end
  return (last >= first ? last - first + 1 : 0) + 1;
}
}  // namespace rh

struct rh_align_batch {
  int32_t device = 0;
  rh::HostArena arena;  // sized without zero-filling
  std::vector<rh::AlignPair> pairs;
  std::vector<uint32_t> order;
  std::vector<rh::AlignResult> results;
  uint64_t ops_bytes = 0;
  uint32_t max_q = 0, max_t = 0;
  rh::DeviceBuffer d_arena, d_pairs, d_ops, d_results, d_order, d_queue, d_scratch;
  rh::DeviceBuffer d_spans, d_points, d_counts;  // breaking points
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool uploaded = false, ran = false;
  // the last rh_align_breaking_points call's points, still on the device
  bool points_ready = false;
  uint32_t points_window = 0;
  uint64_t points_total = 0;
  uint32_t band0 = 0;  // first band guess of every pair (0: per pair, automatic)
  rh_align_stats stats{};
};

namespace rh {

bool break_view(const rh_align_batch* b, BreakView* v) {
  if (b == nullptr || v == nullptr || !b->points_ready) return false;
  v->spans = b->d_spans.as<BreakSpan>();
  v->counts = b->d_counts.as<uint32_t>();
  v->points = b->d_points.as<uint32_t>();
  v->n = static_cast<uint32_t>(b->pairs.size());
  v->window_length = b->points_window;
  v->points_total = b->points_total;
  v->stream = b->stream;
  v->device = b->device;
  return true;
}

}  // namespace rh

extern "C" {

int32_t rh_align_create(int32_t device, rh_align_batch** batch) {
  rh::clear_error();
  if (batch == nullptr) {
    rh::set_error("rh_align_create: null batch pointer");
    return RH_ERR_ARGUMENT;
  }
  *batch = nullptr;
  int n = 0;
  RH_HIP_CHECK(hipGetDeviceCount(&n), RH_ERR_DEVICE);
  if (device < 0 || device >= n) {
    rh::set_error("rh_align_create: no such device");
    return RH_ERR_DEVICE;
  }
  RH_HIP_CHECK(hipSetDevice(device), RH_ERR_DEVICE);
  rh_align_batch* b = new rh_align_batch();
  b->device = device;
  if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess) {
    rh::set_error("rh_align_create: stream/event creation failed");
    delete b;
    return RH_ERR_DEVICE;
  }
  *batch = b;
  return RH_OK;
}

void rh_align_destroy(rh_align_batch* b) {
  if (b == nullptr) {
    return;
  }
  (void)hipSetDevice(b->device);
  if (b->stream != nullptr) (void)hipStreamSynchronize(b->stream);
  b->d_arena.release();
  b->d_pairs.release();
  b->d_ops.release();
  b->d_results.release();
  b->d_spans.release();
  b->d_points.release();
  b->d_counts.release();
  b->d_order.release();
  b->d_queue.release();
  b->d_scratch.release();
  if (b->ev0 != nullptr) (void)hipEventDestroy(b->ev0);
  if (b->ev1 != nullptr) (void)hipEventDestroy(b->ev1);
  if (b->stream != nullptr) (void)hipStreamDestroy(b->stream);
  delete b;
}

int32_t rh_align_upload(rh_align_batch* b, uint32_t n, const rh_pair* pairs) {
  rh::clear_error();
  if (b == nullptr || (n > 0 && pairs == nullptr)) {
    rh::set_error("rh_align_upload: null argument");
    return RH_ERR_ARGUMENT;
  }
  RH_HIP_CHECK(hipSetDevice(b->device), RH_ERR_DEVICE);
  auto t0 = std::chrono::steady_clock::now();
  b->stats = rh_align_stats{};
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if ((pairs[i].query_length > 0 && pairs[i].query == nullptr) ||
        (pairs[i].target_length > 0 && pairs[i].target == nullptr)) {
      rh::set_error("rh_align_upload: pair " + std::to_string(i) + " is malformed");
      return RH_ERR_ARGUMENT;
    }
    total += pairs[i].query_length + pairs[i].target_length;
  }
  // offsets in pair order first, then the byte copies (on host threads)
  b->arena.clear();
  b->arena.resize(total);
  std::vector<rh::ArenaCopy> copies;
  copies.reserve(2ull * n);
  b->pairs.assign(n, rh::AlignPair{});
  b->max_q = b->max_t = 0;
  uint64_t out = 0;
  uint64_t cur = 0;
  for (uint32_t i = 0; i < n; ++i) {
    rh::AlignPair& p = b->pairs[i];
    p.q_off = cur;
    copies.push_back(rh::ArenaCopy{pairs[i].query, cur, pairs[i].query_length});
    cur += pairs[i].query_length;
    p.t_off = cur;
    copies.push_back(rh::ArenaCopy{pairs[i].target, cur, pairs[i].target_length});
    cur += pairs[i].target_length;
    p.q_len = pairs[i].query_length;
    p.t_len = pairs[i].target_length;
    p.out_off = out;
    out += rh::align_up_a(static_cast<uint64_t>(p.q_len) + p.t_len, 16);
    b->max_q = std::max(b->max_q, p.q_len);
    b->max_t = std::max(b->max_t, p.t_len);
    b->stats.cells += static_cast<uint64_t>(p.q_len) * p.t_len;
  }
  b->ops_bytes = out;
  rh::run_copies(b->arena.data(), copies, total);
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0u);
  std::stable_sort(b->order.begin(), b->order.end(), [&](uint32_t x, uint32_t y) {
    return static_cast<uint64_t>(b->pairs[x].q_len) * b->pairs[x].t_len >
           static_cast<uint64_t>(b->pairs[y].q_len) * b->pairs[y].t_len;
  });
  b->results.assign(n, rh::AlignResult{0, -1, rh::kAlignNotRun, 0});
  b->stats.pairs = n;
  b->stats.pack_ms = rh::ms_since(t0);

  auto t1 = std::chrono::steady_clock::now();
  RH_HIP_CHECK(b->d_arena.reserve(std::max<uint64_t>(b->arena.size(), 1), true), RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_pairs.reserve(std::max<uint64_t>(n, 1) * sizeof(rh::AlignPair), true),
               RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_results.reserve(std::max<uint64_t>(n, 1) * sizeof(rh::AlignResult), true),
               RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_ops.reserve(std::max<uint64_t>(out, 1), true), RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_order.reserve(std::max<uint64_t>(n, 1) * sizeof(uint32_t), true),
               RH_ERR_MEMORY);
  if (!b->arena.empty()) {
    RH_HIP_CHECK(hipMemcpyAsync(b->d_arena.ptr, b->arena.data(), b->arena.size(),
                                hipMemcpyHostToDevice, b->stream),
                 RH_ERR_DEVICE);
  }
  if (n > 0) {
    RH_HIP_CHECK(hipMemcpyAsync(b->d_pairs.ptr, b->pairs.data(), n * sizeof(rh::AlignPair),
                                hipMemcpyHostToDevice, b->stream),
                 RH_ERR_DEVICE);
    RH_HIP_CHECK(hipMemcpyAsync(b->d_order.ptr, b->order.data(), n * sizeof(uint32_t),
                                hipMemcpyHostToDevice, b->stream),
                 RH_ERR_DEVICE);
  }
  RH_HIP_CHECK(hipStreamSynchronize(b->stream), RH_ERR_DEVICE);
  b->stats.h2d_ms = rh::ms_since(t1);
  b->uploaded = true;
  b->ran = false;
  b->points_ready = false;
  return RH_OK;
}

int32_t rh_align_run(rh_align_batch* b) {
  rh::clear_error();
  if (b == nullptr || !b->uploaded) {
    rh::set_error("rh_align_run: nothing uploaded");
    return RH_ERR_STATE;
  }
  RH_HIP_CHECK(hipSetDevice(b->device), RH_ERR_DEVICE);
  const uint32_t n = static_cast<uint32_t>(b->pairs.size());
  b->stats.kernel_ms = 0;
  b->stats.launches = 0;
  if (n == 0) {
    b->ran = true;
    b->points_ready = false;
    return RH_OK;
  }
  const rh::AlignSlotCaps caps = rh::align_caps_for(b->max_q, b->max_t);
  const rh::AlignSlotLayout layout = rh::align_slot_layout(caps);
  const uint64_t slot_bytes = rh::align_up_a(layout.total, 4096);

  int per_cu = 0, cus = 0;
  RH_HIP_CHECK(rh::align_occupancy(&per_cu), RH_ERR_DEVICE);
  RH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device),
               RH_ERR_DEVICE);
  size_t free_b = 0, total_b = 0;
  RH_HIP_CHECK(hipMemGetInfo(&free_b, &total_b), RH_ERR_DEVICE);
  uint64_t slots = static_cast<uint64_t>(std::max(1, per_cu)) * std::max(1, cus);
  slots = std::min<uint64_t>(slots, std::max<uint64_t>(
                                        1, (free_b + b->d_scratch.bytes) * 6 / 10 / slot_bytes));
  slots = std::min<uint64_t>(slots, n);
  RH_HIP_CHECK(b->d_scratch.reserve(slots * slot_bytes), RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_queue.reserve(sizeof(uint32_t)), RH_ERR_MEMORY);
  RH_HIP_CHECK(hipMemsetAsync(b->d_queue.ptr, 0, sizeof(uint32_t), b->stream), RH_ERR_DEVICE);

  rh::AlignParams p{};
  p.pairs = b->d_pairs.as<rh::AlignPair>();
  p.arena = b->d_arena.as<uint8_t>();
  p.ops = b->d_ops.as<uint8_t>();
  p.results = b->d_results.as<rh::AlignResult>();
  p.order = b->d_order.as<uint32_t>();
  p.n_order = n;
  p.queue_head = b->d_queue.as<uint32_t>();
  p.scratch = b->d_scratch.as<uint8_t>();
  p.slot_bytes = slot_bytes;
  p.caps = caps;
  p.layout = layout;
  p.band0 = b->band0;

  RH_HIP_CHECK(hipEventRecord(b->ev0, b->stream), RH_ERR_DEVICE);
  RH_HIP_CHECK(rh::launch_align_batch(p, static_cast<uint32_t>(slots), b->stream),
               RH_ERR_KERNEL);
  RH_HIP_CHECK(hipEventRecord(b->ev1, b->stream), RH_ERR_DEVICE);
  RH_HIP_CHECK(hipEventSynchronize(b->ev1), RH_ERR_KERNEL);
  float ms = 0;
  RH_HIP_CHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1), RH_ERR_DEVICE);
  b->stats.kernel_ms = ms;
  b->stats.launches = 1;
  RH_HIP_CHECK(rh::copy_sync(b->results.data(), b->d_results.ptr, n * sizeof(rh::AlignResult),
                             hipMemcpyDeviceToHost, b->stream),
               RH_ERR_DEVICE);
  b->stats.ops = 0;
  b->stats.steps = 0;
  b->stats.rejected = 0;
  for (uint32_t i = 0; i < n; ++i) {
    b->stats.steps += b->results[i].pad;
    if (b->results[i].status != rh::kAlignOk) {
      // left to the caller's CPU path (edit distance -1 in rh_align_download)
      b->stats.rejected++;
      continue;
    }
    b->stats.ops += b->results[i].n_ops;
  }
  b->ran = true;
  b->points_ready = false;
  return RH_OK;
}

int32_t rh_align_set_band(rh_align_batch* b, uint32_t band) {
  if (b == nullptr) {
    rh::set_error("rh_align_set_band: null batch");
    return RH_ERR_ARGUMENT;
  }
  b->band0 = band;
  return RH_OK;
}

uint64_t rh_align_output_bound(const rh_align_batch* b) {
  if (b == nullptr) {
    return 0;
  }
  uint64_t total = 0;
  for (const auto& p : b->pairs) {
    total += static_cast<uint64_t>(p.q_len) + p.t_len;
  }
  return total;
}

int32_t rh_align_download(rh_align_batch* b, uint8_t* ops, uint64_t capacity,
                          uint64_t* offsets, uint32_t* n_ops, int32_t* edit_distance) {
  rh::clear_error();
  if (b == nullptr || !b->ran) {
    rh::set_error("rh_align_download: batch has not run");
    return RH_ERR_STATE;
  }
  if (ops == nullptr || offsets == nullptr || n_ops == nullptr || edit_distance == nullptr) {
    rh::set_error("rh_align_download: null argument");
    return RH_ERR_ARGUMENT;
  }
  RH_HIP_CHECK(hipSetDevice(b->device), RH_ERR_DEVICE);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<uint8_t> dev(b->ops_bytes);
  if (b->ops_bytes > 0) {
    RH_HIP_CHECK(rh::copy_sync(dev.data(), b->d_ops.ptr, b->ops_bytes, hipMemcpyDeviceToHost,
                               b->stream),
                 RH_ERR_DEVICE);
  }
  uint64_t pos = 0;
  for (size_t i = 0; i < b->pairs.size(); ++i) {
    const bool ok = b->results[i].status == rh::kAlignOk;
    const uint32_t len = ok ? b->results[i].n_ops : 0;
    if (pos + len > capacity) {
      rh::set_error("rh_align_download: output capacity too small");
      return RH_ERR_ARGUMENT;
    }
    memcpy(ops + pos, dev.data() + b->pairs[i].out_off, len);
    offsets[i] = pos;
    n_ops[i] = len;
    edit_distance[i] = ok ? b->results[i].edit_distance : -1;
    pos += len;
  }
  b->stats.d2h_ms = rh::ms_since(t0);
  return RH_OK;
}

uint64_t rh_breaking_points_bound(uint32_t window_length, const rh_break_span* spans, uint32_t n) {
  if (spans == nullptr || window_length == 0) return 0;
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) total += 2 * rh::span_windows(window_length, spans[i]);
  return total;
}

int32_t rh_align_breaking_points(rh_align_batch* b, uint32_t window_length,
                                 const rh_break_span* spans, uint64_t* offsets, uint32_t* counts,
                                 uint32_t* points, uint64_t capacity) {
  rh::clear_error();
  if (b == nullptr || !b->ran) {
    rh::set_error("rh_align_breaking_points: batch has not run");
    return RH_ERR_STATE;
  }
  if (window_length == 0 || spans == nullptr || offsets == nullptr || counts == nullptr ||
      points == nullptr) {
    rh::set_error("rh_align_breaking_points: null argument or zero window length");
    return RH_ERR_ARGUMENT;
  }
  const uint32_t n = static_cast<uint32_t>(b->pairs.size());
  if (n == 0) return RH_OK;
  std::vector<rh::BreakSpan> bs(n);
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t cap = 2 * rh::span_windows(window_length, spans[i]);
    bs[i] = rh::BreakSpan{spans[i].t_begin, spans[i].t_end, spans[i].q_first,
                          static_cast<uint32_t>(cap), total};
    offsets[i] = total;
    total += cap;
  }
  if (total > capacity) {
    rh::set_error("rh_align_breaking_points: output capacity too small");
    return RH_ERR_ARGUMENT;
  }
  RH_HIP_CHECK(hipSetDevice(b->device), RH_ERR_DEVICE);
  rh::DeviceBuffer& d_spans = b->d_spans;
  rh::DeviceBuffer& d_points = b->d_points;
  rh::DeviceBuffer& d_counts = b->d_counts;
  RH_HIP_CHECK(d_spans.reserve(n * sizeof(rh::BreakSpan)), RH_ERR_MEMORY);
  RH_HIP_CHECK(d_points.reserve(std::max<uint64_t>(total, 1) * 2 * sizeof(uint32_t)),
               RH_ERR_MEMORY);
  RH_HIP_CHECK(d_counts.reserve(n * sizeof(uint32_t)), RH_ERR_MEMORY);
  RH_HIP_CHECK(hipMemcpyAsync(d_spans.ptr, bs.data(), n * sizeof(rh::BreakSpan),
                              hipMemcpyHostToDevice, b->stream),
               RH_ERR_DEVICE);
  RH_HIP_CHECK(rh::launch_breaking_points(b->d_pairs.as<rh::AlignPair>(),
                                          b->d_results.as<rh::AlignResult>(),
                                          b->d_ops.as<uint8_t>(), d_spans.as<rh::BreakSpan>(), n,
                                          window_length, d_points.as<uint32_t>(),
                                          d_counts.as<uint32_t>(), b->stream),
               RH_ERR_KERNEL);
  RH_HIP_CHECK(hipMemcpyAsync(counts, d_counts.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              b->stream),
               RH_ERR_DEVICE);
  if (total > 0) {
    RH_HIP_CHECK(hipMemcpyAsync(points, d_points.ptr, total * 2 * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, b->stream),
                 RH_ERR_DEVICE);
  }
  RH_HIP_CHECK(hipStreamSynchronize(b->stream), RH_ERR_DEVICE);
  b->points_ready = true;
  b->points_window = window_length;
  b->points_total = total;
  for (uint32_t i = 0; i < n; ++i) {
    if (counts[i] == 0xffffffffu) {
      rh::set_error("rh_align_breaking_points: pair " + std::to_string(i) +
                    " has more points than its span allows");
      return RH_ERR_KERNEL;
    }
  }
  return RH_OK;
}

int32_t rh_align_get_stats(const rh_align_batch* b, rh_align_stats* stats) {
  if (b == nullptr || stats == nullptr) {
    rh::set_error("rh_align_get_stats: null argument");
    return RH_ERR_ARGUMENT;
  }
  *stats = b->stats;
  return RH_OK;
}

}  // extern "C"
