// C ABI: window consensus batches (racon_hip.h, rh_poa_*).
//
// Host side of the POA path: orders each window's layers the way
// Window::generate_consensus does (std::sort by begin, src/window.cpp:78-85),
// packs windows into one device arena, sizes per-wave scratch slots, runs the
// persistent kernel over a work queue ordered by estimated cost (largest
// first) or one launch per window, and reruns the few windows whose graph
// outgrew a default slot in slots sized from their exact bounds.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "poa_host.hpp"

namespace rh {

hipError_t launch_poa_batch(const PoaParams* d_params, uint32_t lds_bytes, uint32_t grid,
                            hipStream_t stream);
hipError_t launch_poa_single(const PoaParams* d_params, uint32_t lds_bytes, uint32_t q,
                             hipStream_t stream);
hipError_t launch_poa_static(const PoaParams* d_params, uint32_t lds_bytes, uint32_t first,
                             uint32_t count, hipStream_t stream);
hipError_t poa_occupancy(int* blocks_per_cu, uint32_t lds_bytes);
// the same kernels with wide aligned sets (poa_kernel_wide.hip)
hipError_t launch_poa_batch_wide(const PoaParams* d_params, uint32_t lds_bytes, uint32_t grid,
                                 hipStream_t stream);
hipError_t launch_poa_single_wide(const PoaParams* d_params, uint32_t lds_bytes, uint32_t q,
                                  hipStream_t stream);
hipError_t launch_poa_static_wide(const PoaParams* d_params, uint32_t lds_bytes, uint32_t first,
                                  uint32_t count, hipStream_t stream);
hipError_t poa_occupancy_wide(int* blocks_per_cu, uint32_t lds_bytes);

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& message) { g_error = message; }
void clear_error() { g_error.clear(); }
const char* last_error() { return g_error.c_str(); }

namespace {

// DP rows kept in LDS (16-bit scores): as many as fit while 16 waves share a
// CU's 160 KiB (12 or 8 waves for long layers), at most 8; a tuned value
// overrides it.
uint32_t ring_rows_for(uint32_t cols, uint32_t tuned) {
  if (tuned != 0) {
    return std::max<uint32_t>(2, std::min<uint32_t>(16, tuned));
  }
  const uint64_t lds_per_cu = 160u * 1024u;
  for (uint32_t waves : {16u, 12u, 8u}) {
    const uint64_t per_wave = lds_per_cu / waves;
    const uint64_t fixed = align_up(cols, 16);
    if (per_wave <= fixed) continue;
    const uint64_t rows = (per_wave - fixed) / (2ull * cols);
    if (rows >= 4 || waves == 8u) {
      return static_cast<uint32_t>(std::max<uint64_t>(2, std::min<uint64_t>(8, rows)));
    }
  }
  return 2;
}

// Windows at the front of the cost-ordered queue whose estimated cost is at
// least 1.5 times the mean: the ones that can outlast the rest of the batch
// on a single wave (2x and 1.25x measured slower on a 10k-window batch).
uint32_t priority_windows(const PoaWindow* host, const std::vector<uint32_t>& order) {
  if (order.empty()) {
    return 0;
  }
  long double total = 0;
  for (uint32_t w : order) total += static_cast<long double>(window_cost(host[w]));
  const long double mean = total / order.size();
  uint32_t n = 0;
  while (n < order.size() && window_cost(host[order[n]]) >= 1.5L * mean) ++n;
  return n;
}

}  // namespace

uint64_t graph_estimate(const PoaWindow& w) {
  const uint64_t layer_bytes = w.node_bound - w.backbone_len;
  return std::min<uint64_t>(
      w.node_bound, w.backbone_len + std::max<uint64_t>(2ull * w.max_layer_len, layer_bytes / 4));
}

uint64_t window_cost(const PoaWindow& w) {
  return static_cast<uint64_t>(w.node_bound) * (w.max_layer_len + 1);
}

bool is_retry_status(int32_t s) {
  return s == kPoaNodeOverflow || s == kPoaRowOverflow || s == kPoaOutOverflow ||
         s == kPoaBandMiss || s == kPoaAlignedFull;
}

namespace {

// Runs body(begin, end) over [0, n) on up to 16 host threads (serially for a
// small n).
template <typename F>
void parallel_ranges(uint32_t n, uint32_t serial_below, F&& body) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (hw <= 1 || n < serial_below) {
    body(0u, n);
    return;
  }
  const uint32_t chunk = (n + 4 * hw - 1) / (4 * hw);
  std::atomic<uint32_t> next{0};
  auto work = [&]() {
    for (;;) {
      const uint32_t b = next.fetch_add(chunk);
      if (b >= n) return;
      body(b, std::min(n, b + chunk));
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < hw; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
}

// Per-window sizes of a packing pass (first pass, in parallel).
struct WindowPlan {
  uint64_t bytes = 0;      // arena bytes
  uint32_t copies = 0;     // arena copies
  uint32_t layers = 0;     // layer records
  uint32_t bad_layer = 0;  // 1 + index of the first invalid layer, else 0
  uint64_t node_bound = 0;
  uint32_t max_len = 0;
  bool malformed = false;
  bool shallow = false;
};

}  // namespace

// Two passes over the windows, each on up to 16 host threads: sizes and
// checks, then (after prefix sums) the window and layer records and the
// arena copies at their final offsets. The layout is the one of appending
// every window in order: backbone, its quality, then each layer's bytes and
// quality in racon's order.
int32_t pack_windows(uint32_t n, const rh_window* windows, bool trim, PackedWindows* out,
                     PinnedBuffer* staging) {
  PackedWindows& P = *out;
  P.arena.clear();
  P.staged_copies.clear();
  P.host_arena = nullptr;
  P.arena_bytes = 0;
  P.windows.assign(n, PoaWindow{});
  P.layers.clear();
  P.gpu_order.clear();
  P.shallow.assign(n, 0);
  P.host_rejected.assign(n, 0);
  P.out_bytes = 0;

  std::vector<WindowPlan> plan(n);
  parallel_ranges(n, 512, [&](uint32_t b, uint32_t e) {
    for (uint32_t i = b; i < e; ++i) {
      const rh_window& w = windows[i];
      WindowPlan& q = plan[i];
      if (w.backbone == nullptr || w.backbone_quality == nullptr || w.backbone_length == 0 ||
          (w.n_layers > 0 && w.layers == nullptr)) {
        q.malformed = true;
        continue;
      }
      for (uint32_t k = 0; k < w.n_layers; ++k) {
        const rh_layer& l = w.layers[k];
        if (l.sequence == nullptr || l.length == 0 || l.begin >= l.end ||
            l.end > w.backbone_length) {
          q.bad_layer = k + 1;
          break;
        }
      }
      q.bytes = 2ull * w.backbone_length;
      q.copies = 2;
      q.shallow = w.n_layers + 1 < 3;
      if (q.shallow) continue;
      q.layers = w.n_layers;
      uint64_t sum_len = 0;
      for (uint32_t k = 0; k < w.n_layers; ++k) {
        const rh_layer& l = w.layers[k];
        const uint32_t parts = l.quality != nullptr ? 2u : 1u;
        q.bytes += static_cast<uint64_t>(parts) * l.length;
        q.copies += parts;
        sum_len += l.length;
        q.max_len = std::max(q.max_len, l.length);
      }
      q.node_bound = w.backbone_length + sum_len;
    }
  });
  for (uint32_t i = 0; i < n; ++i) {
    if (plan[i].malformed) {
      set_error("window " + std::to_string(i) + " is malformed");
      return RH_ERR_ARGUMENT;
    }
  }
  for (uint32_t i = 0; i < n; ++i) {
    if (plan[i].bad_layer != 0) {
      set_error("window " + std::to_string(i) + " layer " +
                std::to_string(plan[i].bad_layer - 1) + " is invalid");
      return RH_ERR_ARGUMENT;
    }
  }

  // prefix sums: arena offset, first copy, first layer record, output offset
  std::vector<uint64_t> at(n + 1), out_at(n + 1);
  std::vector<uint32_t> copy_at(n + 1), layer_at(n + 1);
  at[0] = out_at[0] = 0;
  copy_at[0] = layer_at[0] = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const WindowPlan& q = plan[i];
    // DP rows are packed in 24 bits of a predecessor entry: a larger graph is
    // refused here (RH_WINDOW_REJECTED, the caller's CPU path takes it); the
    // in-edge index of a move code (13 bits) is checked on the device
    const bool gpu = !q.shallow && q.node_bound < (1ull << 24) - 1;
    const uint64_t out_cap =
        gpu ? std::min<uint64_t>(q.node_bound, 2ull * windows[i].backbone_length + 1024) : 0;
    at[i + 1] = at[i] + q.bytes;
    copy_at[i + 1] = copy_at[i] + q.copies;
    layer_at[i + 1] = layer_at[i] + q.layers;
    out_at[i + 1] = out_at[i] + (gpu ? align_up(out_cap, 16) : 0);
    if (q.shallow) P.shallow[i] = 1;
    if (!q.shallow && !gpu) P.host_rejected[i] = 1;
  }
  const uint64_t cur = at[n];
  std::vector<ArenaCopy> copies(copy_at[n]);
  P.layers.resize(layer_at[n]);

  parallel_ranges(n, 512, [&](uint32_t b, uint32_t e) {
    std::vector<uint32_t> rank;
    for (uint32_t i = b; i < e; ++i) {
      const rh_window& w = windows[i];
      const WindowPlan& q = plan[i];
      PoaWindow& d = P.windows[i];
      uint64_t pos = at[i];
      ArenaCopy* cp = copies.data() + copy_at[i];
      auto place = [&](const char* src, uint32_t len) {
        *cp++ = ArenaCopy{src, pos, len};
        const uint64_t off = pos;
        pos += len;
        return off;
      };
      d.backbone_off = place(w.backbone, w.backbone_length);
      d.backbone_qual = place(w.backbone_quality, w.backbone_length);
      d.backbone_len = w.backbone_length;
      d.n_layers = q.shallow ? 0 : w.n_layers;
      d.layer_first = layer_at[i];
      d.offset = static_cast<uint32_t>(0.01 * w.backbone_length);
      d.flags = (w.tgs ? 1u : 0u) | (trim ? 2u : 0u);
      if (q.shallow) continue;
      // racon's layer order: std::sort of indices 1..n by begin (window.cpp:78-85)
      rank.resize(w.n_layers + 1);
      std::iota(rank.begin(), rank.end(), 0u);
      std::sort(rank.begin() + 1, rank.end(), [&](uint32_t lhs, uint32_t rhs) {
        return w.layers[lhs - 1].begin < w.layers[rhs - 1].begin;
      });
      PoaLayer* pl = P.layers.data() + layer_at[i];
      for (uint32_t k = 1; k <= w.n_layers; ++k) {
        const rh_layer& l = w.layers[rank[k] - 1];
        PoaLayer r{};
        r.seq_off = place(l.sequence, l.length);
        r.qual_off = l.quality != nullptr ? place(l.quality, l.length) : kNoQuality;
        r.len = l.length;
        r.begin = l.begin;
        r.end = l.end;
        *pl++ = r;
      }
      if (P.host_rejected[i]) continue;
      d.node_bound = static_cast<uint32_t>(q.node_bound);
      d.edge_bound = static_cast<uint32_t>(q.node_bound + w.n_layers);
      d.max_layer_len = q.max_len;
      d.out_cap = static_cast<uint32_t>(
          std::min<uint64_t>(q.node_bound, 2ull * w.backbone_length + 1024));
      d.out_off = out_at[i];
    }
  });
  for (uint32_t i = 0; i < n; ++i) {
    if (!P.shallow[i] && !P.host_rejected[i]) P.gpu_order.push_back(i);
  }
  P.out_bytes = out_at[n];
  P.arena_bytes = cur;
  bool staged = false;
  if (staging != nullptr) {
    staged = staging->reserve(std::max<uint64_t>(cur, 1)) == hipSuccess;
    if (!staged) (void)hipGetLastError();  // no pinned memory: the pageable path
  }
  if (staged) {
    P.staged_copies = std::move(copies);
    P.host_arena = staging->ptr;
  } else {
    P.arena.resize(cur);
    run_copies(P.arena.data(), copies, cur);
    P.host_arena = P.arena.data();
  }
  // largest estimated cost first (longest-processing-time order)
  std::stable_sort(P.gpu_order.begin(), P.gpu_order.end(), [&](uint32_t x, uint32_t y) {
    return window_cost(P.windows[x]) > window_cost(P.windows[y]);
  });
  return RH_OK;
}

hipError_t PoaLane::init(int32_t dev) {
  device = dev;
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess) e = hipStreamCreate(&stream);
  if (e == hipSuccess) e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  int limit = 0;
  if (e == hipSuccess) {
    e = hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  }
  lds_limit = static_cast<uint32_t>(limit);
  return e;
}

void PoaLane::release() {
  (void)hipSetDevice(device);
  if (stream != nullptr) (void)hipStreamSynchronize(stream);
  d_scratch.release();
  d_order.release();
  d_queue.release();
  d_params.release();
  if (ev0 != nullptr) (void)hipEventDestroy(ev0);
  if (ev1 != nullptr) (void)hipEventDestroy(ev1);
  if (stream != nullptr) (void)hipStreamDestroy(stream);
  ev0 = ev1 = nullptr;
  stream = nullptr;
}

bool plan_pass(const PoaWindow* host, const std::vector<uint32_t>& ws, uint32_t row_cap,
               uint32_t lds_limit, uint32_t tune_ring, bool wide, PassPlan* plan) {
  PoaSlotCaps caps{};
  uint32_t max_nodes = 1, max_edges = 1, max_cols = 1;
  for (uint32_t w : ws) {
    uint32_t nodes = host[w].node_bound, edges = host[w].edge_bound;
    if (row_cap != 0) {
      const uint64_t est = graph_estimate(host[w]);
      nodes = static_cast<uint32_t>(std::min<uint64_t>(nodes, est));
      edges = static_cast<uint32_t>(
          std::min<uint64_t>(edges, 2 * est + host[w].edge_bound - host[w].node_bound));
    }
    max_nodes = std::max(max_nodes, nodes);
    max_edges = std::max(max_edges, edges);
    max_cols = std::max(max_cols, host[w].max_layer_len + 1);
  }
  caps.nodes = max_nodes;
  caps.edges = max_edges;
  caps.cols = static_cast<uint32_t>(align_up(max_cols - 1 + kPoaColPad, 16));
  caps.rows = row_cap == 0 ? max_nodes + 1 : std::min<uint32_t>(max_nodes + 1, row_cap);
  caps.hrows = row_cap == 0 ? 0 : std::min<uint32_t>(caps.rows, kPoaHPoolRows);
  caps.aligned = wide ? kMaxAlignedWide : kMaxAligned;
  caps.stack = 2 * (max_edges + (caps.aligned + 1) * max_nodes) + 64;
  uint32_t ring = ring_rows_for(caps.cols, tune_ring);
  while (ring > 0 && poa_lds_bytes(caps.cols, ring) > lds_limit) --ring;
  if (ring == 1) ring = 0;  // one 16-bit row is no slot in 32-bit mode; keep none
  if (poa_lds_bytes(caps.cols, ring) > lds_limit) {
    return false;
  }
  plan->caps = caps;
  plan->ring = ring;
  plan->lds = poa_lds_bytes(caps.cols, ring);
  plan->wide = wide;
  return true;
}

int32_t size_pass(PoaLane& lane, uint64_t n_windows, const PassPlan& plan, int32_t mode,
                  const PoaConfig& c, PassLaunch* out) {
  int per_cu = 0, cus = 0;
  RH_HIP_CHECK(plan.wide ? poa_occupancy_wide(&per_cu, plan.lds)
                         : poa_occupancy(&per_cu, plan.lds),
               RH_ERR_DEVICE);
  RH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, lane.device),
               RH_ERR_DEVICE);
  uint64_t max_slots = static_cast<uint64_t>(std::max(1, per_cu)) * std::max(1, cus);
  max_slots = std::max<uint64_t>(1, max_slots / std::max<uint32_t>(1, c.slot_share));
  if (c.tune_slots != 0) {
    max_slots = c.tune_slots;
  }
  out->plan = plan;
  out->layout = poa_slot_layout(plan.caps);
  out->slot_bytes = align_up(out->layout.total, 4096);
  size_t free_b = 0, total_b = 0;
  RH_HIP_CHECK(hipMemGetInfo(&free_b, &total_b), RH_ERR_DEVICE);
  // keep the slots within ~60% of free memory (plus what the lane holds)
  const uint64_t budget = static_cast<uint64_t>((free_b + lane.d_scratch.bytes) * 0.6);
  uint64_t slots = std::max<uint64_t>(1, budget / out->slot_bytes);
  if (mode != RH_POA_STATIC) {
    slots = std::min<uint64_t>(slots, max_slots);
  }
  slots = std::min<uint64_t>(slots, n_windows);
  if (mode == RH_POA_PLAIN) {
    slots = 1;
  }
  out->slots = slots;
  return RH_OK;
}

int32_t plan_bins(PoaLane& lane, const PoaWindow* host, const std::vector<uint32_t>& ws,
                  uint32_t row_cap, int32_t mode, const PoaConfig& c,
                  std::vector<std::vector<uint32_t>>* orders, std::vector<PassLaunch>* passes,
                  std::vector<uint32_t>* refused, bool wide) {
  std::vector<uint32_t> bins[2];
  for (uint32_t w : ws) {
    bins[layer_bin(host[w])].push_back(w);
  }
  for (int k = 0; k < 2; ++k) {
    std::vector<uint32_t>& bin = bins[k];
    if (bin.empty()) continue;
    if (mode == RH_POA_STATIC) std::sort(bin.begin(), bin.end());
    PassPlan plan{};
    const uint32_t ring = c.bin_ring[k] != 0 ? c.bin_ring[k] : c.tune_ring;
    if (!plan_pass(host, bin, row_cap, lane.lds_limit, ring, wide, &plan)) {
      refused->insert(refused->end(), bin.begin(), bin.end());
      continue;
    }
    PassLaunch pl;
    const int32_t rc = size_pass(lane, bin.size(), plan, mode, c, &pl);
    if (rc != RH_OK) return rc;
    passes->push_back(pl);
    orders->push_back(std::move(bin));
  }
  return RH_OK;
}

int32_t launch_passes(PoaLane& lane, const PoaTables& t, const PoaConfig& c,
                      const std::vector<const std::vector<uint32_t>*>& orders,
                      const std::vector<PassLaunch>& passes, int32_t mode, uint32_t* launches) {
  const size_t np = passes.size();
  uint64_t scratch = 0, total_order = 0;
  for (size_t k = 0; k < np; ++k) {
    scratch = std::max(scratch, passes[k].slots * passes[k].slot_bytes);
    total_order += orders[k]->size();
  }
  RH_HIP_CHECK(lane.d_scratch.reserve(scratch), RH_ERR_MEMORY);
  RH_HIP_CHECK(lane.d_order.reserve(std::max<uint64_t>(total_order, 1) * sizeof(uint32_t)),
               RH_ERR_MEMORY);
  RH_HIP_CHECK(lane.d_queue.reserve(std::max<size_t>(np, 1) * sizeof(uint32_t)), RH_ERR_MEMORY);
  RH_HIP_CHECK(lane.d_params.reserve(std::max<size_t>(np, 1) * sizeof(PoaParams)),
               RH_ERR_MEMORY);
  RH_HIP_CHECK(hipEventRecord(lane.ev0, lane.stream), RH_ERR_DEVICE);
  if (np > 0) {
    RH_HIP_CHECK(hipMemsetAsync(lane.d_queue.ptr, 0, np * sizeof(uint32_t), lane.stream),
                 RH_ERR_DEVICE);
  }
  uint64_t order_at = 0;
  for (size_t k = 0; k < np; ++k) {
    const std::vector<uint32_t>& order = *orders[k];
    const PassLaunch& pl = passes[k];
    uint32_t* d_order = lane.d_order.as<uint32_t>() + order_at;
    order_at += order.size();
    RH_HIP_CHECK(hipMemcpyAsync(d_order, order.data(), order.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, lane.stream),
                 RH_ERR_DEVICE);
    PoaParams p{};
    p.windows = t.windows;
    p.layers = t.layers;
    p.arena = t.arena;
    p.out = t.out;
    p.results = t.results;
    p.order = d_order;
    p.n_order = static_cast<uint32_t>(order.size());
    p.queue_head = lane.d_queue.as<uint32_t>() + k;
    p.scratch = lane.d_scratch.as<uint8_t>();
    p.slot_bytes = pl.slot_bytes;
    p.caps = pl.plan.caps;
    p.layout = pl.layout;
    p.match = c.match;
    p.mismatch = c.mismatch;
    p.gap = c.gap;
    p.n_slots = static_cast<uint32_t>(pl.slots);
    p.ring_rows = pl.plan.ring;
    p.lds_bytes = pl.plan.lds;
    p.n_priority = priority_windows(t.host, order);
    p.band = c.band;
    PoaParams* dp = lane.d_params.as<PoaParams>() + k;
    RH_HIP_CHECK(hipMemcpyAsync(dp, &p, sizeof(PoaParams), hipMemcpyHostToDevice, lane.stream),
                 RH_ERR_DEVICE);
    if (mode == RH_POA_PLAIN) {
      for (uint32_t q = 0; q < order.size(); ++q) {
        RH_HIP_CHECK(pl.plan.wide ? launch_poa_single_wide(dp, p.lds_bytes, q, lane.stream)
                                  : launch_poa_single(dp, p.lds_bytes, q, lane.stream),
                     RH_ERR_KERNEL);
        ++*launches;
      }
    } else if (mode == RH_POA_STATIC) {
      // one workgroup (and slot) per window; launches of `slots` windows when
      // the slots of all windows do not fit in memory at once
      for (uint64_t first = 0; first < order.size(); first += pl.slots) {
        const uint64_t count = std::min<uint64_t>(pl.slots, order.size() - first);
        const uint32_t f = static_cast<uint32_t>(first), n = static_cast<uint32_t>(count);
        RH_HIP_CHECK(pl.plan.wide ? launch_poa_static_wide(dp, p.lds_bytes, f, n, lane.stream)
                                  : launch_poa_static(dp, p.lds_bytes, f, n, lane.stream),
                     RH_ERR_KERNEL);
        ++*launches;
      }
    } else {
      const uint32_t grid = static_cast<uint32_t>(pl.slots);
      RH_HIP_CHECK(pl.plan.wide ? launch_poa_batch_wide(dp, p.lds_bytes, grid, lane.stream)
                                : launch_poa_batch(dp, p.lds_bytes, grid, lane.stream),
                   RH_ERR_KERNEL);
      ++*launches;
    }
  }
  RH_HIP_CHECK(hipEventRecord(lane.ev1, lane.stream), RH_ERR_DEVICE);
  return RH_OK;
}

int32_t plan_retries(PoaLane& lane, const PoaWindow* host, const std::vector<uint32_t>& ws,
                     const std::vector<int32_t>& status, const PoaConfig& c,
                     std::vector<std::vector<uint32_t>>* orders,
                     std::vector<PassLaunch>* passes, std::vector<uint32_t>* refused) {
  std::vector<uint32_t> narrow, wide;
  for (size_t k = 0; k < ws.size(); ++k) {
    (status[k] == kPoaAlignedFull ? wide : narrow).push_back(ws[k]);
  }
  int32_t rc = plan_bins(lane, host, narrow, 0, RH_POA_BATCHED, c, orders, passes, refused);
  if (rc != RH_OK) return rc;
  return plan_bins(lane, host, wide, 0, RH_POA_BATCHED, c, orders, passes, refused, true);
}

int32_t finish_passes(PoaLane& lane, double* kernel_ms) {
  RH_HIP_CHECK(hipEventSynchronize(lane.ev1), RH_ERR_KERNEL);
  float ms = 0;
  RH_HIP_CHECK(hipEventElapsedTime(&ms, lane.ev0, lane.ev1), RH_ERR_DEVICE);
  *kernel_ms += ms;
  return RH_OK;
}

}  // namespace rh

struct rh_poa_batch {
  bool trim = true;
  rh::PoaConfig config;
  rh::PoaLane lane;

  rh::PackedWindows packed;  // host copy of the batch
  std::vector<rh::PoaResult> results;

  rh::DeviceBuffer d_arena, d_windows, d_layers, d_out, d_results;
  // pinned staging of the input bytes, from the batch's second upload on
  // (a batch object that is reused pays the pinning once); it is allocated
  // on a helper thread right after the first upload, so that the pinning
  // runs beside the first run instead of inside the second upload
  rh::PinnedBuffer staging;
  std::thread staging_maker;
  uint64_t uploads = 0;
  bool uploaded = false;
  bool ran = false;
  bool submitted = false;  // between rh_poa_submit and rh_poa_wait
  bool launched = false;   // the submit queued kernel passes
  std::vector<std::vector<uint32_t>> pass_orders;  // window order of each queued pass
  rh_poa_stats stats{};

  rh::PoaTables tables() const {
    rh::PoaTables t;
    t.host = packed.windows.data();
    t.windows = d_windows.as<rh::PoaWindow>();
    t.layers = d_layers.as<rh::PoaLayer>();
    t.arena = d_arena.as<uint8_t>();
    t.out = d_out.as<uint8_t>();
    t.results = d_results.as<rh::PoaResult>();
    return t;
  }
};

namespace {

using rh::PassLaunch;
using rh::PoaResult;
using rh::PoaWindow;

PoaResult make_result() {
  PoaResult r{};
  r.status = rh::kPoaNotRun;
  return r;
}

}  // namespace

extern "C" {

const char* rh_last_error(void) { return rh::last_error(); }

int32_t rh_version(void) { return 100; }

int32_t rh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    return 0;
  }
  return n;
}

int32_t rh_poa_create(int32_t device, int8_t match, int8_t mismatch, int8_t gap, int32_t trim,
                      rh_poa_batch** batch) {
  rh::clear_error();
  if (batch == nullptr) {
    rh::set_error("rh_poa_create: null batch pointer");
    return RH_ERR_ARGUMENT;
  }
  *batch = nullptr;
  int n = 0;
  RH_HIP_CHECK(hipGetDeviceCount(&n), RH_ERR_DEVICE);
  if (device < 0 || device >= n) {
    rh::set_error("rh_poa_create: no such device");
    return RH_ERR_DEVICE;
  }
  rh_poa_batch* b = new rh_poa_batch();
  b->config.match = match;
  b->config.mismatch = mismatch;
  b->config.gap = gap;
  b->trim = trim != 0;
  if (b->lane.init(device) != hipSuccess) {
    rh::set_error("rh_poa_create: stream/event creation failed");
    b->lane.release();
    delete b;
    return RH_ERR_DEVICE;
  }
  *batch = b;
  return RH_OK;
}

void rh_poa_destroy(rh_poa_batch* b) {
  if (b == nullptr) {
    return;
  }
  b->lane.release();
  b->d_arena.release();
  b->d_windows.release();
  b->d_layers.release();
  b->d_out.release();
  b->d_results.release();
  if (b->staging_maker.joinable()) b->staging_maker.join();
  b->staging.release();
  delete b;
}

int32_t rh_poa_set_tuning(rh_poa_batch* b, uint32_t slots, uint32_t rows, uint32_t ring_rows) {
  if (b == nullptr) {
    rh::set_error("rh_poa_set_tuning: null batch");
    return RH_ERR_ARGUMENT;
  }
  b->config.tune_slots = slots;
  b->config.tune_rows = rows;
  b->config.tune_ring = ring_rows;
  return RH_OK;
}

int32_t rh_poa_upload(rh_poa_batch* b, uint32_t n, const rh_window* windows) {
  rh::clear_error();
  if (b == nullptr || (n > 0 && windows == nullptr)) {
    rh::set_error("rh_poa_upload: null argument");
    return RH_ERR_ARGUMENT;
  }
  if (b->submitted) {
    rh::set_error("rh_poa_upload: the batch is running (call rh_poa_wait first)");
    return RH_ERR_STATE;
  }
  RH_HIP_CHECK(hipSetDevice(b->lane.device), RH_ERR_DEVICE);
  auto t0 = std::chrono::steady_clock::now();
  b->stats = rh_poa_stats{};
  b->uploaded = false;
  if (b->staging_maker.joinable()) b->staging_maker.join();
  const int32_t rc =
      rh::pack_windows(n, windows, b->trim, &b->packed, b->uploads > 0 ? &b->staging : nullptr);
  if (rc != RH_OK) {
    rh::set_error(std::string("rh_poa_upload: ") + rh::last_error());
    return rc;
  }
  ++b->uploads;
  const rh::PackedWindows& P = b->packed;
  b->results.assign(n, make_result());
  for (uint32_t i = 0; i < n; ++i) {
    if (P.host_rejected[i]) b->results[i].status = rh::kPoaRejectedOnHost;
  }
  b->stats.pack_ms = rh::ms_since(t0);
  b->stats.windows = n;
  b->stats.gpu_windows = P.gpu_order.size();
  b->stats.input_bytes = P.arena_bytes;

  auto t1 = std::chrono::steady_clock::now();
  hipStream_t st = b->lane.stream;
  RH_HIP_CHECK(b->d_arena.reserve(std::max<uint64_t>(P.arena_bytes, 1), true), RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_windows.reserve(std::max<uint64_t>(n, 1) * sizeof(PoaWindow), true),
               RH_ERR_MEMORY);
  RH_HIP_CHECK(
      b->d_layers.reserve(std::max<size_t>(P.layers.size(), 1) * sizeof(rh::PoaLayer), true),
      RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_results.reserve(std::max<uint64_t>(n, 1) * sizeof(PoaResult), true),
               RH_ERR_MEMORY);
  RH_HIP_CHECK(b->d_out.reserve(std::max<uint64_t>(P.out_bytes, 1), true), RH_ERR_MEMORY);
  if (!P.staged_copies.empty()) {
    RH_HIP_CHECK(rh::copy_and_upload(b->staging.ptr, P.staged_copies, P.arena_bytes,
                                     b->d_arena.as<uint8_t>(), st),
                 RH_ERR_DEVICE);
  } else if (P.arena_bytes > 0) {
    RH_HIP_CHECK(hipMemcpyAsync(b->d_arena.ptr, P.host_arena, P.arena_bytes,
                                hipMemcpyHostToDevice, st),
                 RH_ERR_DEVICE);
  }
  if (n > 0) {
    RH_HIP_CHECK(hipMemcpyAsync(b->d_windows.ptr, P.windows.data(), n * sizeof(PoaWindow),
                                hipMemcpyHostToDevice, st),
                 RH_ERR_DEVICE);
  }
  if (!P.layers.empty()) {
    RH_HIP_CHECK(hipMemcpyAsync(b->d_layers.ptr, P.layers.data(),
                                P.layers.size() * sizeof(rh::PoaLayer), hipMemcpyHostToDevice,
                                st),
                 RH_ERR_DEVICE);
  }
  RH_HIP_CHECK(hipStreamSynchronize(st), RH_ERR_DEVICE);
  b->stats.h2d_ms = rh::ms_since(t1);
  b->uploaded = true;
  b->ran = false;
  if (b->uploads == 1 && P.arena_bytes > 0) {
    // (a failed pinning leaves the staging empty: the next upload tries
    // again and falls back to pageable copies)
    const uint64_t bytes = P.arena_bytes;
    rh::PinnedBuffer* staging = &b->staging;
    b->staging_maker = std::thread([staging, bytes] {
      if (staging->reserve(bytes) != hipSuccess) (void)hipGetLastError();
    });
  }
  return RH_OK;
}

int32_t rh_poa_submit(rh_poa_batch* b, int32_t mode) {
  rh::clear_error();
  if (b == nullptr || !b->uploaded) {
    rh::set_error("rh_poa_submit: nothing uploaded");
    return RH_ERR_STATE;
  }
  if (b->submitted) {
    rh::set_error("rh_poa_submit: the batch is already running (call rh_poa_wait)");
    return RH_ERR_STATE;
  }
  if (mode != RH_POA_BATCHED && mode != RH_POA_PLAIN && mode != RH_POA_STATIC) {
    rh::set_error("rh_poa_submit: unknown mode");
    return RH_ERR_ARGUMENT;
  }
  RH_HIP_CHECK(hipSetDevice(b->lane.device), RH_ERR_DEVICE);
  const uint32_t n = static_cast<uint32_t>(b->packed.windows.size());
  b->stats.kernel_ms = 0;
  b->stats.launches = 0;
  b->stats.retried = 0;
  b->stats.rejected = 0;
  b->stats.band_reruns = 0;
  b->ran = false;
  b->pass_orders.clear();
  b->submitted = true;
  if (n == 0) {
    return RH_OK;
  }
  // fresh results: every kernel window starts as not-run, the ones refused
  // at upload keep their status
  for (uint32_t w = 0; w < n; ++w) {
    if (b->results[w].status != rh::kPoaRejectedOnHost) b->results[w] = make_result();
  }
  // bins by layer length: the main pass and the long-layer pass (windows in
  // the cost order of the queue either way, in upload order for STATIC)
  const uint32_t row_cap = b->config.tune_rows ? b->config.tune_rows : 4096;
  std::vector<PassLaunch> passes;
  std::vector<uint32_t> refused;
  int32_t rc = rh::plan_bins(b->lane, b->packed.windows.data(), b->packed.gpu_order, row_cap,
                             mode, b->config, &b->pass_orders, &passes, &refused);
  if (rc != RH_OK) return rc;
  for (uint32_t w : refused) b->results[w].status = rh::kPoaRejectedOnHost;
  RH_HIP_CHECK(hipMemcpyAsync(b->d_results.ptr, b->results.data(), n * sizeof(PoaResult),
                              hipMemcpyHostToDevice, b->lane.stream),
               RH_ERR_DEVICE);
  std::vector<const std::vector<uint32_t>*> orders;
  for (const auto& o : b->pass_orders) orders.push_back(&o);
  if (!passes.empty()) {
    b->stats.slots = static_cast<uint32_t>(passes[0].slots);
    rc = rh::launch_passes(b->lane, b->tables(), b->config, orders, passes, mode,
                           &b->stats.launches);
    if (rc != RH_OK) return rc;
  }
  b->launched = !passes.empty();
  return RH_OK;
}

int32_t rh_poa_wait(rh_poa_batch* b) {
  rh::clear_error();
  if (b == nullptr || !b->submitted) {
    rh::set_error("rh_poa_wait: nothing submitted");
    return RH_ERR_STATE;
  }
  b->submitted = false;
  RH_HIP_CHECK(hipSetDevice(b->lane.device), RH_ERR_DEVICE);
  std::vector<PoaWindow>& windows = b->packed.windows;
  const uint32_t n = static_cast<uint32_t>(windows.size());
  if (n == 0) {
    b->ran = true;
    return RH_OK;
  }
  double kernel_ms = 0;
  if (b->launched) {
    const int32_t rc = rh::finish_passes(b->lane, &kernel_ms);
    if (rc != RH_OK) return rc;
  }
  RH_HIP_CHECK(rh::copy_sync(b->results.data(), b->d_results.ptr, n * sizeof(PoaResult),
                             hipMemcpyDeviceToHost, b->lane.stream),
               RH_ERR_DEVICE);

  // windows that outgrew their slot: rerun with exact bounds (rows = node
  // bound + 1), in one pass per layer-length bin
  std::vector<uint32_t> retry;
  for (uint32_t w : b->packed.gpu_order) {
    if (rh::is_retry_status(b->results[w].status)) retry.push_back(w);
    if (b->results[w].status == rh::kPoaBandMiss) b->stats.band_reruns++;
  }
  if (!retry.empty()) {
    b->stats.retried = retry.size();
    uint64_t extra_out = 0;
    for (uint32_t w : retry) extra_out += rh::align_up(windows[w].node_bound, 16);
    // bigger output regions for the retried windows, appended to d_out
    const uint64_t old_out = b->packed.out_bytes;
    rh::DeviceBuffer grown;
    RH_HIP_CHECK(grown.reserve(old_out + extra_out), RH_ERR_MEMORY);
    RH_HIP_CHECK(rh::copy_sync(grown.ptr, b->d_out.ptr, old_out, hipMemcpyDeviceToDevice,
                               b->lane.stream),
                 RH_ERR_DEVICE);
    b->d_out.release();
    b->d_out = grown;
    uint64_t off = old_out;
    for (uint32_t w : retry) {
      windows[w].out_off = off;
      windows[w].out_cap = windows[w].node_bound;
      off += rh::align_up(windows[w].node_bound, 16);
      RH_HIP_CHECK(rh::copy_sync(b->d_windows.as<PoaWindow>() + w, &windows[w],
                                 sizeof(PoaWindow), hipMemcpyHostToDevice, b->lane.stream),
                   RH_ERR_DEVICE);
    }
    b->packed.out_bytes = off;
    std::vector<std::vector<uint32_t>> retry_orders;
    std::vector<PassLaunch> passes;
    std::vector<uint32_t> refused;  // stay in their overflow status: rejected below
    rh::PoaConfig exact = b->config;
    exact.band = 0;  // a window whose path left the band reruns unbanded
    std::vector<int32_t> why;
    for (uint32_t w : retry) why.push_back(b->results[w].status);
    int32_t rc = rh::plan_retries(b->lane, windows.data(), retry, why, exact, &retry_orders,
                                  &passes, &refused);
    if (rc != RH_OK) return rc;
    if (!passes.empty()) {
      std::vector<const std::vector<uint32_t>*> orders;
      for (const auto& o : retry_orders) orders.push_back(&o);
      rc = rh::launch_passes(b->lane, b->tables(), exact, orders, passes, RH_POA_BATCHED,
                             &b->stats.launches);
      if (rc != RH_OK) return rc;
      rc = rh::finish_passes(b->lane, &kernel_ms);
      if (rc != RH_OK) return rc;
    }
    RH_HIP_CHECK(rh::copy_sync(b->results.data(), b->d_results.ptr, n * sizeof(PoaResult),
                               hipMemcpyDeviceToHost, b->lane.stream),
                 RH_ERR_DEVICE);
  }
  b->stats.kernel_ms = kernel_ms;
  b->stats.cells = 0;
  for (uint32_t w = 0; w < n; ++w) {
    if (b->packed.shallow[w]) continue;
    const PoaResult& r = b->results[w];
    if (r.status == rh::kPoaPolished || r.status == rh::kPoaPolishedUntrimmed) {
      b->stats.cells += r.cells;
    } else {
      b->stats.rejected++;  // outside the device limits: racon's CPU path takes it
    }
  }
  b->ran = true;
  return RH_OK;
}

int32_t rh_poa_run(rh_poa_batch* b, int32_t mode) {
  const int32_t rc = rh_poa_submit(b, mode);
  if (rc != RH_OK) {
    if (b != nullptr) b->submitted = false;
    return rc;
  }
  return rh_poa_wait(b);
}

int32_t rh_poa_set_band(rh_poa_batch* b, uint32_t band) {
  rh::clear_error();
  if (b == nullptr || (band != 0 && band != 256)) {
    rh::set_error("rh_poa_set_band: null batch or a band other than 0 and 256");
    return RH_ERR_ARGUMENT;
  }
  b->config.band = band;
  return RH_OK;
}

int32_t rh_poa_autotune(rh_poa_batch* b, rh_poa_tuning* out) {
  rh::clear_error();
  if (b == nullptr || out == nullptr || !b->uploaded) {
    rh::set_error("rh_poa_autotune: null argument or nothing uploaded");
    return RH_ERR_ARGUMENT;
  }
  if (b->submitted) {
    rh::set_error("rh_poa_autotune: the batch is running (call rh_poa_wait first)");
    return RH_ERR_STATE;
  }
  RH_HIP_CHECK(hipSetDevice(b->lane.device), RH_ERR_DEVICE);
  *out = rh_poa_tuning{};
  std::vector<uint32_t> bins[2];
  for (uint32_t w : b->packed.gpu_order) bins[rh::layer_bin(b->packed.windows[w])].push_back(w);
  // DP rows kept in LDS per wave: deeper rings keep more rows on chip but
  // leave room for fewer waves per CU (occupancy follows the LDS bytes)
  static const uint32_t kMainRings[] = {0, 4, 6, 8, 10, 12, 16};
  static const uint32_t kLongRings[] = {0, 2, 3, 4, 6, 8};
  const uint32_t row_cap = b->config.tune_rows ? b->config.tune_rows : 4096;
  for (int k = 0; k < 2; ++k) {
    if (bins[k].empty()) continue;
    const uint32_t* rings = k == 0 ? kMainRings : kLongRings;
    const size_t n_rings = k == 0 ? sizeof(kMainRings) / 4 : sizeof(kLongRings) / 4;
    double best = -1;
    uint32_t best_ring = 0;
    std::vector<uint32_t> tried;
    for (size_t r = 0; r < n_rings; ++r) {
      rh::PoaConfig c = b->config;
      c.bin_ring[k] = rings[r];
      std::vector<std::vector<uint32_t>> orders;
      std::vector<PassLaunch> passes;
      std::vector<uint32_t> refused;
      int32_t rc = rh::plan_bins(b->lane, b->packed.windows.data(), bins[k], row_cap,
                                 RH_POA_BATCHED, c, &orders, &passes, &refused);
      if (rc != RH_OK) return rc;
      if (passes.empty()) break;  // the bin does not fit the LDS limit at all
      const uint32_t ring = passes[0].plan.ring;  // after the LDS limit
      if (std::find(tried.begin(), tried.end(), ring) != tried.end()) continue;
      tried.push_back(ring);
      std::vector<const std::vector<uint32_t>*> ptrs{&orders[0]};
      double ms = 0;
      rc = rh::launch_passes(b->lane, b->tables(), c, ptrs, passes, RH_POA_BATCHED,
                             &b->stats.launches);
      if (rc == RH_OK) rc = rh::finish_passes(b->lane, &ms);
      if (rc != RH_OK) return rc;
      out->candidates++;
      if (r == 0) out->default_ms[k] = ms;
      if (best < 0 || ms < best) {
        best = ms;
        best_ring = rings[r];
        out->ring_rows[k] = ring;
      }
    }
    b->config.bin_ring[k] = best_ring;
    out->kernel_ms[k] = best < 0 ? 0 : best;
  }
  b->ran = false;  // the timing runs leave no complete results: run the batch
  return RH_OK;
}

int32_t rh_poa_capacity(int32_t device, uint64_t* free_bytes, uint64_t* total_bytes) {
  rh::clear_error();
  if (free_bytes == nullptr || total_bytes == nullptr) {
    rh::set_error("rh_poa_capacity: null argument");
    return RH_ERR_ARGUMENT;
  }
  RH_HIP_CHECK(hipSetDevice(device), RH_ERR_DEVICE);
  size_t f = 0, t = 0;
  RH_HIP_CHECK(hipMemGetInfo(&f, &t), RH_ERR_DEVICE);
  *free_bytes = f;
  *total_bytes = t;
  return RH_OK;
}

uint64_t rh_poa_device_bytes(const rh_poa_batch* b) {
  if (b == nullptr) return 0;
  return b->d_arena.bytes + b->d_windows.bytes + b->d_layers.bytes + b->d_out.bytes +
         b->d_results.bytes + b->lane.device_bytes();
}

uint64_t rh_poa_output_bound(const rh_poa_batch* b) {
  if (b == nullptr) {
    return 0;
  }
  const rh::PackedWindows& P = b->packed;
  uint64_t total = 0;
  for (size_t i = 0; i < P.windows.size(); ++i) {
    const bool host_rejected = b->results[i].status == rh::kPoaRejectedOnHost;
    total += P.shallow[i] || host_rejected ? P.windows[i].backbone_len : P.windows[i].out_cap;
  }
  return total;
}

int32_t rh_poa_download(rh_poa_batch* b, char* out, uint64_t capacity, uint64_t* offsets,
                        uint32_t* lengths, uint8_t* status) {
  rh::clear_error();
  if (b == nullptr || !b->ran) {
    rh::set_error("rh_poa_download: batch has not run");
    return RH_ERR_STATE;
  }
  if (out == nullptr || offsets == nullptr || lengths == nullptr || status == nullptr) {
    rh::set_error("rh_poa_download: null argument");
    return RH_ERR_ARGUMENT;
  }
  RH_HIP_CHECK(hipSetDevice(b->lane.device), RH_ERR_DEVICE);
  auto t0 = std::chrono::steady_clock::now();
  const rh::PackedWindows& P = b->packed;
  std::vector<uint8_t> dev_out(P.out_bytes);
  if (P.out_bytes > 0) {
    RH_HIP_CHECK(rh::copy_sync(dev_out.data(), b->d_out.ptr, P.out_bytes,
                               hipMemcpyDeviceToHost, b->lane.stream),
                 RH_ERR_DEVICE);
  }
  uint64_t pos = 0;
  for (size_t i = 0; i < P.windows.size(); ++i) {
    const uint8_t* src;
    uint32_t len;
    uint8_t st;
    const int32_t rs = b->results[i].status;
    if (P.shallow[i]) {
      src = P.host_arena + P.windows[i].backbone_off;
      len = P.windows[i].backbone_len;
      st = RH_WINDOW_UNPOLISHED;
    } else if (rs != rh::kPoaPolished && rs != rh::kPoaPolishedUntrimmed) {
      // outside the device limits: the backbone, for the caller's CPU path
      src = P.host_arena + P.windows[i].backbone_off;
      len = P.windows[i].backbone_len;
      st = RH_WINDOW_REJECTED;
    } else {
      src = dev_out.data() + P.windows[i].out_off;
      len = b->results[i].length;
      st = rs == rh::kPoaPolishedUntrimmed ? RH_WINDOW_POLISHED_UNTRIMMED : RH_WINDOW_POLISHED;
    }
    if (pos + len > capacity) {
      rh::set_error("rh_poa_download: output capacity too small");
      return RH_ERR_ARGUMENT;
    }
    memcpy(out + pos, src, len);
    offsets[i] = pos;
    lengths[i] = len;
    status[i] = st;
    pos += len;
  }
  b->stats.d2h_ms = rh::ms_since(t0);
  return RH_OK;
}

int32_t rh_poa_phase_ticks(const rh_poa_batch* b, uint64_t* ticks, uint32_t n,
                           uint64_t* ticks_per_second) {
  if (b == nullptr || ticks == nullptr) {
    rh::set_error("rh_poa_phase_ticks: null argument");
    return RH_ERR_ARGUMENT;
  }
  if (ticks_per_second != nullptr) {
    int khz = 0;
    RH_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->lane.device),
                 RH_ERR_DEVICE);
    *ticks_per_second = static_cast<uint64_t>(khz) * 1000ull;
  }
  for (uint32_t k = 0; k < n; ++k) ticks[k] = 0;
  for (uint32_t w : b->packed.gpu_order) {
    for (uint32_t k = 0; k < n && k < static_cast<uint32_t>(rh::kPoaPhases); ++k) {
      ticks[k] += b->results[w].phase[k];
    }
  }
  return RH_OK;
}

int32_t rh_poa_window_costs(const rh_poa_batch* b, uint64_t* cells, uint64_t* ticks,
                            uint64_t n) {
  if (b == nullptr || cells == nullptr || ticks == nullptr) {
    rh::set_error("rh_poa_window_costs: null argument");
    return RH_ERR_ARGUMENT;
  }
  if (n > b->results.size()) {
    rh::set_error("rh_poa_window_costs: more windows asked for than uploaded");
    return RH_ERR_ARGUMENT;
  }
  for (uint64_t w = 0; w < n; ++w) {
    const rh::PoaResult& r = b->results[w];
    const bool ran = r.status != rh::kPoaNotRun;
    cells[w] = ran ? r.cells : 0;
    uint64_t t = 0;
    for (int k = 0; k < rh::kPoaPhases; ++k) t += r.phase[k];
    ticks[w] = ran ? t : 0;
  }
  return RH_OK;
}

int32_t rh_poa_get_stats(const rh_poa_batch* b, rh_poa_stats* stats) {
  if (b == nullptr || stats == nullptr) {
    rh::set_error("rh_poa_get_stats: null argument");
    return RH_ERR_ARGUMENT;
  }
  *stats = b->stats;
  return RH_OK;
}

}  // extern "C"
