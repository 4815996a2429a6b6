// C ABI: the serving loop (racon_hip.h, rh_server_*). No counterpart in
// racon, which polishes one job per process.
//
// Requests (windows of polishing jobs) are packed on the caller's thread,
// admitted first come, first served into a device pool of the server's
// budget, where they stay resident until their last window is back, and
// served by executor threads: each owns a HIP stream with its share of the
// persistent kernel's resident slots and launches the next windows of the
// admitted requests (admission order, largest estimated cost first inside a
// launch). Kernels of different executors run side by side on the device,
// so a request that arrives while others are on the GPU starts as soon as
// an executor is free and its bytes fit.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "poa_host.hpp"

namespace {

using rh::PoaLayer;
using rh::PoaResult;
using rh::PoaWindow;
using Clock = std::chrono::steady_clock;

constexpr uint64_t kRegionAlign = 256;

// First-fit allocator over [0, capacity).
class RangePool {
 public:
  void reset(uint64_t capacity) {
    free_.clear();
    if (capacity > 0) free_[0] = capacity;
    capacity_ = capacity;
    used_ = 0;
  }
  bool alloc(uint64_t n, uint64_t* at) {
    if (n == 0) {
      *at = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *at = it->first;
      const uint64_t start = it->first + n, rest = it->second - n;
      free_.erase(it);
      if (rest > 0) free_[start] = rest;
      used_ += n;
      return true;
    }
    return false;
  }
  void release(uint64_t at, uint64_t n) {
    if (n == 0) return;
    used_ -= n;
    uint64_t start = at, len = n;
    auto next = free_.lower_bound(at);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == at) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && at + n == next->first) {
      len += next->second;
      free_.erase(next);
    }
    free_[start] = len;
  }
  uint64_t used() const { return used_; }
  uint64_t capacity() const { return capacity_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t capacity_ = 0, used_ = 0;
};

enum RequestState : uint32_t { kWaiting = 0, kAdmitted = 1, kDone = 2 };

struct Request {
  uint64_t id = 0;
  uint32_t n = 0;
  rh::PackedWindows packed;  // records rebased to the pool once admitted
  // pool region: input arena | layer records | consensus space
  uint64_t region_at = 0, region_bytes = 0, layers_at = 0, out_at = 0;
  uint64_t win_at = 0;  // first window id
  uint64_t device_bytes = 0;
  size_t next = 0;  // gpu_order position of the next window to launch
  size_t left = 0;  // kernel windows not back yet
  std::vector<PoaResult> results;
  std::vector<std::pair<uint64_t, uint64_t>> extra;  // retried windows' consensus (at, bytes)
  std::vector<uint8_t> out_host;                     // consensus space, downloaded at the end
  std::unordered_map<uint32_t, std::vector<uint8_t>> retried_out;  // by window
  uint32_t state = kWaiting;
  int32_t error = RH_OK;
  std::string error_text;
  double arrival = 0, admitted = 0, done = 0;
  uint32_t launches = 0, shared = 0;
};

struct Executor {
  rh::PoaLane lane;
  rh::PoaConfig config;
  std::thread thread;
};

uint64_t record_bytes() { return sizeof(PoaWindow) + sizeof(PoaResult); }

}  // namespace

struct rh_server {
  int32_t device = 0;
  bool trim = true;
  rh::PoaConfig config;
  uint32_t launch_windows = 2048;

  rh::DeviceBuffer d_pool, d_windows, d_results;
  RangePool pool, ids;
  uint64_t window_capacity = 0;
  std::vector<PoaWindow> host_windows;  // by window id (planning)
  hipStream_t upload_stream = nullptr;

  std::vector<std::unique_ptr<Executor>> executors;
  std::thread admitter;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> waiting;  // submitted, not admitted
  std::deque<Request*> ready;    // admitted, windows left to launch
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  uint64_t next_id = 1;
  uint32_t uploading = 0;
  bool closing = false;
  bool held = false;     // executors start no new launch
  bool started = false;  // a request was submitted (warm is refused after)
  rh_server_stats stats{};
  Clock::time_point t0 = Clock::now();

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  rh::PoaTables tables() const {
    rh::PoaTables t;
    t.host = host_windows.data();
    t.windows = d_windows.as<PoaWindow>();
    t.layers = d_pool.as<PoaLayer>();
    t.arena = d_pool.as<uint8_t>();
    t.out = d_pool.as<uint8_t>();
    t.results = d_results.as<PoaResult>();
    return t;
  }
};

namespace {

// Region layout of a packed request inside the pool.
void layout_region(Request* r) {
  const rh::PackedWindows& P = r->packed;
  const uint64_t arena = rh::align_up(std::max<uint64_t>(P.arena.size(), 1), kRegionAlign);
  const uint64_t layers = rh::align_up(P.layers.size() * sizeof(PoaLayer), kRegionAlign);
  const uint64_t out = rh::align_up(P.out_bytes, kRegionAlign);
  r->layers_at = arena;  // relative until admitted
  r->out_at = arena + layers;
  r->region_bytes = arena + layers + out;
  r->device_bytes = r->region_bytes + uint64_t(r->n) * record_bytes();
}

// Reserves pool bytes and window ids for the request at the head of the
// queue. Caller holds the lock.
bool reserve(rh_server* s, Request* r) {
  uint64_t at = 0, win = 0;
  if (!s->pool.alloc(r->region_bytes, &at)) return false;
  if (!s->ids.alloc(r->n, &win)) {
    s->pool.release(at, r->region_bytes);
    return false;
  }
  r->region_at = at;
  r->win_at = win;
  return true;
}

void release_request(rh_server* s, Request* r) {
  s->pool.release(r->region_at, r->region_bytes);
  s->ids.release(r->win_at, r->n);
  for (const auto& e : r->extra) s->pool.release(e.first, e.second);
  r->extra.clear();
}

void finish(rh_server* s, Request* r) {  // caller holds the lock
  release_request(s, r);
  r->state = kDone;
  r->done = s->now();
  s->stats.completed++;
  s->cv.notify_all();
}

// Rebases the request's records to its pool region and window ids, and
// copies inputs and records to the device (upload stream).
int32_t upload(rh_server* s, Request* r) {
  rh::PackedWindows& P = r->packed;
  const uint64_t base = r->region_at;
  const uint64_t layer_first = (base + r->layers_at) / sizeof(PoaLayer);
  for (PoaLayer& l : P.layers) {
    l.seq_off += base;
    if (l.qual_off != rh::kNoQuality) l.qual_off += base;
  }
  for (uint32_t i = 0; i < r->n; ++i) {
    PoaWindow& w = P.windows[i];
    w.backbone_off += base;
    w.backbone_qual += base;
    w.layer_first += static_cast<uint32_t>(layer_first);
    w.out_off += base + r->out_at;
    s->host_windows[r->win_at + i] = w;
  }
  uint8_t* pool = s->d_pool.as<uint8_t>();
  if (!P.arena.empty()) {
    RH_HIP_CHECK(hipMemcpyAsync(pool + base, P.arena.data(), P.arena.size(),
                                hipMemcpyHostToDevice, s->upload_stream),
                 RH_ERR_DEVICE);
  }
  if (!P.layers.empty()) {
    RH_HIP_CHECK(hipMemcpyAsync(pool + base + r->layers_at, P.layers.data(),
                                P.layers.size() * sizeof(PoaLayer), hipMemcpyHostToDevice,
                                s->upload_stream),
                 RH_ERR_DEVICE);
  }
  if (r->n > 0) {
    RH_HIP_CHECK(hipMemcpyAsync(s->d_windows.as<PoaWindow>() + r->win_at, P.windows.data(),
                                r->n * sizeof(PoaWindow), hipMemcpyHostToDevice,
                                s->upload_stream),
                 RH_ERR_DEVICE);
  }
  RH_HIP_CHECK(hipStreamSynchronize(s->upload_stream), RH_ERR_DEVICE);
  return RH_OK;
}

void admitter_loop(rh_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  for (;;) {
    Request* r = nullptr;
    s->cv.wait(lock, [&] {
      return (!s->waiting.empty() && reserve(s, s->waiting.front())) ||
             (s->closing && s->waiting.empty());
    });
    if (s->waiting.empty()) return;  // closing and nothing left to admit
    r = s->waiting.front();
    s->waiting.pop_front();
    s->uploading++;
    lock.unlock();
    const int32_t rc = upload(s, r);
    const std::string text = rc == RH_OK ? std::string() : rh::last_error();
    lock.lock();
    s->uploading--;
    r->admitted = s->now();
    s->stats.admitted++;
    if (rc != RH_OK) {
      r->error = rc;
      r->error_text = "upload: " + text;
      finish(s, r);
      continue;
    }
    r->state = kAdmitted;
    if (r->left == 0) {
      finish(s, r);  // nothing for the kernel (shallow or refused windows)
    } else {
      s->ready.push_back(r);
    }
    s->cv.notify_all();
  }
}

// One launch over the windows `ws_in` (ids): largest estimated cost first,
// a pass per layer-length bin, waited for.
int32_t run_launch(rh_server* s, Executor* ex, const std::vector<uint32_t>& ws_in,
                   double* kernel_ms) {
  rh::PoaLane& lane = ex->lane;
  std::vector<uint32_t> ws = ws_in;
  std::stable_sort(ws.begin(), ws.end(), [&](uint32_t x, uint32_t y) {
    return rh::window_cost(s->host_windows[x]) > rh::window_cost(s->host_windows[y]);
  });
  const uint32_t row_cap = ex->config.tune_rows ? ex->config.tune_rows : 4096;
  std::vector<std::vector<uint32_t>> orders;
  std::vector<rh::PassLaunch> passes;
  std::vector<uint32_t> refused;
  int32_t rc = rh::plan_bins(lane, s->host_windows.data(), ws, row_cap, RH_POA_BATCHED,
                             ex->config, &orders, &passes, &refused);
  if (rc != RH_OK) return rc;
  // windows no pass can hold are never launched: their results say so
  PoaResult rejected{};
  rejected.status = rh::kPoaRejectedOnHost;
  for (uint32_t w : refused) {
    RH_HIP_CHECK(rh::copy_sync(s->d_results.as<PoaResult>() + w, &rejected, sizeof(PoaResult),
                               hipMemcpyHostToDevice, lane.stream),
                 RH_ERR_DEVICE);
  }
  uint32_t launches = 0;
  if (!passes.empty()) {
    std::vector<const std::vector<uint32_t>*> ptrs;
    for (const auto& o : orders) ptrs.push_back(&o);
    rc = rh::launch_passes(lane, s->tables(), ex->config, ptrs, passes, RH_POA_BATCHED,
                           &launches);
    if (rc != RH_OK) return rc;
    rc = rh::finish_passes(lane, kernel_ms);
    if (rc != RH_OK) return rc;
  } else {
    RH_HIP_CHECK(hipStreamSynchronize(lane.stream), RH_ERR_DEVICE);
  }
  return RH_OK;
}

struct Part {
  Request* r;
  size_t first;  // gpu_order position
  size_t count;
};

// Reads back the results of windows `ws` (ids, in the order of `parts`):
// one copy per part, of the id range its windows span.
int32_t read_results(rh_server* s, rh::PoaLane& lane, const std::vector<Part>& parts,
                     const std::vector<uint32_t>& ws, std::vector<PoaResult>* out) {
  out->resize(ws.size());
  std::vector<std::vector<PoaResult>> spans(parts.size());
  std::vector<uint32_t> lows(parts.size());
  size_t k = 0;
  for (size_t p = 0; p < parts.size(); ++p) {
    uint32_t lo = ws[k], hi = ws[k];
    for (size_t i = 0; i < parts[p].count; ++i) {
      lo = std::min(lo, ws[k + i]);
      hi = std::max(hi, ws[k + i]);
    }
    lows[p] = lo;
    spans[p].resize(hi - lo + 1);
    RH_HIP_CHECK(hipMemcpyAsync(spans[p].data(), s->d_results.as<PoaResult>() + lo,
                                spans[p].size() * sizeof(PoaResult), hipMemcpyDeviceToHost,
                                lane.stream),
                 RH_ERR_DEVICE);
    k += parts[p].count;
  }
  RH_HIP_CHECK(hipStreamSynchronize(lane.stream), RH_ERR_DEVICE);
  k = 0;
  for (size_t p = 0; p < parts.size(); ++p) {
    for (size_t i = 0; i < parts[p].count; ++i, ++k) (*out)[k] = spans[p][ws[k] - lows[p]];
  }
  return RH_OK;
}

// Reruns the windows whose graph outgrew the default slot, in slots sized
// from their exact bounds and with consensus space of their node bound.
int32_t retry(rh_server* s, Executor* ex, const std::vector<uint32_t>& ws,
              const std::vector<Request*>& owner, const std::vector<int32_t>& why,
              double* kernel_ms, uint64_t* retried) {
  rh::PoaLane& lane = ex->lane;
  std::vector<uint32_t> again;
  std::vector<int32_t> again_why;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    for (size_t k = 0; k < ws.size(); ++k) {
      PoaWindow& w = s->host_windows[ws[k]];
      const uint64_t bytes = rh::align_up(w.node_bound, kRegionAlign);
      uint64_t at = 0;
      if (!s->pool.alloc(bytes, &at)) continue;  // no room: stays rejected
      owner[k]->extra.emplace_back(at, bytes);
      w.out_off = at;
      w.out_cap = w.node_bound;
      again.push_back(ws[k]);
      again_why.push_back(why[k]);
    }
  }
  *retried += again.size();
  for (uint32_t id : again) {
    RH_HIP_CHECK(hipMemcpyAsync(s->d_windows.as<PoaWindow>() + id, &s->host_windows[id],
                                sizeof(PoaWindow), hipMemcpyHostToDevice, lane.stream),
                 RH_ERR_DEVICE);
  }
  if (again.empty()) return RH_OK;
  std::vector<std::vector<uint32_t>> orders;
  std::vector<rh::PassLaunch> passes;
  std::vector<uint32_t> refused;
  rh::PoaConfig c = ex->config;
  c.tune_slots = 0;  // exact slots are large: as many as fit
  c.band = 0;
  int32_t rc = rh::plan_retries(lane, s->host_windows.data(), again, again_why, c, &orders,
                                &passes, &refused);
  if (rc != RH_OK) return rc;
  if (!passes.empty()) {
    std::vector<const std::vector<uint32_t>*> ptrs;
    for (const auto& o : orders) ptrs.push_back(&o);
    uint32_t launches = 0;
    rc = rh::launch_passes(lane, s->tables(), c, ptrs, passes, RH_POA_BATCHED, &launches);
    if (rc != RH_OK) return rc;
    rc = rh::finish_passes(lane, kernel_ms);
    if (rc != RH_OK) return rc;
  }
  return RH_OK;
}

// Downloads a finished request's consensus space and retried windows.
int32_t download(rh_server* s, rh::PoaLane& lane, Request* r) {
  const uint64_t out_bytes = r->packed.out_bytes;
  r->out_host.resize(out_bytes);
  if (out_bytes > 0) {
    RH_HIP_CHECK(hipMemcpyAsync(r->out_host.data(),
                                s->d_pool.as<uint8_t>() + r->region_at + r->out_at, out_bytes,
                                hipMemcpyDeviceToHost, lane.stream),
                 RH_ERR_DEVICE);
  }
  for (uint32_t i = 0; i < r->n; ++i) {
    const PoaWindow& w = s->host_windows[r->win_at + i];
    const PoaResult& res = r->results[i];
    if (w.out_off != r->packed.windows[i].out_off &&
        (res.status == rh::kPoaPolished || res.status == rh::kPoaPolishedUntrimmed)) {
      std::vector<uint8_t>& dst = r->retried_out[i];
      dst.resize(res.length);
      if (res.length > 0) {
        RH_HIP_CHECK(hipMemcpyAsync(dst.data(), s->d_pool.as<uint8_t>() + w.out_off, res.length,
                                    hipMemcpyDeviceToHost, lane.stream),
                     RH_ERR_DEVICE);
      }
    }
  }
  RH_HIP_CHECK(hipStreamSynchronize(lane.stream), RH_ERR_DEVICE);
  return RH_OK;
}

void executor_loop(rh_server* s, Executor* ex) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  for (;;) {
    s->cv.wait(lock, [&] {
      return (!s->held && !s->ready.empty()) ||
             (s->closing && s->waiting.empty() && s->uploading == 0 && s->ready.empty());
    });
    if (s->ready.empty()) return;  // closing, drained
    std::vector<Part> parts;
    std::vector<uint32_t> ws;
    while (!s->ready.empty() && ws.size() < s->launch_windows) {
      Request* r = s->ready.front();
      const size_t total = r->packed.gpu_order.size();
      const size_t k = std::min<size_t>(total - r->next, s->launch_windows - ws.size());
      for (size_t i = r->next; i < r->next + k; ++i) {
        ws.push_back(static_cast<uint32_t>(r->win_at + r->packed.gpu_order[i]));
      }
      parts.push_back(Part{r, r->next, k});
      r->next += k;
      if (r->next == total) s->ready.pop_front();
    }
    s->stats.launches++;
    s->stats.windows_launched += ws.size();
    if (parts.size() > 1) s->stats.shared_launches++;
    lock.unlock();

    double kernel_ms = 0;
    uint64_t retried = 0;
    std::vector<PoaResult> res;
    int32_t rc = run_launch(s, ex, ws, &kernel_ms);
    if (rc == RH_OK) rc = read_results(s, ex->lane, parts, ws, &res);
    if (rc == RH_OK) {
      std::vector<uint32_t> again;
      std::vector<Request*> owner;
      std::vector<int32_t> why;
      size_t k = 0;
      for (const Part& p : parts) {
        for (size_t i = 0; i < p.count; ++i, ++k) {
          if (rh::is_retry_status(res[k].status)) {
            again.push_back(ws[k]);
            why.push_back(res[k].status);
            owner.push_back(p.r);
          }
        }
      }
      if (!again.empty()) {
        rc = retry(s, ex, again, owner, why, &kernel_ms, &retried);
        if (rc == RH_OK) rc = read_results(s, ex->lane, parts, ws, &res);
      }
    }
    const std::string text = rc == RH_OK ? std::string() : rh::last_error();

    std::vector<Request*> done;
    lock.lock();
    s->stats.kernel_ms += kernel_ms;
    s->stats.retried += retried;
    size_t k = 0;
    for (const Part& p : parts) {
      Request* r = p.r;
      r->launches++;
      if (parts.size() > 1) r->shared++;
      for (size_t i = 0; i < p.count; ++i, ++k) {
        const uint32_t local = r->packed.gpu_order[p.first + i];
        if (rc == RH_OK) r->results[local] = res[k];
      }
      if (rc != RH_OK && r->error == RH_OK) {
        r->error = rc;
        r->error_text = text;
      }
      r->left -= p.count;
      if (r->left == 0) done.push_back(r);
    }
    if (!done.empty()) {
      lock.unlock();
      std::vector<int32_t> drc(done.size(), RH_OK);
      std::vector<std::string> dtext(done.size());
      for (size_t d = 0; d < done.size(); ++d) {
        if (done[d]->error != RH_OK) continue;
        drc[d] = download(s, ex->lane, done[d]);
        if (drc[d] != RH_OK) dtext[d] = rh::last_error();
      }
      lock.lock();
      for (size_t d = 0; d < done.size(); ++d) {
        Request* r = done[d];
        if (drc[d] != RH_OK && r->error == RH_OK) {
          r->error = drc[d];
          r->error_text = "download: " + dtext[d];
        }
        for (uint32_t i = 0; i < r->n; ++i) {
          const int32_t st = r->results[i].status;
          if (!r->packed.shallow[i] && st != rh::kPoaPolished && st != rh::kPoaPolishedUntrimmed) {
            s->stats.rejected++;
          }
        }
        finish(s, r);
      }
    }
    s->cv.notify_all();
  }
}

Request* find(rh_server* s, uint64_t id) {
  auto it = s->requests.find(id);
  return it == s->requests.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int32_t rh_server_create(int32_t device, int8_t match, int8_t mismatch, int8_t gap,
                         int32_t trim, uint64_t budget_bytes, uint32_t executors,
                         uint32_t launch_windows, rh_server** server) {
  rh::clear_error();
  if (server == nullptr) {
    rh::set_error("rh_server_create: null server pointer");
    return RH_ERR_ARGUMENT;
  }
  *server = nullptr;
  if (budget_bytes < (1u << 20) || executors == 0 || executors > 16 || launch_windows == 0) {
    rh::set_error("rh_server_create: budget below 1 MiB, or executors not in 1..16, or no "
                  "windows per launch");
    return RH_ERR_ARGUMENT;
  }
  int n = 0;
  RH_HIP_CHECK(hipGetDeviceCount(&n), RH_ERR_DEVICE);
  if (device < 0 || device >= n) {
    rh::set_error("rh_server_create: no such device");
    return RH_ERR_DEVICE;
  }
  RH_HIP_CHECK(hipSetDevice(device), RH_ERR_DEVICE);
  auto s = std::make_unique<rh_server>();
  s->device = device;
  s->trim = trim != 0;
  s->config.match = match;
  s->config.mismatch = mismatch;
  s->config.gap = gap;
  s->launch_windows = launch_windows;
  // 1/32 of the budget for window and result records, the rest for regions
  s->window_capacity = std::max<uint64_t>(1, (budget_bytes / 32) / record_bytes());
  const uint64_t pool_bytes = budget_bytes - s->window_capacity * record_bytes();
  auto fail = [&](const char* what) {
    rh::set_error(std::string("rh_server_create: ") + what);
    s->d_pool.release();
    s->d_windows.release();
    s->d_results.release();
    if (s->upload_stream != nullptr) (void)hipStreamDestroy(s->upload_stream);
    for (auto& e : s->executors) e->lane.release();
    return RH_ERR_MEMORY;
  };
  if (s->d_pool.reserve(pool_bytes) != hipSuccess ||
      s->d_windows.reserve(s->window_capacity * sizeof(PoaWindow)) != hipSuccess ||
      s->d_results.reserve(s->window_capacity * sizeof(PoaResult)) != hipSuccess) {
    return fail("device pool allocation failed");
  }
  if (hipStreamCreate(&s->upload_stream) != hipSuccess) return fail("stream creation failed");
  s->pool.reset(pool_bytes);
  s->ids.reset(s->window_capacity);
  s->host_windows.assign(s->window_capacity, PoaWindow{});
  s->stats.pool_bytes = pool_bytes;
  s->stats.window_capacity = s->window_capacity;
  for (uint32_t e = 0; e < executors; ++e) {
    auto ex = std::make_unique<Executor>();
    ex->config = s->config;
    ex->config.slot_share = executors;  // the executors split the resident waves
    if (ex->lane.init(device) != hipSuccess) {
      ex->lane.release();
      return fail("stream/event creation failed");
    }
    s->executors.push_back(std::move(ex));
  }
  rh_server* raw = s.release();
  raw->admitter = std::thread(admitter_loop, raw);
  for (auto& ex : raw->executors) ex->thread = std::thread(executor_loop, raw, ex.get());
  *server = raw;
  return RH_OK;
}

int32_t rh_server_warm(rh_server* s, uint32_t n, const rh_window* windows) {
  rh::clear_error();
  if (s == nullptr || (n > 0 && windows == nullptr)) {
    rh::set_error("rh_server_warm: null argument");
    return RH_ERR_ARGUMENT;
  }
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->started) {
      rh::set_error("rh_server_warm: requests were already submitted");
      return RH_ERR_STATE;
    }
  }
  RH_HIP_CHECK(hipSetDevice(s->device), RH_ERR_DEVICE);
  rh::PackedWindows P;
  int32_t rc = rh::pack_windows(n, windows, s->trim, &P);
  if (rc != RH_OK) {
    rh::set_error(std::string("rh_server_warm: ") + rh::last_error());
    return rc;
  }
  if (P.gpu_order.empty()) return RH_OK;
  rh::DeviceBuffer arena, win, lay, out, res;
  auto cleanup = [&] {
    arena.release();
    win.release();
    lay.release();
    out.release();
    res.release();
  };
  if (arena.reserve(std::max<uint64_t>(P.arena.size(), 1)) != hipSuccess ||
      win.reserve(n * sizeof(PoaWindow)) != hipSuccess ||
      lay.reserve(std::max<uint64_t>(P.layers.size(), 1) * sizeof(PoaLayer)) != hipSuccess ||
      out.reserve(std::max<uint64_t>(P.out_bytes, 1)) != hipSuccess ||
      res.reserve(n * sizeof(PoaResult)) != hipSuccess) {
    cleanup();
    rh::set_error("rh_server_warm: device allocation failed");
    return RH_ERR_MEMORY;
  }
  rh::PoaTables t;
  t.host = P.windows.data();
  t.windows = win.as<PoaWindow>();
  t.layers = lay.as<PoaLayer>();
  t.arena = arena.as<uint8_t>();
  t.out = out.as<uint8_t>();
  t.results = res.as<PoaResult>();
  for (auto& ex : s->executors) {
    rh::PoaLane& lane = ex->lane;
    if (hipMemcpyAsync(arena.ptr, P.arena.data(), P.arena.size(), hipMemcpyHostToDevice,
                       lane.stream) != hipSuccess ||
        hipMemcpyAsync(win.ptr, P.windows.data(), n * sizeof(PoaWindow), hipMemcpyHostToDevice,
                       lane.stream) != hipSuccess ||
        (!P.layers.empty() &&
         hipMemcpyAsync(lay.ptr, P.layers.data(), P.layers.size() * sizeof(PoaLayer),
                        hipMemcpyHostToDevice, lane.stream) != hipSuccess)) {
      cleanup();
      rh::set_error("rh_server_warm: copy failed");
      return RH_ERR_DEVICE;
    }
    std::vector<std::vector<uint32_t>> orders;
    std::vector<rh::PassLaunch> passes;
    std::vector<uint32_t> refused;
    rc = rh::plan_bins(lane, P.windows.data(), P.gpu_order, 4096, RH_POA_BATCHED, ex->config,
                       &orders, &passes, &refused);
    if (rc == RH_OK && !passes.empty()) {
      std::vector<const std::vector<uint32_t>*> ptrs;
      for (const auto& o : orders) ptrs.push_back(&o);
      uint32_t launches = 0;
      double ms = 0;
      rc = rh::launch_passes(lane, t, ex->config, ptrs, passes, RH_POA_BATCHED, &launches);
      if (rc == RH_OK) rc = rh::finish_passes(lane, &ms);
    }
    if (rc != RH_OK) {
      const std::string text = rh::last_error();
      cleanup();
      rh::set_error("rh_server_warm: " + text);
      return rc;
    }
  }
  cleanup();
  return RH_OK;
}

int32_t rh_server_submit(rh_server* s, uint32_t n, const rh_window* windows,
                         uint64_t* request) {
  rh::clear_error();
  if (s == nullptr || request == nullptr || (n > 0 && windows == nullptr)) {
    rh::set_error("rh_server_submit: null argument");
    return RH_ERR_ARGUMENT;
  }
  const double arrival = s->now();
  auto r = std::make_unique<Request>();
  r->n = n;
  int32_t rc = rh::pack_windows(n, windows, s->trim, &r->packed);
  if (rc != RH_OK) {
    rh::set_error(std::string("rh_server_submit: ") + rh::last_error());
    return rc;
  }
  layout_region(r.get());
  if (r->region_bytes > s->pool.capacity() || n > s->window_capacity) {
    rh::set_error("rh_server_submit: request of " + std::to_string(r->device_bytes) +
                  " bytes (" + std::to_string(n) + " windows) exceeds the server's budget");
    return RH_ERR_ARGUMENT;
  }
  r->results.assign(n, PoaResult{});
  for (uint32_t i = 0; i < n; ++i) {
    r->results[i].status = r->packed.host_rejected[i] ? rh::kPoaRejectedOnHost : rh::kPoaNotRun;
  }
  r->left = r->packed.gpu_order.size();
  r->arrival = arrival;
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->closing) {
    rh::set_error("rh_server_submit: the server is closing");
    return RH_ERR_STATE;
  }
  r->id = s->next_id++;
  *request = r->id;
  s->started = true;
  s->stats.submitted++;
  s->waiting.push_back(r.get());
  s->requests.emplace(r->id, std::move(r));
  s->cv.notify_all();
  return RH_OK;
}

int32_t rh_server_hold(rh_server* s, int32_t hold) {
  rh::clear_error();
  if (s == nullptr) {
    rh::set_error("rh_server_hold: null server");
    return RH_ERR_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(s->mu);
  s->held = hold != 0;
  s->cv.notify_all();
  return RH_OK;
}

int32_t rh_server_wait(rh_server* s, uint64_t id, int32_t timeout_ms) {
  rh::clear_error();
  if (s == nullptr) {
    rh::set_error("rh_server_wait: null server");
    return RH_ERR_ARGUMENT;
  }
  std::unique_lock<std::mutex> lock(s->mu);
  Request* r = find(s, id);
  if (r == nullptr) {
    rh::set_error("rh_server_wait: unknown request");
    return RH_ERR_ARGUMENT;
  }
  auto is_done = [&] { return r->state == kDone; };
  if (timeout_ms < 0) {
    s->cv.wait(lock, is_done);
  } else if (!s->cv.wait_for(lock, std::chrono::milliseconds(timeout_ms), is_done)) {
    return RH_TIMEOUT;
  }
  if (r->error != RH_OK) {
    rh::set_error("request " + std::to_string(id) + ": " + r->error_text);
    return r->error;
  }
  return RH_OK;
}

int32_t rh_server_request_info(rh_server* s, uint64_t id, rh_request_info* info) {
  rh::clear_error();
  if (s == nullptr || info == nullptr) {
    rh::set_error("rh_server_request_info: null argument");
    return RH_ERR_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(s->mu);
  Request* r = find(s, id);
  if (r == nullptr) {
    rh::set_error("rh_server_request_info: unknown request");
    return RH_ERR_ARGUMENT;
  }
  *info = rh_request_info{};
  info->state = r->state;
  info->windows = r->n;
  info->device_bytes = r->device_bytes;
  uint64_t bound = 0;
  for (uint32_t i = 0; i < r->n; ++i) {
    bound += std::max<uint64_t>(r->packed.windows[i].backbone_len, r->packed.windows[i].node_bound);
  }
  info->output_bound = bound;
  info->launches = r->launches;
  info->shared_launches = r->shared;
  info->arrival_s = r->arrival;
  info->admitted_s = r->state == kWaiting ? 0.0 : r->admitted;
  info->done_s = r->state == kDone ? r->done : 0.0;
  return RH_OK;
}

int32_t rh_server_fetch(rh_server* s, uint64_t id, char* out, uint64_t capacity,
                        uint64_t* offsets, uint32_t* lengths, uint8_t* status) {
  rh::clear_error();
  if (s == nullptr || (out == nullptr && capacity > 0)) {
    rh::set_error("rh_server_fetch: null argument");
    return RH_ERR_ARGUMENT;
  }
  std::unique_ptr<Request> owned;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    Request* r = find(s, id);
    if (r == nullptr) {
      rh::set_error("rh_server_fetch: unknown request");
      return RH_ERR_ARGUMENT;
    }
    if (r->state != kDone) {
      rh::set_error("rh_server_fetch: request not done");
      return RH_ERR_STATE;
    }
    if (r->error != RH_OK) {
      rh::set_error("rh_server_fetch: request failed: " + r->error_text);
      return r->error;
    }
    if (r->n > 0 && (offsets == nullptr || lengths == nullptr || status == nullptr)) {
      rh::set_error("rh_server_fetch: null argument");
      return RH_ERR_ARGUMENT;
    }
    // check the capacity before forgetting the request
    uint64_t need = 0;
    for (uint32_t i = 0; i < r->n; ++i) {
      const int32_t st = r->results[i].status;
      const bool polished = !r->packed.shallow[i] &&
                            (st == rh::kPoaPolished || st == rh::kPoaPolishedUntrimmed);
      need += polished ? r->results[i].length : r->packed.windows[i].backbone_len;
    }
    if (need > capacity) {
      rh::set_error("rh_server_fetch: output capacity too small");
      return RH_ERR_ARGUMENT;
    }
    auto it = s->requests.find(id);
    owned = std::move(it->second);
    s->requests.erase(it);
  }
  Request* r = owned.get();
  const rh::PackedWindows& P = r->packed;
  uint64_t pos = 0;
  for (uint32_t i = 0; i < r->n; ++i) {
    const uint8_t* src;
    uint32_t len;
    uint8_t st;
    const int32_t rs = r->results[i].status;
    // the backbone bytes sit at their (rebased) arena offset minus the region base
    const uint64_t bb = P.windows[i].backbone_off - r->region_at;
    if (P.shallow[i]) {
      src = P.arena.data() + bb;
      len = P.windows[i].backbone_len;
      st = RH_WINDOW_UNPOLISHED;
    } else if (rs != rh::kPoaPolished && rs != rh::kPoaPolishedUntrimmed) {
      src = P.arena.data() + bb;
      len = P.windows[i].backbone_len;
      st = RH_WINDOW_REJECTED;
    } else {
      auto it = r->retried_out.find(i);
      if (it != r->retried_out.end()) {
        src = it->second.data();
      } else {
        src = r->out_host.data() + (P.windows[i].out_off - r->region_at - r->out_at);
      }
      len = r->results[i].length;
      st = rs == rh::kPoaPolishedUntrimmed ? RH_WINDOW_POLISHED_UNTRIMMED : RH_WINDOW_POLISHED;
    }
    if (len > 0) std::memcpy(out + pos, src, len);
    offsets[i] = pos;
    lengths[i] = len;
    status[i] = st;
    pos += len;
  }
  return RH_OK;
}

int32_t rh_server_get_stats(rh_server* s, rh_server_stats* stats) {
  rh::clear_error();
  if (s == nullptr || stats == nullptr) {
    rh::set_error("rh_server_get_stats: null argument");
    return RH_ERR_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(s->mu);
  *stats = s->stats;
  stats->pool_in_use = s->pool.used() + s->ids.used() * record_bytes();
  uint64_t scratch = 0;
  for (const auto& ex : s->executors) scratch += ex->lane.device_bytes();
  stats->scratch_bytes = scratch;
  return RH_OK;
}

void rh_server_destroy(rh_server* s) {
  if (s == nullptr) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->closing = true;
    s->held = false;
    s->cv.notify_all();
  }
  if (s->admitter.joinable()) s->admitter.join();
  for (auto& ex : s->executors) {
    if (ex->thread.joinable()) ex->thread.join();
    ex->lane.release();
  }
  (void)hipSetDevice(s->device);
  if (s->upload_stream != nullptr) (void)hipStreamDestroy(s->upload_stream);
  s->d_pool.release();
  s->d_windows.release();
  s->d_results.release();
  delete s;
}

}  // extern "C"
