// Host side of the POA kernels shared by the batch entry points (rh_poa_*)
// and the serving loop (rh_server_*): packing windows the way
// Window::generate_consensus orders them (src/window.cpp:65-87), planning
// and sizing kernel passes, and launching them on one HIP stream.
#pragma once

#include <vector>

#include "poa_types.hpp"
#include "rh_common.hpp"

namespace rh {

// Windows packed for the device: the input arena (bytes and qualities),
// window and layer records (arena offsets relative to the arena start,
// layer_first relative to the layer records, out_off relative to an output
// region of out_bytes) and, per window, whether it needs the kernel.
struct PackedWindows {
  HostArena arena;                      // the input bytes, unless staged (below)
  // With a pinned staging buffer the bytes go there instead: the copies are
  // left to the upload (copy_and_upload), which overlaps them with the copy
  // to the device. host_arena points at the bytes either way.
  std::vector<ArenaCopy> staged_copies;
  const uint8_t* host_arena = nullptr;
  uint64_t arena_bytes = 0;
  std::vector<PoaWindow> windows;
  std::vector<PoaLayer> layers;
  std::vector<uint8_t> shallow;         // 1: fewer than 3 sequences (backbone kept)
  std::vector<uint8_t> host_rejected;   // 1: outside the device limits, never launched
  std::vector<uint32_t> gpu_order;      // windows for the kernel, largest cost first
  uint64_t out_bytes = 0;
};

// Packs n windows (racon's std::sort of layers by begin, window.cpp:78-85;
// layer checks of Window::add_layer, window.cpp:45-63). Fails with
// RH_ERR_ARGUMENT on a malformed window or layer.
int32_t pack_windows(uint32_t n, const rh_window* windows, bool trim, PackedWindows* out,
                     PinnedBuffer* staging = nullptr);

// Estimated cost of a window: graph bound x (longest layer + 1).
uint64_t window_cost(const PoaWindow& w);

// Graph nodes a main pass sizes a window's slot for: the backbone plus the
// larger of twice the longest layer and a quarter of all layer bytes (the
// node bound counts every layer byte as a new node; racon's sample reads and
// a 5 Mbp 50x ONT contig grow at most 0.14 and 0.08 nodes per layer byte).
// A graph that outgrows it ends kPoaNodeOverflow and reruns in exact slots.
uint64_t graph_estimate(const PoaWindow& w);

// Statuses after which a window reruns in a slot sized from its exact bounds.
bool is_retry_status(int32_t s);

// Layers up to this length run in the main pass; longer ones run in a pass
// of their own with a wider row pitch (and a shallower LDS ring), so that one
// long layer does not cut the LDS ring and the occupancy of every window.
constexpr uint32_t kMainPassMaxLayer = 3072;

// Scores and tuning of the passes (0: library default). A lane that runs
// its kernels beside slot_share - 1 other lanes takes that share of the
// device's resident waves.
struct PoaConfig {
  int32_t match = 3, mismatch = -5, gap = -4;
  uint32_t tune_slots = 0, tune_rows = 0, tune_ring = 0;
  uint32_t slot_share = 1;
  uint32_t bin_ring[2] = {0, 0};  // per layer-length bin ring rows (autotuned; 0: tune_ring)
  uint32_t band = 0;              // racon -b: band rows (256 columns) for long layers
};

// Layer-length bin of a window: 0 main pass, 1 long-layer pass.
inline int layer_bin(const PoaWindow& w) { return w.max_layer_len <= kMainPassMaxLayer ? 0 : 1; }

// One HIP stream with the device buffers its passes use: per-wave scratch
// slots, the queue orders, queue counters and parameter blocks.
struct PoaLane {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DeviceBuffer d_scratch, d_order, d_queue, d_params;
  uint32_t lds_limit = 0;  // dynamic LDS a workgroup may use
  hipError_t init(int32_t dev);
  void release();
  uint64_t device_bytes() const {
    return d_scratch.bytes + d_order.bytes + d_queue.bytes + d_params.bytes;
  }
};

// Device tables the kernel reads and writes (windows indexed by window id)
// and the host copy of the window records, used for planning.
struct PoaTables {
  const PoaWindow* host = nullptr;
  const PoaWindow* windows = nullptr;
  const PoaLayer* layers = nullptr;
  const uint8_t* arena = nullptr;
  uint8_t* out = nullptr;
  PoaResult* results = nullptr;
};

// One kernel pass: slot capacities and the DP rows kept in LDS per wave.
struct PassPlan {
  PoaSlotCaps caps;
  uint32_t ring;  // LDS ring rows (0: every DP row in the slot's global matrix)
  uint32_t lds;   // dynamic LDS bytes per wave
  bool wide;      // the wide-aligned-set kernel (kMaxAlignedWide)
};

// A pass ready to launch: its plan, resident slots and slot size.
struct PassLaunch {
  PassPlan plan;
  uint64_t slots = 0;
  uint64_t slot_bytes = 0;
  PoaSlotLayout layout{};
};

// The plan for windows `ws`: graph arrays for the largest (by the graph
// estimate, or by the exact node bound when row_cap is 0), the row pitch of
// the longest layer, DP rows capped at row_cap (0: exact, node bound + 1)
// with a pool of kPoaHPoolRows global score rows (0: one per DP row),
// and the deepest ring that fits lds_limit, for the default or the wide
// kernel. False when even a ring of 0 rows does not fit (the layer bytes
// alone exceed the limit).
bool plan_pass(const PoaWindow* host, const std::vector<uint32_t>& ws, uint32_t row_cap,
               uint32_t lds_limit, uint32_t tune_ring, bool wide, PassPlan* plan);

// Resident slots of a pass: the kernel's occupancy times the CUs divided by
// the lane's share (all windows for RH_POA_STATIC, one for RH_POA_PLAIN),
// within ~60% of free device memory plus the scratch the lane already holds.
int32_t size_pass(PoaLane& lane, uint64_t n_windows, const PassPlan& plan, int32_t mode,
                  const PoaConfig& c, PassLaunch* out);

// Queues passes back to back on the lane's stream between its events
// (orders, queue counters and parameter blocks side by side; the passes
// share the scratch, sized for the largest) and returns without waiting.
int32_t launch_passes(PoaLane& lane, const PoaTables& t, const PoaConfig& c,
                      const std::vector<const std::vector<uint32_t>*>& orders,
                      const std::vector<PassLaunch>& passes, int32_t mode, uint32_t* launches);

// Waits for the passes of launch_passes and adds their device time.
int32_t finish_passes(PoaLane& lane, double* kernel_ms);

// Splits windows into the main and the long-layer bin, plans and sizes a pass
// per non-empty bin (row_cap 0: exact rows; wide: the wide kernel). Windows
// of a bin whose plan does not fit the LDS limit are appended to `refused`.
int32_t plan_bins(PoaLane& lane, const PoaWindow* host, const std::vector<uint32_t>& ws,
                  uint32_t row_cap, int32_t mode, const PoaConfig& c,
                  std::vector<std::vector<uint32_t>>* orders, std::vector<PassLaunch>* passes,
                  std::vector<uint32_t>* refused, bool wide = false);

// Plans the rerun of `ws` (windows that ended in the retry statuses
// `status`) in exact slots: windows whose aligned sets were full in the wide
// kernel, the others in the default one. `c` is the rerun's configuration
// (the exact DP: band 0). Appends the orders and passes.
int32_t plan_retries(PoaLane& lane, const PoaWindow* host, const std::vector<uint32_t>& ws,
                     const std::vector<int32_t>& status, const PoaConfig& c,
                     std::vector<std::vector<uint32_t>>* orders,
                     std::vector<PassLaunch>* passes, std::vector<uint32_t>* refused);

}  // namespace rh
