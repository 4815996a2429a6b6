// C ABI: allocation-free overlap alignment (racon_hip.h, rh_align_run_noalloc).
//
// The same kernels as rh_align_run, for callers that cannot allocate device
// memory: every input and output is caller-owned device memory, and the
// kernels' scratch (the work-queue counter and the per-wave slots) lives in
// the file-scope device array `scratch` below. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in symbol): the
// device address of each device's copy is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "align_types.hpp"
#include "rh_common.hpp"

namespace rh {
hipError_t launch_align_batch_cleared(const AlignParams& params, uint32_t grid,
                                      hipStream_t stream);
hipError_t align_occupancy(int* blocks_per_cu);
}  // namespace rh

// The scratch of every rh_align_run_noalloc call on a device. Held by each
// device the library runs on until the process ends; calls that overlap on
// one device share it and race.
static __device__ __align__(256) uint8_t scratch[RH_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device `device` (looked up on
// the first call on that device).
int32_t scratch_address(int32_t device, uint8_t** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    RH_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), RH_ERR_DEVICE);
    if (ptr == nullptr) {
      rh::set_error("rh_align_run_noalloc: no device address for the scratch array");
      return RH_ERR_DEVICE;
    }
    g_scratch_ptr[device] = ptr;
  }
  *out = static_cast<uint8_t*>(g_scratch_ptr[device]);
  return RH_OK;
}

}  // namespace

extern "C" int32_t rh_align_run_noalloc(int32_t device, void* stream, uint32_t n,
                                        const rh_dev_pair* pairs, const uint8_t* arena,
                                        uint32_t max_query_length, uint32_t max_target_length,
                                        uint8_t* ops, rh_dev_align_result* results) {
  rh::clear_error();
  static_assert(sizeof(rh_dev_pair) == sizeof(rh::AlignPair), "rh_dev_pair layout");
  static_assert(sizeof(rh_dev_align_result) == sizeof(rh::AlignResult), "result layout");
  if (n == 0) {
    return RH_OK;
  }
  if (pairs == nullptr || arena == nullptr || ops == nullptr || results == nullptr) {
    rh::set_error("rh_align_run_noalloc: null device pointer");
    return RH_ERR_ARGUMENT;
  }
  int count = 0;
  RH_HIP_CHECK(hipGetDeviceCount(&count), RH_ERR_DEVICE);
  if (device < 0 || device >= count || device >= kMaxDevices) {
    rh::set_error("rh_align_run_noalloc: no such device");
    return RH_ERR_DEVICE;
  }
  RH_HIP_CHECK(hipSetDevice(device), RH_ERR_DEVICE);
  uint8_t* base = nullptr;
  const int32_t rc = scratch_address(device, &base);
  if (rc != RH_OK) {
    return rc;
  }

  // scratch: the queue counter, then as many slots as fit (at most one per
  // resident wave and one per pair)
  const rh::AlignSlotCaps caps = rh::align_caps_for(max_query_length, max_target_length);
  const rh::AlignSlotLayout layout = rh::align_slot_layout(caps);
  const uint64_t slot_bytes = rh::align_up_a(layout.total, 256);
  const uint64_t head_bytes = 256;
  const uint64_t room = RH_NOALLOC_SCRATCH_BYTES - head_bytes;
  if (slot_bytes > room) {
    rh::set_error("rh_align_run_noalloc: pairs of " + std::to_string(max_query_length) + " x " +
                  std::to_string(max_target_length) + " bytes need a " +
                  std::to_string(slot_bytes) + "-byte slot, more than the " +
                  std::to_string(room) + " bytes of scratch");
    return RH_ERR_ARGUMENT;
  }
  int per_cu = 0, cus = 0;
  RH_HIP_CHECK(rh::align_occupancy(&per_cu), RH_ERR_DEVICE);
  RH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
               RH_ERR_DEVICE);
  uint64_t slots = room / slot_bytes;
  slots = std::min<uint64_t>(slots, static_cast<uint64_t>(std::max(1, per_cu)) * std::max(1, cus));
  slots = std::min<uint64_t>(slots, n);

  rh::AlignParams p{};
  p.pairs = reinterpret_cast<const rh::AlignPair*>(pairs);
  p.arena = arena;
  p.ops = ops;
  p.results = reinterpret_cast<rh::AlignResult*>(results);
  p.order = nullptr;  // queue position = pair index
  p.n_order = n;
  p.queue_head = reinterpret_cast<uint32_t*>(base);
  p.scratch = base + head_bytes;
  p.slot_bytes = slot_bytes;
  p.caps = caps;
  p.layout = layout;
  RH_HIP_CHECK(rh::launch_align_batch_cleared(p, static_cast<uint32_t>(slots),
                                              static_cast<hipStream_t>(stream)),
               RH_ERR_KERNEL);
  return RH_OK;
}
