// Layer extraction on the device: the loop of racon's Polisher::initialize
// that turns each overlap's breaking points into window layers
// (src/polisher.cpp:403-457), over the points an aligner batch leaves in
// device memory.
//
// A set holds one record slot per breaking-point pair the batches can give
// (the rooms of rh_align_breaking_points). rh_layers_add fills the slots of
// one batch: one wave per aligned pair, its pairs (windows) one after the
// other, the wave summing a pair's quality bytes 64 at a time. A slot gets
// the layer's window as its key, or the "none" key (n_windows) when the pair
// fails a filter or the room is unused. rh_layers_finish sorts the slots by
// key with a stable radix sort (slot order is overlap order, so each window
// keeps racon's layer order) and gathers the kept records in that order.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "align_types.hpp"
#include "rh_common.hpp"

struct rh_layer_set {
  int32_t device = 0;
  uint32_t window_length = 0;
  double quality_threshold = 0;
  uint32_t none_key = 0;  // n_windows: the key of a slot without a layer
  uint64_t capacity = 0;  // slots
  uint64_t used = 0;      // slots filled by rh_layers_add
  uint64_t kept = 0;      // layers, after rh_layers_finish
  uint64_t quality_size = 0;
  bool finished = false;
  rh::DeviceBuffer d_quality, d_sources, d_keys, d_keys_sorted, d_slots, d_slots_sorted,
      d_records, d_grouped, d_kept, d_temp;
  hipStream_t stream = nullptr;
};

namespace rh {
namespace {

static_assert(sizeof(rh_layer_record) == 24, "layer record layout");
static_assert(sizeof(rh_layer_source) == 32, "layer source layout");

constexpr int kLayerWave = 64;
constexpr int kLayerWavesPerBlock = 4;

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int d = kLayerWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kLayerWave);
  return x;
}

// One wave per aligned pair of the batch. Pair i's points are the
// (t, q) values points[2 * (out + k)], out = spans[i].out, k < counts[i]:
// breaking-point pair e is points k = 2e (first match) and 2e + 1 (after the
// last match). Its slots are slot_base + out / 2 + e for e < spans[i].cap / 2.
//
// The filters of src/polisher.cpp:411-429: the query span q1 - q0 must reach
// 0.02 * window_length, and when the read has quality bytes their mean of
// (byte - 33) over the span must reach the threshold. racon adds the terms
// as uint32 values (the char sign-extended, minus 33) to a double one by
// one; the wave adds them as uint64 and converts once, which is the same
// double while the sum stays below 2^53 (every quality string of printable
// bytes, and any span under 2^21 bytes).
__global__ __launch_bounds__(kLayerWave * kLayerWavesPerBlock) void layer_select_kernel(
    const BreakSpan* __restrict__ spans, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ points, const rh_layer_source* __restrict__ sources, uint32_t n,
    uint32_t window_length, double min_span, double threshold,
    const uint8_t* __restrict__ quality, uint64_t quality_size, uint64_t slot_base,
    uint32_t none_key, uint32_t* __restrict__ keys, uint32_t* __restrict__ slots,
    rh_layer_record* __restrict__ records, unsigned long long* __restrict__ kept) {
  const uint32_t i = blockIdx.x * kLayerWavesPerBlock + threadIdx.x / kLayerWave;
  const int lane = threadIdx.x & (kLayerWave - 1);
  if (i >= n) return;  // (uniform over the wave)
  const BreakSpan sp = spans[i];
  const rh_layer_source src = sources[i];
  const uint32_t count = counts[i];
  const uint32_t room = sp.cap / 2;
  const uint32_t pairs = count >= RH_PAIR_REJECTED ? 0 : min(count / 2, room);
  const uint64_t base = slot_base + sp.out / 2;
  const uint32_t* pt = points + 2 * sp.out;
  // unused room
  for (uint32_t e = pairs + lane; e < room; e += kLayerWave) {
    keys[base + e] = none_key;
    slots[base + e] = static_cast<uint32_t>(base + e);
  }
  const bool has_quality = src.quality != RH_NO_QUALITY;
  uint32_t n_kept = 0;
  for (uint32_t e = 0; e < pairs; ++e) {
    const uint32_t t0 = pt[4 * e], q0 = pt[4 * e + 1];
    const uint32_t t1 = pt[4 * e + 2], q1 = pt[4 * e + 3];
    const uint32_t span = q1 - q0;
    bool keep = !(static_cast<double>(span) < min_span);
    if (keep && has_quality) {
      // the span in the stored bytes: mirrored bytes run from the read's end
      uint64_t a = src.mirrored ? static_cast<uint64_t>(src.query_length) - q1 : q0;
      uint64_t b = src.mirrored ? static_cast<uint64_t>(src.query_length) - q0 : q1;
      // (a span outside the read, which well-formed inputs never give, reads
      // nothing past the stored bytes)
      const uint64_t limit = quality_size - src.quality;
      a = min(a, limit);
      b = min(max(a, b), limit);
      const uint8_t* qb = quality + src.quality;
      uint64_t sum = 0;
      for (uint64_t k = a + lane; k < b; k += kLayerWave) {
        const uint32_t c = static_cast<uint32_t>(static_cast<int32_t>(static_cast<int8_t>(qb[k])));
        sum += static_cast<uint64_t>(static_cast<uint32_t>(c - 33u));
      }
      sum = wave_sum_u64(sum);
      const double average = static_cast<double>(sum) / static_cast<double>(span);
      keep = !(average < threshold);
    }
    const uint32_t w = t0 / window_length;
    const uint64_t window = src.first_window + w;
    if (keep && window >= none_key) {
      keep = false;  // not a window of the set: counted, rh_layers_finish fails
      if (lane == 0) atomicAdd(kept + 1, 1ull);
    }
    if (lane == 0) {
      const uint64_t slot = base + e;
      slots[slot] = static_cast<uint32_t>(slot);
      if (keep) {
        const uint32_t start = w * window_length;
        keys[slot] = static_cast<uint32_t>(window);
        records[slot] = rh_layer_record{static_cast<uint32_t>(window), src.overlap, q0, q1,
                                        t0 - start, t1 - start - 1};
      } else {
        keys[slot] = none_key;
      }
    }
    n_kept += keep ? 1u : 0u;
  }
  if (lane == 0 && n_kept > 0) atomicAdd(kept, static_cast<unsigned long long>(n_kept));
}

__global__ void layer_gather_kernel(const uint32_t* __restrict__ order,
                                    const rh_layer_record* __restrict__ records, uint64_t n,
                                    rh_layer_record* __restrict__ out) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) out[k] = records[order[k]];
}

int key_bits(uint32_t none_key) {
  int bits = 1;
  while (bits < 32 && (none_key >> bits) != 0) ++bits;
  return bits;
}

}  // namespace
}  // namespace rh

extern "C" {

int32_t rh_layers_create(int32_t device, uint32_t window_length, double quality_threshold,
                         uint64_t n_windows, uint64_t capacity, const char* qualities,
                         uint64_t qualities_size, rh_layer_set** set) {
  rh::clear_error();
  if (set == nullptr) {
    rh::set_error("rh_layers_create: null set pointer");
    return RH_ERR_ARGUMENT;
  }
  *set = nullptr;
  if (window_length == 0 || n_windows >= 0xffffffffull || capacity >= 0xffffffffull ||
      (qualities == nullptr && qualities_size != 0)) {
    rh::set_error(
        "rh_layers_create: zero window length, too many windows or slots (< 2^32 - 1), or "
        "null qualities");
    return RH_ERR_ARGUMENT;
  }
  int n = 0;
  RH_HIP_CHECK(hipGetDeviceCount(&n), RH_ERR_DEVICE);
  if (device < 0 || device >= n) {
    rh::set_error("rh_layers_create: no such device");
    return RH_ERR_DEVICE;
  }
  RH_HIP_CHECK(hipSetDevice(device), RH_ERR_DEVICE);
  std::unique_ptr<rh_layer_set> s(new rh_layer_set());
  s->device = device;
  s->window_length = window_length;
  s->quality_threshold = quality_threshold;
  s->none_key = static_cast<uint32_t>(n_windows);
  s->capacity = capacity;
  if (hipStreamCreate(&s->stream) != hipSuccess) {
    rh::set_error("rh_layers_create: stream creation failed");
    return RH_ERR_DEVICE;
  }
  const uint64_t slots = std::max<uint64_t>(capacity, 1);
  RH_HIP_CHECK(s->d_quality.reserve(std::max<uint64_t>(qualities_size, 1)), RH_ERR_MEMORY);
  RH_HIP_CHECK(s->d_keys.reserve(4 * slots), RH_ERR_MEMORY);
  RH_HIP_CHECK(s->d_keys_sorted.reserve(4 * slots), RH_ERR_MEMORY);
  RH_HIP_CHECK(s->d_slots.reserve(4 * slots), RH_ERR_MEMORY);
  RH_HIP_CHECK(s->d_slots_sorted.reserve(4 * slots), RH_ERR_MEMORY);
  RH_HIP_CHECK(s->d_records.reserve(sizeof(rh_layer_record) * slots), RH_ERR_MEMORY);
  RH_HIP_CHECK(s->d_kept.reserve(2 * sizeof(unsigned long long)), RH_ERR_MEMORY);
  if (qualities_size > 0) {
    RH_HIP_CHECK(hipMemcpyAsync(s->d_quality.ptr, qualities, qualities_size,
                                hipMemcpyHostToDevice, s->stream),
                 RH_ERR_DEVICE);
  }
  const unsigned long long zero[2] = {0, 0};
  RH_HIP_CHECK(hipMemcpyAsync(s->d_kept.ptr, zero, sizeof(zero), hipMemcpyHostToDevice,
                              s->stream),
               RH_ERR_DEVICE);
  RH_HIP_CHECK(hipStreamSynchronize(s->stream), RH_ERR_DEVICE);
  s->quality_size = qualities_size;
  *set = s.release();
  return RH_OK;
}

int32_t rh_layers_add(rh_layer_set* s, rh_align_batch* batch, const rh_layer_source* sources,
                      uint32_t n) {
  rh::clear_error();
  if (s == nullptr || batch == nullptr || (sources == nullptr && n != 0)) {
    rh::set_error("rh_layers_add: null argument");
    return RH_ERR_ARGUMENT;
  }
  if (s->finished) {
    rh::set_error("rh_layers_add: set already finished");
    return RH_ERR_STATE;
  }
  rh::BreakView v{};
  if (!rh::break_view(batch, &v)) {
    rh::set_error("rh_layers_add: batch has no breaking points (call rh_align_breaking_points)");
    return RH_ERR_STATE;
  }
  if (v.n != n || v.window_length != s->window_length || v.device != s->device) {
    rh::set_error("rh_layers_add: sources, window length or device differ from the batch's");
    return RH_ERR_ARGUMENT;
  }
  const uint64_t add = v.points_total / 2;
  if (s->used + add > s->capacity) {
    rh::set_error("rh_layers_add: set capacity too small");
    return RH_ERR_ARGUMENT;
  }
  for (uint32_t i = 0; i < n; ++i) {
    if (sources[i].quality != RH_NO_QUALITY &&
        (sources[i].quality > s->quality_size ||
         sources[i].query_length > s->quality_size - sources[i].quality)) {
      rh::set_error("rh_layers_add: source " + std::to_string(i) +
                    " reads past the quality bytes");
      return RH_ERR_ARGUMENT;
    }
  }
  if (n == 0) return RH_OK;
  RH_HIP_CHECK(hipSetDevice(s->device), RH_ERR_DEVICE);
  RH_HIP_CHECK(s->d_sources.reserve(sizeof(rh_layer_source) * n), RH_ERR_MEMORY);
  // on the batch's stream, after its breaking points
  RH_HIP_CHECK(hipMemcpyAsync(s->d_sources.ptr, sources, sizeof(rh_layer_source) * n,
                              hipMemcpyHostToDevice, v.stream),
               RH_ERR_DEVICE);
  const uint32_t blocks = (n + rh::kLayerWavesPerBlock - 1) / rh::kLayerWavesPerBlock;
  hipLaunchKernelGGL(rh::layer_select_kernel, dim3(blocks),
                     dim3(rh::kLayerWave * rh::kLayerWavesPerBlock), 0, v.stream, v.spans,
                     v.counts, v.points, s->d_sources.as<rh_layer_source>(), n,
                     s->window_length, 0.02 * s->window_length, s->quality_threshold,
                     s->d_quality.as<uint8_t>(), s->quality_size, s->used, s->none_key,
                     s->d_keys.as<uint32_t>(), s->d_slots.as<uint32_t>(),
                     s->d_records.as<rh_layer_record>(), s->d_kept.as<unsigned long long>());
  RH_HIP_CHECK(hipGetLastError(), RH_ERR_KERNEL);
  RH_HIP_CHECK(hipStreamSynchronize(v.stream), RH_ERR_DEVICE);
  s->used += add;
  return RH_OK;
}

int32_t rh_layers_finish(rh_layer_set* s, uint64_t* n_layers) {
  rh::clear_error();
  if (s == nullptr || n_layers == nullptr) {
    rh::set_error("rh_layers_finish: null argument");
    return RH_ERR_ARGUMENT;
  }
  RH_HIP_CHECK(hipSetDevice(s->device), RH_ERR_DEVICE);
  if (!s->finished) {
    unsigned long long tally[2] = {0, 0};  // layers kept, layers past the last window
    RH_HIP_CHECK(hipMemcpyAsync(tally, s->d_kept.ptr, sizeof(tally), hipMemcpyDeviceToHost,
                                s->stream),
                 RH_ERR_DEVICE);
    RH_HIP_CHECK(hipStreamSynchronize(s->stream), RH_ERR_DEVICE);
    const unsigned long long kept = tally[0];
    if (tally[1] != 0) {
      rh::set_error("rh_layers_finish: " + std::to_string(tally[1]) +
                    " layers fall past the set's windows (wrong first_window)");
      return RH_ERR_ARGUMENT;
    }
    if (kept > s->used) {
      rh::set_error("rh_layers_finish: more layers than slots");
      return RH_ERR_KERNEL;
    }
    s->kept = kept;
    if (s->used > 0) {
      const int bits = rh::key_bits(s->none_key);
      const uint32_t items = static_cast<uint32_t>(s->used);
      size_t temp = 0;
      RH_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
                       nullptr, temp, s->d_keys.as<uint32_t>(), s->d_keys_sorted.as<uint32_t>(),
                       s->d_slots.as<uint32_t>(), s->d_slots_sorted.as<uint32_t>(), items, 0,
                       bits, s->stream),
                   RH_ERR_DEVICE);
      RH_HIP_CHECK(s->d_temp.reserve(std::max<size_t>(temp, 1)), RH_ERR_MEMORY);
      RH_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
                       s->d_temp.ptr, temp, s->d_keys.as<uint32_t>(),
                       s->d_keys_sorted.as<uint32_t>(), s->d_slots.as<uint32_t>(),
                       s->d_slots_sorted.as<uint32_t>(), items, 0, bits, s->stream),
                   RH_ERR_DEVICE);
      RH_HIP_CHECK(s->d_grouped.reserve(sizeof(rh_layer_record) * std::max<uint64_t>(kept, 1)),
                   RH_ERR_MEMORY);
      if (kept > 0) {
        const uint32_t threads = 256;
        const uint64_t blocks = (kept + threads - 1) / threads;
        hipLaunchKernelGGL(rh::layer_gather_kernel, dim3(static_cast<uint32_t>(blocks)),
                           dim3(threads), 0, s->stream, s->d_slots_sorted.as<uint32_t>(),
                           s->d_records.as<rh_layer_record>(), static_cast<uint64_t>(kept),
                           s->d_grouped.as<rh_layer_record>());
        RH_HIP_CHECK(hipGetLastError(), RH_ERR_KERNEL);
      }
      RH_HIP_CHECK(hipStreamSynchronize(s->stream), RH_ERR_DEVICE);
    }
    s->finished = true;
  }
  *n_layers = s->kept;
  return RH_OK;
}

int32_t rh_layers_download(rh_layer_set* s, rh_layer_record* records, uint64_t capacity) {
  rh::clear_error();
  if (s == nullptr || (records == nullptr && capacity != 0)) {
    rh::set_error("rh_layers_download: null argument");
    return RH_ERR_ARGUMENT;
  }
  if (!s->finished) {
    rh::set_error("rh_layers_download: call rh_layers_finish first");
    return RH_ERR_STATE;
  }
  if (capacity < s->kept) {
    rh::set_error("rh_layers_download: capacity too small");
    return RH_ERR_ARGUMENT;
  }
  if (s->kept == 0) return RH_OK;
  RH_HIP_CHECK(hipSetDevice(s->device), RH_ERR_DEVICE);
  RH_HIP_CHECK(hipMemcpyAsync(records, s->d_grouped.ptr, sizeof(rh_layer_record) * s->kept,
                              hipMemcpyDeviceToHost, s->stream),
               RH_ERR_DEVICE);
  RH_HIP_CHECK(hipStreamSynchronize(s->stream), RH_ERR_DEVICE);
  return RH_OK;
}

void rh_layers_destroy(rh_layer_set* s) {
  if (s == nullptr) return;
  (void)hipSetDevice(s->device);
  if (s->stream != nullptr) (void)hipStreamSynchronize(s->stream);
  for (rh::DeviceBuffer* b : {&s->d_quality, &s->d_sources, &s->d_keys, &s->d_keys_sorted,
                              &s->d_slots, &s->d_slots_sorted, &s->d_records, &s->d_grouped,
                              &s->d_kept, &s->d_temp}) {
    b->release();
  }
  if (s->stream != nullptr) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // extern "C"
