// Shared host/device layouts of the overlap-alignment batch.
#pragma once

#include <cstdint>

namespace rh {

struct AlignPair {
  uint64_t q_off;   // arena offset of the query bytes
  uint64_t t_off;   // arena offset of the target bytes
  uint64_t out_off; // output offset of the ops (capacity q_len + t_len)
  uint32_t q_len;
  uint32_t t_len;
};

// Breaking-point input of one aligned pair: the target span [t_begin, t_end)
// on the contig, the query position before the first aligned base minus one
// plus one (q_first: the walk starts at q_first - 1), and where its points go
// (out, in points of two uint32 values) with room for cap points.
struct BreakSpan {
  uint32_t t_begin;
  uint32_t t_end;
  uint32_t q_first;
  uint32_t cap;
  uint64_t out;
};

struct AlignResult {
  uint32_t n_ops;
  int32_t edit_distance;
  int32_t status;  // 0 ok, < 0 error
  uint32_t pad;    // wave steps the pair took (each step: one column of up to 64 blocks)
};

enum : int32_t {
  kAlignOk = 0,
  kAlignNoSplit = -1,    // no optimal split found (inconsistent score)
  kAlignStack = -2,      // subproblem stack exhausted
  kAlignScratch = -3,    // scratch slot too small
  kAlignNotRun = -100,
};

// Leaf size rule of the path search: a subproblem is traced back directly
// when (2 * 8 + 4) * ceil(|q| / 64) * |t| < 1 MiB, else split at |t| / 2.
constexpr uint64_t kLeafBytes = 1024 * 1024;
constexpr uint32_t kAlignStackDepth = 128;

struct AlignSlotLayout {
  uint64_t col_left;   // int32 [q cap]
  uint64_t col_right;  // int32 [q cap]
  uint64_t carry;      // int8  [t cap + 64]
  uint64_t store_p;    // uint64 [store words]
  uint64_t store_m;    // uint64 [store words]
  uint64_t store_s;    // int32  [store words]
  uint64_t ops_tmp;    // uint8 [q cap + t cap]
  uint64_t total;
};

struct AlignSlotCaps {
  uint32_t q;            // longest query
  uint32_t t;            // longest target
  uint64_t store_words;  // stored (column, block) entries for a leaf
};

inline uint64_t align_up_a(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

inline AlignSlotLayout align_slot_layout(const AlignSlotCaps& c) {
  AlignSlotLayout l{};
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    uint64_t at = o;
    o = align_up_a(o + bytes, 256);
    return at;
  };
  l.col_left = take(4ull * (c.q + 64));
  l.col_right = take(4ull * (c.q + 64));
  l.carry = take(uint64_t(c.t) + 128);
  l.store_p = take(8ull * c.store_words);
  l.store_m = take(8ull * c.store_words);
  l.store_s = take(4ull * c.store_words);
  l.ops_tmp = take(uint64_t(c.q) + c.t + 64);
  l.total = o;
  return l;
}

// Slot capacities for pairs of at most max_q query and max_t target bytes:
// the score columns and carries, and the stored columns of the largest leaf
// (under kLeafBytes at 20 bytes per stored word) plus one block row's slack.
inline AlignSlotCaps align_caps_for(uint32_t max_q, uint32_t max_t) {
  AlignSlotCaps caps{};
  caps.q = max_q < 64 ? 64 : max_q;
  caps.t = max_t < 64 ? 64 : max_t;
  const uint64_t max_blocks = (caps.q + 63) / 64 + 1;
  caps.store_words = kLeafBytes / 20 + 64 * (max_blocks + 2);
  return caps;
}

struct AlignParams {
  const AlignPair* pairs;
  const uint8_t* arena;
  uint8_t* ops;
  AlignResult* results;
  const uint32_t* order;  // queue position -> pair index; null: identity
  uint32_t n_order;
  uint32_t* queue_head;
  uint8_t* scratch;
  uint64_t slot_bytes;
  AlignSlotCaps caps;
  AlignSlotLayout layout;
  uint32_t band0;  // first band guess (distance bound) of a whole pair; 0: automatic
  uint32_t pad;
};

}  // namespace rh
