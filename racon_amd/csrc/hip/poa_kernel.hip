// Partial-order alignment consensus of racon windows on gfx950.
//
// One 64-lane wave owns one window at a time (a "slot" of device scratch).
// Per layer, in racon's order (window.cpp:89-113):
//   1. the subgraph the layer is aligned to: nodes reached backwards from
//      node `end` through in-edges and aligned links without passing below
//      node id `begin` (whole graph when the layer spans the window);
//   2. its DFS topological order (roots in increasing node id, predecessors
//      then aligned nodes pushed, an aligned set emitted right after the
//      node that reached it first);
//   3. linear-gap NW of the layer against that order: lanes own 8-column
//      strips of each DP row, predecessor rows stream from the slot, the
//      horizontal gap chain is a wave prefix-max, and each cell records the
//      first matching move (diagonal over predecessors in in-edge order,
//      then vertical, then horizontal) as a 16-bit code;
//   4. the traceback from the first sink row with the highest last-column
//      score, following the codes;
//   5. insertion of the layer into the graph (edge weight = w[i-1] + w[i],
//      w = phred - 33 or 1) and a new topological order.
// After the last layer: heaviest-bundle path with branch completion,
// per-position coverage and racon's trimming (window.cpp:115-139).
//
// The graph walks (steps 1, 2, 4, 5 and the consensus) are inherently
// sequential; the wave runs them lane-uniformly (every lane executes the same
// loads and branches) so no lane-0 divergence or broadcast is needed, and the
// slot stays L2-resident for small windows. Parallelism comes from thousands
// of windows in flight (persistent kernel over a size-ordered queue) and from
// the 64-lane DP rows.
#include <hip/hip_runtime.h>

#include "poa_types.hpp"

namespace rh {

namespace {

constexpr int kWave = 64;
// Launch-invariant data (parameters, slot layout, window and layer tables)
// is read through constant-address-space references: uniform scalar loads
// that no store of the kernel can clobber, hoisted and kept in scalar
// registers.
typedef __attribute__((address_space(1))) uint8_t GByte;
// Device buffers named by the parameters, as global-memory pointers (they
// are loaded from memory, so the compiler cannot infer their address space).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* global(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}
typedef const __attribute__((address_space(4))) PoaParams CParams;
typedef const __attribute__((address_space(4))) PoaSlotLayout CLayout;
typedef const __attribute__((address_space(4))) PoaSlotCaps CCaps;
typedef const __attribute__((address_space(4))) PoaWindow CWindow;
typedef const __attribute__((address_space(4))) PoaLayer CLayer;

// Resident waves per SIMD the kernel is register-allocated for (168 VGPRs
// at 3, 128 at 4).
#ifndef RH_POA_WAVES_PER_EU
#define RH_POA_WAVES_PER_EU 4
#endif

// Aligned-set entries kept per node. The default build keeps kMaxAligned;
// poa_kernel_wide.hip compiles this file again with kMaxAlignedWide (and
// "_wide" kernel and launcher names) for the windows whose columns hold more
// distinct bytes (IUPAC codes), which the host reruns there.
#ifndef RH_POA_ALIGNED
#define RH_POA_ALIGNED kMaxAligned
#endif
#ifdef RH_POA_WIDE
#define RH_POA_SYM(name) name##_wide
#else
#define RH_POA_SYM(name) name
#endif
constexpr uint32_t kAligned = RH_POA_ALIGNED;

constexpr int kStrip = 4;                  // DP columns per lane per tile
constexpr int kTile = kWave * kStrip;      // 256 columns per tile
constexpr int32_t kNegInf = -(1 << 29);
constexpr int32_t kNegKey = -(1 << 30);  // below every (score, 255 - k) key
// Move codes (16 bits per DP cell): bits 15-14 the move (diagonal,
// vertical, horizontal); for diagonal and vertical moves bits 12-0 name the
// predecessor row, as the row delta r - p, or, with kCodeViaK set, as the
// index k of the predecessor in the row's in-edge list.
constexpr uint16_t kCodeDiag = 0x0000;
constexpr uint16_t kCodeVert = 0x4000;
constexpr uint16_t kCodeHorz = 0x8000;
constexpr uint16_t kCodeViaK = 0x2000;
constexpr uint32_t kCodePayload = 0x1fff;
// Rows with one predecessor in the packed DP store only the move of each
// cell, as a 2-bit kind 2 fd + fv from the flags fd (not diagonal) and fv
// (not vertical): 0 or 1 diagonal, 2 vertical, 3 horizontal (column j at
// bits 2 (j - 1) % 8 of byte (j - 1) / 4 of the row's code memory). The
// predecessor is implied, and column 0 is always vertical.
constexpr uint32_t kKindVert = 2;
constexpr uint32_t kKindHorz = 3;
// Bits 16..22 of a row's meta word: the row's entry in the slot's pool of
// global score rows (a row that found no LDS slot, in a pass with a pool).
constexpr int kRowMetaGShift = 16;
constexpr int kRowMetaGMask = 127;
static_assert(kPoaHPoolRows <= kRowMetaGMask + 1, "pool entries fit the meta bits");

__device__ __forceinline__ int lane_id() { return threadIdx.x; }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// v_writelane: `old` with lane `lane` set to the uniform `value` (the
// intrinsic has no clang builtin here, so it is bound by name)
extern "C" __device__ int32_t rh_amdgcn_writelane(int32_t value, int32_t lane, int32_t old) __asm(
    "llvm.amdgcn.writelane.i32");
__device__ __forceinline__ int lane_write(int value, int lane, int old) {
  return rh_amdgcn_writelane(value, lane, old);
}

__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// Phase clock: accumulates device constant-clock ticks into R.phase[k].
struct PhaseClock {
  uint64_t t;
  __device__ PhaseClock() : t(wall_clock64()) {}
  __device__ void lap(PoaResult& R, int k) {
    const uint64_t now = wall_clock64();
    R.phase[k] += now - t;
    t = now;
  }
};

// Views of one slot's arrays: one base pointer plus the layout's offsets
// (kernel arguments), so the array addresses are rebuilt where used instead of
// occupying registers.
struct Slot {
  GByte* p;  // slot base (global memory)
  CLayout* L;
  __device__ uint8_t* base() const { return (uint8_t*)(p + L->base); }
  __device__ int32_t* in_head() const { return (int32_t*)(p + L->in_head); }
  __device__ int32_t* out_head() const { return (int32_t*)(p + L->out_head); }
  __device__ int32_t* in_tail() const { return (int32_t*)(p + L->in_tail); }
  __device__ int32_t* out_tail() const { return (int32_t*)(p + L->out_tail); }
  __device__ uint8_t* n_aligned() const { return (uint8_t*)(p + L->n_aligned); }
  __device__ int32_t* aligned() const { return (int32_t*)(p + L->aligned); }
  __device__ int32_t* cov() const { return (int32_t*)(p + L->cov); }
  __device__ int32_t* last_label() const { return (int32_t*)(p + L->last_label); }
  __device__ int32_t* e_tail() const { return (int32_t*)(p + L->e_tail); }
  __device__ int32_t* e_head() const { return (int32_t*)(p + L->e_head); }
  __device__ int32_t* e_weight() const { return (int32_t*)(p + L->e_weight); }
  __device__ int32_t* e_next_in() const { return (int32_t*)(p + L->e_next_in); }
  __device__ int32_t* e_next_out() const { return (int32_t*)(p + L->e_next_out); }
  __device__ int32_t* rank_to_node() const { return (int32_t*)(p + L->rank_to_node); }
  __device__ int32_t* node_rank() const { return (int32_t*)(p + L->node_rank); }
  __device__ uint8_t* mark() const { return (uint8_t*)(p + L->mark); }
  __device__ uint8_t* lead() const { return (uint8_t*)(p + L->lead); }
  __device__ int32_t* stack() const { return (int32_t*)(p + L->stack); }
  __device__ uint8_t* in_sub() const { return (uint8_t*)(p + L->in_sub); }
  __device__ int32_t* row_of() const { return (int32_t*)(p + L->row_of); }
  __device__ int32_t* sub_order() const { return (int32_t*)(p + L->sub_order); }
  __device__ int32_t* pred_off() const { return (int32_t*)(p + L->pred_off); }
  __device__ int32_t* pred_rows() const { return (int32_t*)(p + L->pred_rows); }
  __device__ int32_t* row_meta() const { return (int32_t*)(p + L->row_meta); }
  __device__ int32_t* H() const { return (int32_t*)(p + L->H); }
  // global score row of DP row `row` (meta: its row meta word): its pool
  // entry when the slot keeps a pool, else row `row` of the matrix
  __device__ int32_t* H_row(int row, int meta, int pitch) const {
    const int64_t at = L->hpool != 0 ? ((meta >> kRowMetaGShift) & kRowMetaGMask) : row;
    return H() + at * pitch;
  }
  __device__ int32_t* H_row(int row, int pitch) const {
    return L->hpool != 0 ? H_row(row, __builtin_amdgcn_readfirstlane(row_meta()[row - 1]), pitch)
                         : H() + static_cast<int64_t>(row) * pitch;
  }
  __device__ uint16_t* code() const { return (uint16_t*)(p + L->code); }
  __device__ int64_t* scores() const { return (int64_t*)(p + L->scores); }
  __device__ int32_t* preds() const { return (int32_t*)(p + L->preds); }
  __device__ int32_t* path() const { return (int32_t*)(p + L->path); }
  __device__ uint8_t* seq() const { return (uint8_t*)(p + L->seq); }
  __device__ int32_t* qual() const { return (int32_t*)(p + L->qual); }
  __device__ int32_t* anchor() const { return (int32_t*)(p + L->anchor); }
  __device__ int32_t* lpath() const { return (int32_t*)(p + L->lpath); }
  __device__ int32_t* sink_score() const { return (int32_t*)(p + L->sink_score); }
  __device__ int32_t* tie_mark() const { return (int32_t*)(p + L->tie_mark); }
  __device__ uint64_t* row_pred4() const { return (uint64_t*)(p + L->row_pred4); }

  int32_t n_nodes;
  int32_t n_edges;
  int32_t n_seqs;
  bool overflow;
};

__device__ __forceinline__ Slot carve(GByte* p, CLayout& l) {
  Slot s;
  s.p = p;
  s.L = &l;
  s.n_nodes = 0;
  s.n_edges = 0;
  s.n_seqs = 0;
  s.overflow = false;
  return s;
}

// ---- graph construction (lane-uniform) ----

__device__ __forceinline__ void touch(Slot& s, int node, int label) {
  if (s.last_label()[node] != label) {
    s.last_label()[node] = label;
    s.cov()[node] = s.cov()[node] + 1;
  }
}

__device__ __forceinline__ int add_node(Slot& s, CCaps& caps, uint8_t base) {
  if (s.n_nodes >= static_cast<int>(caps.nodes)) {
    s.overflow = true;
    return 0;
  }
  const int v = s.n_nodes++;
  s.base()[v] = base;
  s.in_head()[v] = -1;
  s.out_head()[v] = -1;
  s.in_tail()[v] = -1;
  s.out_tail()[v] = -1;
  s.n_aligned()[v] = 0;
  s.cov()[v] = 0;
  s.last_label()[v] = -1;
  return v;
}

__device__ __forceinline__ void add_edge(Slot& s, CCaps& caps, int tail, int head, int weight) {
  for (int e = s.out_head()[tail]; e != -1; e = s.e_next_out()[e]) {
    if (s.e_head()[e] == head) {
      s.e_weight()[e] = s.e_weight()[e] + weight;
      touch(s, tail, s.n_seqs);
      touch(s, head, s.n_seqs);
      return;
    }
  }
  if (s.n_edges >= static_cast<int>(caps.edges)) {
    s.overflow = true;
    return;
  }
  const int e = s.n_edges++;
  s.e_tail()[e] = tail;
  s.e_head()[e] = head;
  s.e_weight()[e] = weight;
  s.e_next_in()[e] = -1;
  s.e_next_out()[e] = -1;
  if (s.out_tail()[tail] == -1) {
    s.out_head()[tail] = e;
  } else {
    s.e_next_out()[s.out_tail()[tail]] = e;
  }
  s.out_tail()[tail] = e;
  if (s.in_tail()[head] == -1) {
    s.in_head()[head] = e;
  } else {
    s.e_next_in()[s.in_tail()[head]] = e;
  }
  s.in_tail()[head] = e;
  touch(s, tail, s.n_seqs);
  touch(s, head, s.n_seqs);
}

// DFS topological order over nodes with keep[v] (all nodes when keep is
// null); writes node ids in rank order to out, returns how many.
__device__ __forceinline__ int topo_sort(Slot& s, const uint8_t* keep, int32_t* out, int stack_cap) {
  const int n = s.n_nodes;
  for (int v = lane_id(); v < n; v += kWave) {
    s.mark()[v] = 0;
    s.lead()[v] = 1;
  }
  wave_sync();
  int count = 0;
  for (int root = 0; root < n; ++root) {
    if (keep != nullptr && keep[root] == 0) {
      continue;
    }
    if (s.mark()[root] != 0) {
      continue;
    }
    int top = 0;
    s.stack()[top++] = root;
    while (top > 0) {
      const int v = uniform(s.stack()[top - 1]);
      bool ready = true;
      if (s.mark()[v] != 2) {
        for (int e = uniform(s.in_head()[v]); e != -1; e = uniform(s.e_next_in()[e])) {
          const int u = uniform(s.e_tail()[e]);
          if (keep != nullptr && keep[u] == 0) {
            continue;
          }
          if (s.mark()[u] != 2) {
            if (top >= stack_cap) {
              s.overflow = true;
              return count;
            }
            s.stack()[top++] = u;
            ready = false;
          }
        }
        if (s.lead()[v] != 0) {
          const int na = s.n_aligned()[v];
          for (int k = 0; k < na; ++k) {
            const int a = uniform(s.aligned()[v * kAligned + k]);
            if (keep != nullptr && keep[a] == 0) {
              continue;
            }
            if (s.mark()[a] != 2) {
              if (top >= stack_cap) {
                s.overflow = true;
                return count;
              }
              s.stack()[top++] = a;
              s.lead()[a] = 0;
              ready = false;
            }
          }
        }
        if (ready) {
          s.mark()[v] = 2;
          if (s.lead()[v] != 0) {
            if (count + 1 + s.n_aligned()[v] > n) {
              s.overflow = true;
              return count;
            }
            out[count++] = v;
            const int na = s.n_aligned()[v];
            for (int k = 0; k < na; ++k) {
              const int a = uniform(s.aligned()[v * kAligned + k]);
              if (keep == nullptr || keep[a] != 0) {
                out[count++] = a;
              }
            }
          }
        } else {
          s.mark()[v] = 1;
        }
      }
      if (ready) {
        --top;
      }
    }
  }
  return count;
}

// ---- DP ----

// Inclusive prefix max over the 64 lanes with DPP row shifts and row
// broadcasts (no LDS crossbar traffic).

__device__ __forceinline__ int32_t wave_prefix_max(int32_t x) {
  // within rows of 16 lanes: shifts 1, 2, 4, 8 (lanes shifted in from
  // outside the row keep the identity). INT32_MIN, the identity of max, lets
  // each move fold into one v_max_i32 with a DPP source.
  constexpr int32_t kId = INT32_MIN;
  x = max(x, __builtin_amdgcn_update_dpp(kId, x, 0x111, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(kId, x, 0x112, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(kId, x, 0x114, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(kId, x, 0x118, 0xf, 0xf, false));
  // across rows: lane 15 into rows 1 and 3, lane 31 into rows 2 and 3
  x = max(x, __builtin_amdgcn_update_dpp(kId, x, 0x142, 0xa, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(kId, x, 0x143, 0xc, 0xf, false));
  return x;
}

// Value of the previous lane (lane 0 gets `first`), via DPP wave shift.
__device__ __forceinline__ int32_t wave_shift_right(int32_t x, int32_t first) {
  return __builtin_amdgcn_update_dpp(first, x, 0x138, 0xf, 0xf, false);
}

// Wavefront-scope ordering of LDS/global traffic between lanes of this wave
// (a single-wave workgroup needs no s_barrier; the fence keeps the compiler
// from moving memory operations across it).
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Waits for this wave's outstanding global stores (vmcnt(0)).
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

constexpr int kRowMetaSink = 0x100;
// The row has one predecessor (or only the virtual start row).
constexpr int kRowMetaSingle = 0x200;
// Bits 10..14 of a row's meta word: where the row's scores live while later
// rows need them. Slots 0..kMaxLdsSlots-1 are LDS rows; kSlotGlobal rows go
// through the slot's global H matrix; kSlotNone rows are read by no later row.
constexpr int kRowMetaSlotShift = 10;
constexpr int kSlotGlobal = 30;
constexpr int kSlotNone = 31;
constexpr int kMaxLdsSlots = 30;
// Predecessor list entries: DP row in the low 24 bits, its slot above.
constexpr int kPredRowBits = 24;
constexpr int kPredRowMask = (1 << kPredRowBits) - 1;

// LDS score rows keep column j at element j + kColShift, so that the four
// columns of a lane's strip (j0 = 1 + 4k) start on an aligned quad.
constexpr int kColShift = 3;

// Code matrix index of column j inside a row: columns 1..L at j - 1 (aligned
// quads per strip), column 0 at the row's last element (pitch >= L + 9).
__device__ __forceinline__ int code_index(int j, int pitch) { return j == 0 ? pitch - 1 : j - 1; }

template <typename T>
struct Quad;

template <>
struct Quad<int16_t> {
  static __device__ __forceinline__ void load(const int16_t* p, int32_t* v) {
    const uint2 q = *reinterpret_cast<const uint2*>(p);
    v[0] = static_cast<int16_t>(q.x & 0xffffu);
    v[1] = static_cast<int32_t>(q.x) >> 16;
    v[2] = static_cast<int16_t>(q.y & 0xffffu);
    v[3] = static_cast<int32_t>(q.y) >> 16;
  }
  static __device__ __forceinline__ void store(int16_t* p, const int32_t* v) {
    uint2 q;
    q.x = (static_cast<uint32_t>(v[0]) & 0xffffu) | (static_cast<uint32_t>(v[1]) << 16);
    q.y = (static_cast<uint32_t>(v[2]) & 0xffffu) | (static_cast<uint32_t>(v[3]) << 16);
    *reinterpret_cast<uint2*>(p) = q;
  }
};

template <>
struct Quad<int32_t> {
  static __device__ __forceinline__ void load(const int32_t* p, int32_t* v) {
    const int4 q = *reinterpret_cast<const int4*>(p);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  }
  static __device__ __forceinline__ void store(int32_t* p, const int32_t* v) {
    *reinterpret_cast<int4*>(p) = make_int4(v[0], v[1], v[2], v[3]);
  }
};

// All DP rows of one layer. Lanes own 4-column strips of 256-column tiles;
// predecessor rows come from their LDS slots (the global H matrix for rows
// that found no free slot); the horizontal gap chain is a wave prefix-max;
// every cell stores the first move that reproduces its score (diagonal,
// then vertical, each over predecessors in in-edge order, then horizontal).
// Returns the DP row of the end cell: the first sink row with the highest
// score in column L.
//
// The global-H branch waits for its own loads before leaving the branch, so
// the common LDS path never waits on vector memory (on gfx9 the vector memory
// counter also counts the traceback-code stores of earlier rows). Strips
// that start past column L are idle; the columns past L inside the last
// strip are computed from padding and never read back.
// General DP (any score width, layer length and in-degree).
template <typename T>
__device__ __forceinline__ int dp_layer(CParams& P, const Slot& s, uint8_t* lds, int V, int L,
                        int* best_score, bool* tied) {
  static_assert(kStrip == 4, "strips are loaded and stored as quads");
  const int lane = lane_id();
  const int pitch = static_cast<int>(P.caps.cols);
  // scores as uniform registers (a select between two parameter fields
  // would otherwise become a per-lane load from the selected address)
  const int32_t gap = uniform(P.gap);
  const int32_t match = uniform(P.match);
  const int32_t mismatch = uniform(P.mismatch);
  T* ring = reinterpret_cast<T*>(lds);
  const uint8_t* seq = lds + align_up(static_cast<uint64_t>(P.ring_rows) * pitch * 2, 16);
  int32_t best = kNegInf;
  int best_row = 0;
  int ties = 0;
  const int total_preds = uniform(s.pred_off()[V]);
  // rows are processed in chunks of 64: lane k holds the metadata of row
  // r0 + k and a 128-entry window of the predecessor list, broadcast per row
  for (int r0 = 1; r0 <= V; r0 += kWave) {
  const int rk = r0 + lane;
  int off_v = 0, end_v = 0, meta_v = 0;
  if (rk <= V) {
    off_v = s.pred_off()[rk - 1];
    end_v = s.pred_off()[rk];
    meta_v = s.row_meta()[rk - 1];
  }
  const int base = __builtin_amdgcn_readlane(off_v, 0);
  const int pr0 = base + lane < total_preds ? s.pred_rows()[base + lane] : 0;
  const int pr1 = base + kWave + lane < total_preds ? s.pred_rows()[base + kWave + lane] : 0;
  // (up to four entries per row of the chunk without a reload)
  const int pr2 =
      base + 2 * kWave + lane < total_preds ? s.pred_rows()[base + 2 * kWave + lane] : 0;
  const int pr3 =
      base + 3 * kWave + lane < total_preds ? s.pred_rows()[base + 3 * kWave + lane] : 0;
  // complete the chunk's loads here: otherwise the row loop below would wait
  // for them (and for every code store before them) on each row
  wait_stores();
  const int chunk = min(kWave, V - r0 + 1);
  for (int rr = 0; rr < chunk; ++rr) {
    const int r = r0 + rr;
    const int off = __builtin_amdgcn_readlane(off_v, rr);
    const int np = __builtin_amdgcn_readlane(end_v, rr) - off;
    const int meta = __builtin_amdgcn_readlane(meta_v, rr);
    auto pred_at = [&](int k) -> int {
      const int idx = off + k - base;
      if (idx < kWave) return __builtin_amdgcn_readlane(pr0, idx);
      if (idx < 2 * kWave) return __builtin_amdgcn_readlane(pr1, idx - kWave);
      if (idx < 3 * kWave) return __builtin_amdgcn_readlane(pr2, idx - 2 * kWave);
      if (idx < 4 * kWave) return __builtin_amdgcn_readlane(pr3, idx - 3 * kWave);
      const int e = uniform(s.pred_rows()[off + k]);
      wait_stores();
      return e;
    };
    const uint32_t b = static_cast<uint32_t>(meta & 0xff);
    const int own = (meta >> kRowMetaSlotShift) & 31;
    T* Hr = ring + (own < kSlotGlobal ? own : 0) * pitch;
    int32_t* Hg = s.H_row(r, meta, pitch);
    uint16_t* Cr = s.code() + static_cast<int64_t>(r) * pitch;

    // column 0: max over predecessors of their column 0, plus gap
    int32_t c0 = kNegInf;
    int k0 = 0;
    for (int k = 0; k < np; ++k) {
      const int e = pred_at(k);
      const int p = e & kPredRowMask;
      const int sl = static_cast<unsigned>(e) >> kPredRowBits;
      int32_t h0;
      if (p == 0) {
        h0 = 0;
      } else if (sl < kSlotGlobal) {
        h0 = ring[sl * pitch + kColShift];
      } else {
        h0 = uniform(s.H_row(p, pitch)[0]);
        wait_stores();
      }
      if (h0 + gap > c0) {
        c0 = h0 + gap;
        k0 = k;
      }
    }
    if (lane == 0) {
      if (own < kSlotGlobal) Hr[kColShift] = static_cast<T>(c0);
      if (own == kSlotGlobal) Hg[0] = c0;
      Cr[pitch - 1] = kCodeVert | kCodeViaK | static_cast<uint16_t>(k0);
    }
    int32_t carry = c0;
    int32_t h_last = kNegInf;

    for (int t0 = 1; t0 <= L; t0 += kTile) {
      const int j0 = t0 + lane * kStrip;
      const bool live = j0 <= L;
      const int jr = live ? j0 : 1;  // idle strips read column 0..4 (in bounds)
      const uint32_t w4 = *reinterpret_cast<const uint32_t*>(seq + (jr - 1));
      int32_t sc[kStrip];
      bool in[kStrip];
#pragma unroll
      for (int u = 0; u < kStrip; ++u) {
        sc[u] = ((w4 >> (8 * u)) & 0xffu) == b ? match : mismatch;
        in[u] = j0 + u <= L;
      }
      int32_t bd[kStrip], bv[kStrip];
      uint32_t kk[kStrip];  // diagonal k (low 16 bits), vertical k (high 16 bits)
#pragma unroll
      for (int u = 0; u < kStrip; ++u) {
        bd[u] = kNegInf;
        bv[u] = kNegInf;
        kk[u] = 0;
      }
      for (int k = 0; k < np; ++k) {
        const int e = pred_at(k);
        const int p = e & kPredRowMask;
        const int sl = static_cast<unsigned>(e) >> kPredRowBits;
        int32_t hp[kStrip + 1];
        if (p == 0) {
#pragma unroll
          for (int u = 0; u <= kStrip; ++u) hp[u] = (j0 - 1 + u) * gap;
        } else if (sl < kSlotGlobal) {
          const T* row = ring + sl * pitch + kColShift;
          hp[0] = row[jr - 1];
          Quad<T>::load(row + jr, hp + 1);
        } else {
          const int32_t* row = s.H_row(p, pitch);
#pragma unroll
          for (int u = 0; u <= kStrip; ++u) {
            const int j = j0 - 1 + u;
            hp[u] = j <= L ? row[j] : kNegInf;
          }
          wait_stores();
        }
#pragma unroll
        for (int u = 0; u < kStrip; ++u) {
          const int32_t d = hp[u] + sc[u];
          const int32_t v = hp[u + 1] + gap;
          const bool dg = d > bd[u];
          const bool vg = v > bv[u];
          bd[u] = dg ? d : bd[u];
          bv[u] = vg ? v : bv[u];
          kk[u] = dg ? ((kk[u] & 0xffff0000u) | static_cast<uint32_t>(k)) : kk[u];
          kk[u] = vg ? ((kk[u] & 0xffffu) | (static_cast<uint32_t>(k) << 16)) : kk[u];
        }
      }
      // horizontal chain: H[j] = j*g + max_{k<=j} (t_k - k*g)
      int32_t pm = kNegInf;
      int32_t pre[kStrip];
#pragma unroll
      for (int u = 0; u < kStrip; ++u) {
        const int j = j0 + u;
        const int32_t t = max(bd[u], bv[u]);
        pre[u] = in[u] ? t - j * gap : kNegInf;
        pm = max(pm, pre[u]);
        pre[u] = pm;
      }
      const int32_t incl = wave_prefix_max(pm);
      const int32_t excl = wave_shift_right(incl, kNegInf);
      const int32_t in_carry = max(excl, carry - (t0 - 1) * gap);
      int32_t h[kStrip];
      uint32_t cd[kStrip];
#pragma unroll
      for (int u = 0; u < kStrip; ++u) {
        h[u] = max(pre[u], in_carry) + (j0 + u) * gap;
        const uint32_t c_diag = kCodeDiag | kCodeViaK | (kk[u] & 0xffffu);
        const uint32_t c_vert = kCodeVert | kCodeViaK | (kk[u] >> 16);
        cd[u] = bd[u] == h[u] ? c_diag : (bv[u] == h[u] ? c_vert : kCodeHorz);
      }
      if (live) {
        uint2 q;
        q.x = cd[0] | (cd[1] << 16);
        q.y = cd[2] | (cd[3] << 16);
        *reinterpret_cast<uint2*>(Cr + (j0 - 1)) = q;
        if (own < kSlotGlobal) {
          Quad<T>::store(Hr + kColShift + j0, h);
        }
      }
      if (own == kSlotGlobal) {
#pragma unroll
        for (int u = 0; u < kStrip; ++u) {
          if (in[u]) Hg[j0 + u] = h[u];
        }
      }
      // value of the tile's last column: carry into the next tile and, in
      // the last tile, the end-cell candidate
      const int last_col = min(L, t0 + kTile - 1);
      const int owner = (last_col - t0) / kStrip;
      const int last_u = (last_col - t0) % kStrip;
      int32_t lv = h[0];
#pragma unroll
      for (int u = 1; u < kStrip; ++u) {
        if (u == last_u) lv = h[u];
      }
      carry = __builtin_amdgcn_readlane(lv, owner);
      h_last = carry;
    }
    if ((meta & kRowMetaSink) != 0) {
      if (lane == 0) {
        s.sink_score()[r] = h_last;
      }
      if (h_last > best) {
        best = h_last;
        best_row = r;
        ties = 0;
      } else if (h_last == best) {
        ++ties;
      }
    }
    if (own == kSlotGlobal) {
      wait_stores();
    }
    wave_fence();
  }
  }
  *best_score = best;
  *tied = ties > 0;
  return best_row;
}

// Strip rows in LDS (the one-pass DP below) are laid out quad-major: quad q
// (strip columns j0 + 4q .. j0 + 4q + 3) of lane k sits at element
// jq + q * qs with jq = 4k and qs = 4 * (lanes with columns <= L), so that
// consecutive lanes touch consecutive 8-byte quads (no LDS bank conflicts);
// column 0 sits at the row's last element. H[p][j0 - 1], the last column
// of the previous lane's strip, comes from that lane with a DPP shift.
__device__ __forceinline__ uint32_t from_prev_lane(uint32_t x) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x138, 0xf, 0xf, false));  // wave_shr:1
}

// Predecessor row p (entry e) at the lane's strip: hp[u] = H[p][j0 - 1 + u],
// u = 0..W. Row 0 is the virtual start row (j * gap).
template <int W>
__device__ __forceinline__ void load_strip(const int16_t* ring, int pitch, const Slot& s, int e,
                                           int j0, int jq, int qs, int L, int32_t gap,
                                           int32_t* hp) {
  const int p = e & kPredRowMask;
  const int sl = static_cast<unsigned>(e) >> kPredRowBits;
  if (p == 0) {
#pragma unroll
    for (int u = 0; u <= W; ++u) hp[u] = (j0 - 1 + u) * gap;
  } else if (sl < kSlotGlobal) {
    const int16_t* row = ring + sl * pitch;
#pragma unroll
    for (int q = 0; q < W / 4; ++q) Quad<int16_t>::load(row + jq + q * qs, hp + 1 + 4 * q);
    const int32_t c0 = row[pitch - 1];
    const int32_t prev = static_cast<int32_t>(from_prev_lane(static_cast<uint32_t>(hp[W])));
    hp[0] = lane_id() == 0 ? c0 : prev;
  } else {
    const int32_t* row = s.H_row(p, pitch);
#pragma unroll
    for (int u = 0; u <= W; ++u) {
      const int j = j0 - 1 + u;
      hp[u] = j <= L ? row[j] : kNegInf;
    }
    wait_stores();
  }
}

// Rest of one DP row once the best diagonal (bd) and vertical (bv) moves of
// every strip column are known: horizontal chain, move codes, row storage.
// With kKeys the moves are (score << 8 | 255 - k) keys of predecessor k,
// else plain scores from predecessor 0. Lane 0's chain starts at column 0
// (key or score c0, code vertical). Returns H[row][L].
template <int W, bool kKeys>
__device__ __forceinline__ int32_t finish_row(const int32_t* bd, const int32_t* bv, int32_t c0,
                                              uint32_t pay0, int lane, int j0, int jq, int qs,
                                              bool live, int L, int32_t gap, int own, int pitch,
                                              int16_t* ring, int32_t* Hg, uint16_t* Cr,
                                              int end_lane, int end_u) {
  auto score = [](int32_t x) { return kKeys ? (x >> 8) : x; };
  // code payload of the move's predecessor: the row's single predecessor
  // (pay0: row delta or kCodeViaK), else kCodeViaK | k from the key
  auto pred = [&](int32_t x) -> uint32_t {
    if (!kKeys) return pay0;
    return kCodeViaK | (255u - (static_cast<uint32_t>(x) & 255u));
  };
  const int32_t h0 = score(c0);
  int32_t pm = lane == 0 ? h0 : kNegInf;
  int32_t pre[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    pm = max(pm, score(max(bd[u], bv[u])) - (j0 + u) * gap);
    pre[u] = pm;
  }
  const int32_t excl = wave_shift_right(wave_prefix_max(pm), kNegInf);
  int32_t h[W];
  uint32_t cd[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    h[u] = max(pre[u], excl) + (j0 + u) * gap;
    const uint32_t c = score(bv[u]) == h[u] ? (kCodeVert | pred(bv[u])) : kCodeHorz;
    cd[u] = score(bd[u]) == h[u] ? (kCodeDiag | pred(bd[u])) : c;
  }
  if (live) {
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
      uint2 v;
      v.x = cd[4 * q] | (cd[4 * q + 1] << 16);
      v.y = cd[4 * q + 2] | (cd[4 * q + 3] << 16);
      *reinterpret_cast<uint2*>(Cr + (j0 - 1) + 4 * q) = v;
    }
    if (own < kSlotGlobal) {
      int16_t* row = ring + own * pitch;
#pragma unroll
      for (int q = 0; q < W / 4; ++q) Quad<int16_t>::store(row + jq + q * qs, h + 4 * q);
    }
  }
  if (own == kSlotGlobal) {
#pragma unroll
    for (int u = 0; u < W; ++u) {
      if (j0 + u <= L) Hg[j0 + u] = h[u];
    }
  }
  if (lane == 0) {
    if (own < kSlotGlobal) ring[own * pitch + pitch - 1] = static_cast<int16_t>(h0);
    if (own == kSlotGlobal) Hg[0] = h0;
    Cr[pitch - 1] = static_cast<uint16_t>(kCodeVert | pred(c0));
  }
  int32_t last = h[0];
#pragma unroll
  for (int u = 1; u < W; ++u) {
    if (u == end_u) last = h[u];
  }
  return __builtin_amdgcn_readlane(last, end_lane);
}

// ---- packed 16-bit pairs: two DP columns per 32-bit register ----

typedef short __attribute__((ext_vector_type(2))) I16x2;
typedef unsigned short __attribute__((ext_vector_type(2))) U16x2;

__device__ __forceinline__ U16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(U16x2, x); }
__device__ __forceinline__ I16x2 as_i16x2(uint32_t x) { return __builtin_bit_cast(I16x2, x); }
__device__ __forceinline__ uint32_t pk_bits(U16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ uint32_t pk_bits(I16x2 x) { return __builtin_bit_cast(uint32_t, x); }
// per-half wrapping add / subtract, signed max, unsigned min, multiply-add
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
  return pk_bits(as_u16x2(a) + as_u16x2(b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
  return pk_bits(as_u16x2(a) - as_u16x2(b));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
  return pk_bits(__builtin_elementwise_max(as_i16x2(a), as_i16x2(b)));
}
__device__ __forceinline__ uint32_t pk_min_u(uint32_t a, uint32_t b) {
  return pk_bits(__builtin_elementwise_min(as_u16x2(a), as_u16x2(b)));
}
__device__ __forceinline__ uint32_t pk_mad(uint32_t a, uint32_t b, uint32_t c) {
  return pk_bits(as_u16x2(a) * as_u16x2(b) + as_u16x2(c));
}
// max of a with the low (resp. high) half of b in both halves
__device__ __forceinline__ uint32_t pk_max_lo(uint32_t a, uint32_t b) {
  const I16x2 v = as_i16x2(b);
  return pk_bits(__builtin_elementwise_max(as_i16x2(a), __builtin_shufflevector(v, v, 0, 0)));
}
__device__ __forceinline__ uint32_t pk_max_hi(uint32_t a, uint32_t b) {
  const I16x2 v = as_i16x2(b);
  return pk_bits(__builtin_elementwise_max(as_i16x2(a), __builtin_shufflevector(v, v, 1, 1)));
}
// (lo, max(lo, hi)): inclusive prefix max inside the pair
__device__ __forceinline__ uint32_t pk_scan(uint32_t a) {
  const I16x2 v = as_i16x2(a);
  return pk_bits(__builtin_elementwise_max(v, __builtin_shufflevector(v, v, 0, 0)));
}
__device__ __forceinline__ uint32_t pk_splat(int32_t v) {
  return (static_cast<uint32_t>(v) & 0xffffu) * 0x10001u;
}
__device__ __forceinline__ int32_t pk_half(uint32_t x, int hi) {
  return hi ? static_cast<int32_t>(x) >> 16 : static_cast<int32_t>(static_cast<int16_t>(x & 0xffffu));
}

// Per-lane constants of the packed rows: pair q holds strip columns
// j0 + 2q (low half) and j0 + 2q + 1 (high half).
template <int W>
struct PackedStrip {
  uint32_t base[W / 2];  // layer bytes of the pair, one per half
  uint32_t match_gap;    // match - gap and gap in both halves
  uint32_t gap;
  uint32_t one;          // 1 in both halves, opaque to the optimizer (a
  uint32_t c4000;        // visible 1 turns min(x, 1) into compares and selects)
  uint32_t diff;         // mismatch - match in both halves
};

template <int W>
__device__ __forceinline__ void packed_strip_init(PackedStrip<W>& ps, const uint32_t* sw, int j0,
                                                  int32_t match, int32_t mismatch, int32_t gap) {
#pragma unroll
  for (int q = 0; q < W / 2; ++q) {
    ps.base[q] = __builtin_amdgcn_perm(sw[q / 2], sw[q / 2], (q & 1) ? 0x0c030c02u : 0x0c010c00u);
  }
  ps.match_gap = pk_splat(match - gap);
  ps.gap = pk_splat(gap);
  ps.one = 0x00010001u;
  asm volatile("" : "+s"(ps.one));
  ps.c4000 = 0x40004000u;
  asm volatile("" : "+v"(ps.c4000));
  ps.diff = pk_splat(mismatch - match);
}

// Predecessor row p (entry e) at the lane's strip as pairs of the row's
// stored values G[p][j] = H[p][j] - j * gap: pp[q] = (G[p][j0 + 2q],
// G[p][j0 + 2q + 1]); returns G[p][j0 - 1] in the low half. Row 0 is the
// virtual start row (H = j * gap, so G = 0).
template <int W>
__device__ __forceinline__ uint32_t load_pairs(const int16_t* ring, int pitch, const Slot& s, int e,
                                               int j0, int jq, int qs, int L, int32_t gap,
                                               const PackedStrip<W>& ps, uint32_t* pp) {
  const int p = e & kPredRowMask;
  const int sl = static_cast<unsigned>(e) >> kPredRowBits;
  if (p == 0) {
#pragma unroll
    for (int q = 0; q < W / 2; ++q) pp[q] = 0u;
    return 0u;
  } else if (sl < kSlotGlobal) {
    const int16_t* row = ring + sl * pitch;
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
      const uint2 v = *reinterpret_cast<const uint2*>(row + jq + q * qs);
      pp[2 * q] = v.x;
      pp[2 * q + 1] = v.y;
    }
    const uint32_t c0 = static_cast<uint16_t>(row[pitch - 1]);
    const uint32_t prev = from_prev_lane(pp[W / 2 - 1]) >> 16;
    return lane_id() == 0 ? c0 : prev;
  } else {
    const int32_t* row = s.H_row(p, pitch);
    int32_t hv[W + 1];
#pragma unroll
    for (int u = 0; u <= W; ++u) {
      const int j = j0 - 1 + u;
      hv[u] = j <= L ? row[j] : 0;
    }
    wait_stores();
#pragma unroll
    for (int q = 0; q < W / 2; ++q) {
      pp[q] = (static_cast<uint32_t>(hv[2 * q + 1]) & 0xffffu) |
              (static_cast<uint32_t>(hv[2 * q + 2]) << 16);
    }
    return static_cast<uint32_t>(hv[0]);
  }
}

// Scores and move codes of one packed row (registers only).
template <int W>
struct PackedOut {
  uint32_t h[W / 2];
  uint32_t cd[W / 2];
  int32_t h0;  // column 0
};

template <int W, bool kMulti>
__device__ __forceinline__ void packed_compute(const PackedStrip<W>& ps, const uint32_t* pp,
                                               uint32_t hp0, const uint32_t* kk, uint32_t k0,
                                               uint32_t b, uint32_t pay0, int lane, int32_t gap,
                                               PackedOut<W>& o) {
  constexpr int H2 = W / 2;
  const uint32_t bb = b * 0x10001u;
  // column 0 (hp0 + gap) heads lane 0's chain
  o.h0 = pk_half(hp0 + static_cast<uint32_t>(gap), 0);
  uint32_t carry = lane == 0 ? pk_splat(o.h0) : 0x80008000u;
  uint32_t bd[H2], bv[H2], pre[H2];
#pragma unroll
  for (int q = 0; q < H2; ++q) {
    // diagonal sources H[p][j - 1] of the pair
    const uint32_t d = q == 0 ? ((pp[0] << 16) | (hp0 & 0xffffu))
                              : __builtin_amdgcn_alignbit(pp[q], pp[q - 1], 16);
    const uint32_t mis = pk_min_u(ps.base[q] ^ bb, ps.one);  // 1 where the bases differ
    bd[q] = pk_mad(mis, ps.diff, pk_add(d, ps.match_gap));
    bv[q] = pk_add(pp[q], ps.gap);
    carry = pk_max_hi(pk_scan(pk_max(bd[q], bv[q])), carry);
    pre[q] = carry;
  }
  const int32_t excl = wave_shift_right(wave_prefix_max(static_cast<int32_t>(carry) >> 16), -32768);
#pragma unroll
  for (int q = 0; q < H2; ++q) {
    o.h[q] = pk_max_lo(pre[q], static_cast<uint32_t>(excl));
    // diagonal first, then vertical, else horizontal, from 0/1 flags
    // fd (not diagonal) and fv (not vertical)
    const uint32_t fd = pk_min_u(pk_sub(o.h[q], bd[q]), ps.one);
    const uint32_t fv = pk_min_u(pk_sub(o.h[q], bv[q]), ps.one);
    if (kMulti) {
      // codes of the first best predecessor: kCodeViaK | k
      const uint32_t kd = q == 0 ? ((kk[0] << 16) | (k0 & 0xffffu))
                                 : __builtin_amdgcn_alignbit(kk[q], kk[q - 1], 16);
      const uint32_t dc = pk_add(kd, 0x20002000u);     // diagonal | kCodeViaK | k
      const uint32_t vc = pk_add(kk[q], 0x60006000u);  // vertical | kCodeViaK | k
      const uint32_t cv = pk_mad(fv, pk_sub(0x80008000u, vc), vc);
      o.cd[q] = pk_mad(fd, pk_sub(cv, dc), dc);
    } else {
      // the move kind 2 fd + fv (both halves at once: no carries)
      o.cd[q] = (fd << 1) | fv;
    }
    // computed here, not sunk into the store branch, so that a second row's
    // scan can interleave with this row's codes
    asm volatile("" : "+v"(o.cd[q]));
  }
}

// Stores of one packed row (codes, the row's scores where later rows read
// them) and H[row][L] as the return value.
template <int W, bool kMulti>
__device__ __forceinline__ int32_t packed_store(const PackedOut<W>& o, uint32_t k0, uint32_t pay0,
                                                int lane, int j0, int jq, int qs, bool live, int L,
                                                int32_t gap, int own, int pitch, int16_t* ring,
                                                int32_t* Hg, uint16_t* Cr, int end_lane,
                                                int end_u) {
  constexpr int H2 = W / 2;
  if (live) {
    if (kMulti) {
#pragma unroll
      for (int q = 0; q < W / 4; ++q) {
        *reinterpret_cast<uint2*>(Cr + (j0 - 1) + 4 * q) = make_uint2(o.cd[2 * q], o.cd[2 * q + 1]);
      }
    } else {
      // move kinds, 2 bits per column: pair q's halves (columns 2q, 2q + 1
      // of the strip) go to bits 4q and 4q + 2
      auto pack8 = [&](int g) -> uint32_t {  // pairs 4g..4g+3: 16 bits
        uint32_t x = o.cd[min(4 * g, H2 - 1)];
#pragma unroll
        for (int q = 1; q < 4; ++q) x |= o.cd[min(4 * g + q, H2 - 1)] << (4 * q);
        return (x & 0xffffu) | (x >> 14);
      };
      uint8_t* K = reinterpret_cast<uint8_t*>(Cr) + (j0 - 1) / 4;
      if constexpr (W == 4) {
        const uint32_t x = o.cd[0] | (o.cd[1] << 4);
        *K = static_cast<uint8_t>(x | (x >> 14));
      } else if constexpr (W == 8) {
        *reinterpret_cast<uint16_t*>(K) = static_cast<uint16_t>(pack8(0));
      } else if constexpr (W == 12) {
        const uint32_t v = pack8(0);
        const uint32_t x = o.cd[4] | (o.cd[5] << 4);
        K[0] = static_cast<uint8_t>(v);
        K[1] = static_cast<uint8_t>(v >> 8);
        K[2] = static_cast<uint8_t>(x | (x >> 14));
      } else {
        *reinterpret_cast<uint32_t*>(K) = pack8(0) | (pack8(1) << 16);
      }
    }
    if (own < kSlotGlobal) {
      int16_t* row = ring + own * pitch;
#pragma unroll
      for (int q = 0; q < W / 4; ++q) {
        *reinterpret_cast<uint2*>(row + jq + q * qs) = make_uint2(o.h[2 * q], o.h[2 * q + 1]);
      }
    }
  }
  if (own == kSlotGlobal) {
#pragma unroll
    for (int u = 0; u < W; ++u) {
      if (j0 + u <= L) Hg[j0 + u] = pk_half(o.h[u / 2], u & 1);
    }
  }
  if (lane == 0) {
    if (own < kSlotGlobal) ring[own * pitch + pitch - 1] = static_cast<int16_t>(o.h0);
    if (own == kSlotGlobal) Hg[0] = o.h0;
    if (kMulti) Cr[pitch - 1] = static_cast<uint16_t>(kCodeVert | kCodeViaK | (k0 & 0xffffu));
  }
  // (each pair opaque, so that the selection stays register selects rather
  // than an indexed load from a stack copy of the row)
  uint32_t lastp = o.h[0];
#pragma unroll
  for (int q = 1; q < H2; ++q) {
    uint32_t v = o.h[q];
    asm volatile("" : "+v"(v));
    if (q == (end_u >> 1)) lastp = v;
  }
  return __builtin_amdgcn_readlane(pk_half(lastp, end_u & 1), end_lane) + L * gap;
}

// One DP row in packed pairs: the moves, codes, stores and end-column value
// of finish_row, two columns per instruction. pp holds the predecessor
// scores at the strip's columns and hp0 (low half) the one at column j0 - 1.
// With one predecessor (kMulti false) every diagonal and vertical move names
// it through pay0. With several, pp and hp0 hold the best predecessor score
// of each column and kk, k0 the first predecessor k (in-edge order) that
// reaches it: the diagonal into column j and the vertical into column j - 1
// both take the first predecessor with the best H[p][j - 1], which is the
// predecessor spoa's traceback prefers. Every value is kept as
// G = H - j * gap (the rows in LDS or global memory too): a horizontal move
// keeps G, a vertical one adds gap and a diagonal one adds s - gap, so the
// horizontal chain is a plain prefix max, and every value of a live column
// lies within max_abs * (V + 2L + 2) of zero; the caller takes this path only
// when that fits 16 bits. Columns past L compute unused values (they flow
// only into later columns). Returns H[row][L].
template <int W, bool kMulti>
__device__ __forceinline__ int32_t packed_row(const PackedStrip<W>& ps, const uint32_t* pp,
                                              uint32_t hp0, const uint32_t* kk, uint32_t k0,
                                              uint32_t b, uint32_t pay0, int lane,
                                              int j0, int jq, int qs, bool live, int L,
                                              int32_t gap, int own, int pitch, int16_t* ring,
                                              int32_t* Hg, uint16_t* Cr, int end_lane,
                                              int end_u) {
  PackedOut<W> o;
  packed_compute<W, kMulti>(ps, pp, hp0, kk, k0, b, pay0, lane, gap, o);
  return packed_store<W, kMulti>(o, k0, pay0, lane, j0, jq, qs, live, L, gap, own, pitch, ring,
                                 Hg, Cr, end_lane, end_u);
}

// DP rows of one layer with 16-bit scores, every row in one pass of the
// wave: lane k owns columns 1 + kW .. (k + 1)W (W = 4, 8, 12, or 16 when
// packed, so that 64W >= L), the layer bytes of its strip stay in
// registers, and column 0 heads lane 0's horizontal chain. Same moves,
// codes and end cell as
// dp_layer. Requires in-degree <= 255, 64W + pad <= pitch and 16-bit scores
// over the whole strip. With kPacked every row runs in packed 16-bit pairs
// (packed_row; a row with several predecessors first folds them with a
// packed max and first argmax), which needs max_abs * (V + 2L + 2) in 16
// bits. Without it (W <= 12), rows with one predecessor take no argmax and
// rows with several keep (score, 255 - k) in one integer so that one max
// keeps the first predecessor with the best score (in-edge order, as spoa's
// traceback prefers).
template <int W, bool kPacked>
__device__ __forceinline__ int dp_layer_strip(CParams& P, const Slot& s, uint8_t* lds, int V, int L,
                              int* best_score, bool* tied) {
  static_assert(W % 4 == 0 && W >= 4 && W <= (kPacked ? 16 : 12), "strip width");
  const int lane = lane_id();
  const int pitch = static_cast<int>(P.caps.cols);
  // scores as uniform registers (a select between two parameter fields
  // would otherwise become a per-lane load from the selected address)
  const int32_t gap = uniform(P.gap);
  const int32_t match = uniform(P.match);
  const int32_t mismatch = uniform(P.mismatch);
  int16_t* ring = reinterpret_cast<int16_t*>(lds);
  const uint8_t* seq = lds + align_up(static_cast<uint64_t>(P.ring_rows) * pitch * 2, 16);
  const int j0 = 1 + lane * W;
  const bool live = j0 <= L;
  const int jr = live ? j0 : 1;  // idle lanes read columns 0..W (in bounds)
  // quad-major LDS rows (see from_prev_lane): this lane's first quad and the
  // quad stride; idle lanes read lane 0's quads
  const int n_live = (L + W - 1) / W;
  const int jq = (live ? lane : 0) * 4;
  const int qs = n_live * 4;
  uint32_t sw[W / 4];            // layer bytes of the strip
#pragma unroll
  for (int q = 0; q < W / 4; ++q) sw[q] = *reinterpret_cast<const uint32_t*>(seq + (jr - 1) + 4 * q);
  PackedStrip<W> ps;
  if (kPacked) packed_strip_init<W>(ps, sw, j0, match, mismatch, gap);
  const int end_lane = (L - 1) / W;
  const int end_u = (L - 1) % W;
  int32_t best = kNegInf;
  int best_row = 0;
  int ties = 0;
  const int total_preds = uniform(s.pred_off()[V]);
  // sink rows: record the end-column score and keep the best (in row order)
  auto account = [&](int32_t h, int row, int m) {
    if ((m & kRowMetaSink) != 0) {
      if (lane == 0) {
        s.sink_score()[row] = h;
      }
      if (h > best) {
        best = h;
        best_row = row;
        ties = 0;
      } else if (h == best) {
        ++ties;
      }
    }
  };
  for (int r0 = 1; r0 <= V; r0 += kWave) {
    const int rk = r0 + lane;
    int off_v = 0, end_v = 0, meta_v = 0;
    if (rk <= V) {
      off_v = s.pred_off()[rk - 1];
      end_v = s.pred_off()[rk];
      meta_v = s.row_meta()[rk - 1];
    }
    const int base = __builtin_amdgcn_readlane(off_v, 0);
    const int pr0 = base + lane < total_preds ? s.pred_rows()[base + lane] : 0;
    const int pr1 = base + kWave + lane < total_preds ? s.pred_rows()[base + kWave + lane] : 0;
    // (up to four entries per row of the chunk without a reload)
    const int pr2 =
        base + 2 * kWave + lane < total_preds ? s.pred_rows()[base + 2 * kWave + lane] : 0;
    const int pr3 =
        base + 3 * kWave + lane < total_preds ? s.pred_rows()[base + 3 * kWave + lane] : 0;
    // complete the chunk's loads here: otherwise the row loop would wait for
    // them (and for every code store issued before them) on each row
    wait_stores();
    const int chunk = min(kWave, V - r0 + 1);
    for (int rr = 0; rr < chunk; ++rr) {
      const int r = r0 + rr;
      const int off = __builtin_amdgcn_readlane(off_v, rr);
      const int np = __builtin_amdgcn_readlane(end_v, rr) - off;
      const int meta = __builtin_amdgcn_readlane(meta_v, rr);
      auto pred_of = [&](int o, int k) -> int {
        const int idx = o + k - base;
        if (idx < kWave) return __builtin_amdgcn_readlane(pr0, idx);
        if (idx < 2 * kWave) return __builtin_amdgcn_readlane(pr1, idx - kWave);
        if (idx < 3 * kWave) return __builtin_amdgcn_readlane(pr2, idx - 2 * kWave);
        if (idx < 4 * kWave) return __builtin_amdgcn_readlane(pr3, idx - 3 * kWave);
        const int e = uniform(s.pred_rows()[o + k]);
        wait_stores();
        return e;
      };
      auto pred_at = [&](int k) -> int { return pred_of(off, k); };
      const uint32_t b = static_cast<uint32_t>(meta & 0xff);
      const int own = (meta >> kRowMetaSlotShift) & 31;
      int32_t* Hg = s.H_row(r, meta, pitch);
      uint16_t* Cr = s.code() + static_cast<int64_t>(r) * pitch;
      // match score of strip column u against the row's base
      auto sc = [&](int u) -> int32_t {
        return ((sw[u / 4] >> (8 * (u % 4))) & 0xffu) == b ? match : mismatch;
      };
      int32_t bd[W], bv[W];
      int32_t h_last = 0;
      // the row the common tail below accounts for (the second of a pair)
      int tail_row = r, tail_meta = meta, tail_own = own;
      if (kPacked && np == 1) {
        const int e = pred_at(0);
        const uint32_t delta = static_cast<uint32_t>(r - (e & kPredRowMask));
        const uint32_t pay0 = delta <= kCodePayload ? delta : kCodeViaK;
        // The next row as well when it has one predecessor other than this
        // row and neither row keeps its scores in global memory: the rows
        // are independent, so both loads go out before either row's
        // arithmetic and the two rows' instructions interleave. (This row's
        // LDS slot holds no row the next one reads: a slot is reused only
        // after the last reader of its row.)
        bool paired = false;
        if (W <= 12 && rr + 1 < chunk && own != kSlotGlobal) {  // (16 would spill)
          const int off2 = __builtin_amdgcn_readlane(off_v, rr + 1);
          const int np2 = __builtin_amdgcn_readlane(end_v, rr + 1) - off2;
          const int meta2 = __builtin_amdgcn_readlane(meta_v, rr + 1);
          const int own2 = (meta2 >> kRowMetaSlotShift) & 31;
          if (np2 == 1 && own2 != kSlotGlobal) {
            const int e2 = pred_of(off2, 0);
            if ((e2 & kPredRowMask) != r) {
              paired = true;
              const uint32_t delta2 = static_cast<uint32_t>(r + 1 - (e2 & kPredRowMask));
              const uint32_t pay2 = delta2 <= kCodePayload ? delta2 : kCodeViaK;
              uint32_t pa[W / 2], pb[W / 2];
              const uint32_t ha0 = load_pairs<W>(ring, pitch, s, e, j0, jq, qs, L, gap, ps, pa);
              const uint32_t hb0 = load_pairs<W>(ring, pitch, s, e2, j0, jq, qs, L, gap, ps, pb);
              PackedOut<W> oa, ob;
              packed_compute<W, false>(ps, pa, ha0, pa, 0u, b, pay0, lane, gap, oa);
              packed_compute<W, false>(ps, pb, hb0, pb, 0u, static_cast<uint32_t>(meta2 & 0xff),
                                       pay2, lane, gap, ob);
              const int32_t h_a = packed_store<W, false>(oa, 0u, pay0, lane, j0, jq, qs, live, L, gap,
                                                         own, pitch, ring, Hg, Cr, end_lane,
                                                         end_u);
              h_last = packed_store<W, false>(ob, 0u, pay2, lane, j0, jq, qs, live, L, gap, own2, pitch,
                                              ring, s.H_row(r + 1, meta2, pitch),
                                              s.code() + static_cast<int64_t>(r + 1) * pitch,
                                              end_lane, end_u);
              account(h_a, r, meta);
              tail_row = r + 1;
              tail_meta = meta2;
              tail_own = own2;
              ++rr;
            }
          }
        }
        if (!paired) {
          uint32_t pp[W / 2];
          const uint32_t hp0 = load_pairs<W>(ring, pitch, s, e, j0, jq, qs, L, gap, ps, pp);
          h_last = packed_row<W, false>(ps, pp, hp0, pp, 0u, b, pay0, lane, j0, jq, qs, live, L, gap,
                                        own, pitch, ring, Hg, Cr, end_lane, end_u);
        }
      } else if (kPacked) {
        // best predecessor score of every column and the first predecessor
        // reaching it (a later one replaces it only when strictly better)
        uint32_t mx[W / 2], kx[W / 2];
        uint32_t m0 = load_pairs<W>(ring, pitch, s, pred_at(0), j0, jq, qs, L, gap, ps, mx);
        uint32_t k0 = 0;
#pragma unroll
        for (int q = 0; q < W / 2; ++q) kx[q] = 0;
        for (int k = 1; k < np; ++k) {
          uint32_t pp[W / 2];
          const uint32_t hk = load_pairs<W>(ring, pitch, s, pred_at(k), j0, jq, qs, L, gap, ps, pp);
          const uint32_t ks = static_cast<uint32_t>(k) * 0x10001u;
#pragma unroll
          for (int q = 0; q < W / 2; ++q) {
            const uint32_t m = pk_max(mx[q], pp[q]);
            kx[q] = pk_mad(pk_min_u(pk_sub(m, mx[q]), ps.one), pk_sub(ks, kx[q]), kx[q]);
            mx[q] = m;
          }
          const uint32_t m = pk_max(m0, hk);
          k0 = pk_mad(pk_min_u(pk_sub(m, m0), ps.one), pk_sub(ks, k0), k0);
          m0 = m;
        }
        h_last = packed_row<W, true>(ps, mx, m0, kx, k0, b, 0u, lane, j0, jq, qs, live, L, gap, own, pitch,
                                     ring, Hg, Cr, end_lane, end_u);
      } else if (np == 1) {
        // move-code payload of the single predecessor: its row delta
        const uint32_t delta = static_cast<uint32_t>(r - (pred_at(0) & kPredRowMask));
        const uint32_t pay0 = delta <= kCodePayload ? delta : kCodeViaK;
        int32_t hp[W + 1];
        load_strip<W>(ring, pitch, s, pred_at(0), j0, jq, qs, L, gap, hp);
#pragma unroll
        for (int u = 0; u < W; ++u) {
          bd[u] = hp[u] + sc(u);
          bv[u] = hp[u + 1] + gap;
        }
        h_last = finish_row<W, false>(bd, bv, hp[0] + gap, pay0, lane, j0, jq, qs, live, L, gap, own,
                                      pitch, ring, Hg, Cr, end_lane, end_u);
      } else {
        int32_t c0k = kNegKey;
#pragma unroll
        for (int u = 0; u < W; ++u) {
          bd[u] = kNegKey;
          bv[u] = kNegKey;
        }
        for (int k = 0; k < np; ++k) {
          int32_t hp[W + 1];
          load_strip<W>(ring, pitch, s, pred_at(k), j0, jq, qs, L, gap, hp);
          const int32_t tag = 255 - k;
#pragma unroll
          for (int u = 0; u < W; ++u) {
            bd[u] = max(bd[u], (hp[u] + sc(u)) * 256 | tag);
            bv[u] = max(bv[u], (hp[u + 1] + gap) * 256 | tag);
          }
          c0k = max(c0k, (hp[0] + gap) * 256 | tag);
        }
        h_last = finish_row<W, true>(bd, bv, c0k, 0u, lane, j0, jq, qs, live, L, gap, own, pitch, ring,
                                     Hg, Cr, end_lane, end_u);
      }
      account(h_last, tail_row, tail_meta);
      if (tail_own == kSlotGlobal) {
        wait_stores();
      }
      wave_fence();
    }
  }
  *best_score = best;
  *tied = ties > 0;
  return best_row;
}

// ---- banded rows: the opt-in approximation of racon -b ----
//
// Each DP row r (1..V) computes only 256 columns, 4 per lane: columns
// lo(r) + 1 .. lo(r) + 256, where lo(r) follows the diagonal of the V x L
// matrix (centred on column r L / V, a multiple of 4, inside [0, L + 3]).
// lo never decreases with r, so a predecessor p < r holds row r's lane k in
// its lane k + (lo(r) - lo(p)) / 4. Values outside a row's band read as
// kBandNeg; a window whose traceback would leave the band is rerun unbanded
// (kPoaBandMiss). Used for layers longer than 256 only, and only in the
// packed 16-bit form (kBandNeg sits below every value that form holds).
constexpr int kBandW = 4;
constexpr int kBand = kWave * kBandW;  // 256 columns
constexpr int32_t kBandNeg = -32768 + 64;

__device__ __forceinline__ int band_lo(int r, int V, int L) {
  // (single-precision r * (L / V): monotone in r, the same in the DP and the
  // traceback)
  const int c = static_cast<int>(static_cast<float>(r) * (static_cast<float>(L) /
                                                           static_cast<float>(V))) - kBand / 2;
  const int top = (L - kBand + 3) & ~3;
  return min(max(c, 0) & ~3, top);
}

// Whether cell (i, j) lies in row i's band (row 0, the virtual start row,
// holds every column).
__device__ __forceinline__ bool in_band(int i, int j, int V, int L) {
  if (i <= 0) return true;
  const int lo = band_lo(i, V, L);
  return (j > lo && j <= lo + kBand) || (j == 0 && lo == 0);
}

// Predecessor row p (entry e) at the lanes of a band row whose band starts
// after column lo: pp[q] = (G[p][j0 + 2q], G[p][j0 + 2q + 1]) with
// j0 = lo + 1 + 4 lane, kBandNeg outside p's band; returns G[p][j0 - 1] in
// the low half.
__device__ __forceinline__ uint32_t load_band(const int16_t* ring, int pitch, const Slot& s, int e,
                                              int lo, int V, int L, int lane, uint32_t* pp) {
  const int p = e & kPredRowMask;
  const int sl = static_cast<unsigned>(e) >> kPredRowBits;
  const uint32_t neg = pk_splat(kBandNeg);
  if (p == 0) {  // the virtual start row: G = 0 everywhere
    pp[0] = pp[1] = 0u;
    return 0u;
  }
  const int lo_p = band_lo(p, V, L);
  const int shift = max(lo - lo_p, 0) >> 2;  // p < r: lo_p <= lo
  const int src = lane + shift;
  if (sl < kSlotGlobal) {
    const int16_t* row = ring + sl * pitch;
    const bool in = src < kWave;
    const uint2 v = *reinterpret_cast<const uint2*>(row + 4 * (in ? src : 0));
    pp[0] = in ? v.x : neg;
    pp[1] = in ? v.y : neg;
    // column lo of p: the last one of p's lane shift - 1, or p's column 0
    uint32_t c0 = neg;
    if (shift > 0 && shift <= kWave) {
      c0 = static_cast<uint16_t>(row[4 * (shift - 1) + 3]);
    } else if (shift == 0 && lo == 0) {
      c0 = static_cast<uint16_t>(row[pitch - 1]);
    }
    const uint32_t prev = from_prev_lane(pp[1]) >> 16;
    return lane == 0 ? c0 : prev;
  }
  // a row kept in the slot's global matrix (int32, absolute columns)
  const int32_t* row = s.H_row(p, pitch);
  int32_t hv[kBandW + 1];
#pragma unroll
  for (int u = 0; u <= kBandW; ++u) {
    const int j = lo + 4 * lane + u;
    const bool valid = (j > lo_p && j <= lo_p + kBand && j <= L) || (j == 0 && lo_p == 0);
    hv[u] = valid ? row[j] : kBandNeg;
  }
  wait_stores();
  pp[0] = (static_cast<uint32_t>(hv[1]) & 0xffffu) | (static_cast<uint32_t>(hv[2]) << 16);
  pp[1] = (static_cast<uint32_t>(hv[3]) & 0xffffu) | (static_cast<uint32_t>(hv[4]) << 16);
  return static_cast<uint32_t>(hv[0]) & 0xffffu;
}

// One band row: packed_compute with the lane-0 chain starting from column 0
// only when the band starts there, values kept at or above kBandNeg, then
// the stores of packed_store at the band's columns. Returns H[row][L], or
// kNegInf when column L is outside the band.
template <bool kMulti>
__device__ __forceinline__ int32_t band_row(const PackedStrip<kBandW>& ps, const uint32_t* pp,
                                            uint32_t hp0, const uint32_t* kk, uint32_t k0,
                                            uint32_t b, uint32_t pay0, int lane, int lo, int L,
                                            int32_t gap, int own, int pitch, int16_t* ring,
                                            int32_t* Hg, uint16_t* Cr) {
  constexpr int H2 = kBandW / 2;
  const uint32_t bb = b * 0x10001u;
  const uint32_t neg = pk_splat(kBandNeg);
  const int32_t h0 = pk_half(hp0 + static_cast<uint32_t>(gap), 0);  // column 0 when lo == 0
  uint32_t carry = lane == 0 ? (lo == 0 ? pk_splat(h0) : neg) : 0x80008000u;
  uint32_t bd[H2], bv[H2], pre[H2], h[H2], cd[H2];
#pragma unroll
  for (int q = 0; q < H2; ++q) {
    const uint32_t d = q == 0 ? ((pp[0] << 16) | (hp0 & 0xffffu))
                              : __builtin_amdgcn_alignbit(pp[q], pp[q - 1], 16);
    const uint32_t mis = pk_min_u(ps.base[q] ^ bb, ps.one);
    bd[q] = pk_mad(mis, ps.diff, pk_add(d, ps.match_gap));
    bv[q] = pk_add(pp[q], ps.gap);
    carry = pk_max_hi(pk_scan(pk_max(bd[q], bv[q])), carry);
    pre[q] = carry;
  }
  const int32_t excl = wave_shift_right(wave_prefix_max(static_cast<int32_t>(carry) >> 16), -32768);
#pragma unroll
  for (int q = 0; q < H2; ++q) {
    h[q] = pk_max(pk_max_lo(pre[q], static_cast<uint32_t>(excl)), neg);
    const uint32_t fd = pk_min_u(pk_sub(h[q], bd[q]), ps.one);
    const uint32_t fv = pk_min_u(pk_sub(h[q], bv[q]), ps.one);
    if (kMulti) {
      const uint32_t kd = q == 0 ? ((kk[0] << 16) | (k0 & 0xffffu))
                                 : __builtin_amdgcn_alignbit(kk[q], kk[q - 1], 16);
      const uint32_t dc = pk_add(kd, 0x20002000u);     // diagonal | kCodeViaK | k
      const uint32_t vc = pk_add(kk[q], 0x60006000u);  // vertical | kCodeViaK | k
      const uint32_t cv = pk_mad(fv, pk_sub(0x80008000u, vc), vc);
      cd[q] = pk_mad(fd, pk_sub(cv, dc), dc);
    } else {
      cd[q] = (fd << 1) | fv;  // the move kind, as in packed_compute
    }
  }
  const int j0 = lo + 1 + kBandW * lane;
  if (kMulti) {
    *reinterpret_cast<uint2*>(Cr + (j0 - 1)) = make_uint2(cd[0], cd[1]);
  } else {
    // 2-bit kinds of the lane's 4 columns: byte (j0 - 1) / 4 of the row
    const uint32_t x = cd[0] | (cd[1] << 4);
    reinterpret_cast<uint8_t*>(Cr)[(j0 - 1) >> 2] = static_cast<uint8_t>(x | (x >> 14));
  }
  if (own < kSlotGlobal) {
    *reinterpret_cast<uint2*>(ring + own * pitch + 4 * lane) = make_uint2(h[0], h[1]);
  } else if (own == kSlotGlobal) {
#pragma unroll
    for (int u = 0; u < kBandW; ++u) {
      if (j0 + u <= L) Hg[j0 + u] = pk_half(h[u / 2], u & 1);
    }
  }
  if (lane == 0) {
    const int16_t c0 = static_cast<int16_t>(lo == 0 ? h0 : kBandNeg);
    if (own < kSlotGlobal) ring[own * pitch + pitch - 1] = c0;
    if (own == kSlotGlobal) Hg[0] = c0;
    if (kMulti && lo == 0) {
      Cr[pitch - 1] = static_cast<uint16_t>(kCodeVert | kCodeViaK | (k0 & 0xffffu));
    }
  }
  const int e = L - 1 - lo;  // column L in the band: lane e / 4, column e % 4
  if (e < 0 || e >= kBand) return kNegInf;
  const int eu = e % kBandW;
  uint32_t v = eu < 2 ? h[0] : h[1];
  return __builtin_amdgcn_readlane(pk_half(v, eu & 1), e / kBandW) + L * gap;
}

// DP rows of one layer in band rows (layers longer than 256, packed form).
// Same codes, sink scores and end cell as dp_layer_strip inside the bands.
__device__ __forceinline__ int dp_layer_band(CParams& P, const Slot& s, uint8_t* lds, int V, int L,
                                             int* best_score, bool* tied) {
  const int lane = lane_id();
  const int pitch = static_cast<int>(P.caps.cols);
  const int32_t gap = uniform(P.gap);
  const int32_t match = uniform(P.match);
  const int32_t mismatch = uniform(P.mismatch);
  int16_t* ring = reinterpret_cast<int16_t*>(lds);
  const uint8_t* seq = lds + align_up(static_cast<uint64_t>(P.ring_rows) * pitch * 2, 16);
  PackedStrip<kBandW> ps;
  {
    uint32_t sw0 = 0;
    packed_strip_init<kBandW>(ps, &sw0, 1, match, mismatch, gap);
  }
  int32_t best = kNegInf;
  int best_row = 0;
  int ties = 0;
  const int total_preds = uniform(s.pred_off()[V]);
  for (int r0 = 1; r0 <= V; r0 += kWave) {
    const int rk = r0 + lane;
    int off_v = 0, end_v = 0, meta_v = 0;
    if (rk <= V) {
      off_v = s.pred_off()[rk - 1];
      end_v = s.pred_off()[rk];
      meta_v = s.row_meta()[rk - 1];
    }
    const int base = __builtin_amdgcn_readlane(off_v, 0);
    const int pr0 = base + lane < total_preds ? s.pred_rows()[base + lane] : 0;
    const int pr1 = base + kWave + lane < total_preds ? s.pred_rows()[base + kWave + lane] : 0;
    wait_stores();
    const int chunk = min(kWave, V - r0 + 1);
    for (int rr = 0; rr < chunk; ++rr) {
      const int r = r0 + rr;
      const int off = __builtin_amdgcn_readlane(off_v, rr);
      const int np = __builtin_amdgcn_readlane(end_v, rr) - off;
      const int meta = __builtin_amdgcn_readlane(meta_v, rr);
      auto pred_at = [&](int k) -> int {
        const int idx = off + k - base;
        if (idx < kWave) return __builtin_amdgcn_readlane(pr0, idx);
        if (idx < 2 * kWave) return __builtin_amdgcn_readlane(pr1, idx - kWave);
        const int e = uniform(s.pred_rows()[off + k]);
        wait_stores();
        return e;
      };
      const int lo = uniform(band_lo(r, V, L));
      // the row's layer bytes: columns lo + 1 + 4 lane .. + 3 (lo + 4 lane
      // + 3 <= L + 2 < pitch: inside the layer's LDS bytes and padding)
      const uint32_t sw = *reinterpret_cast<const uint32_t*>(seq + lo + kBandW * lane);
      ps.base[0] = __builtin_amdgcn_perm(sw, sw, 0x0c010c00u);
      ps.base[1] = __builtin_amdgcn_perm(sw, sw, 0x0c030c02u);
      const uint32_t b = static_cast<uint32_t>(meta & 0xff);
      const int own = (meta >> kRowMetaSlotShift) & 31;
      int32_t* Hg = s.H_row(r, meta, pitch);
      uint16_t* Cr = s.code() + static_cast<int64_t>(r) * pitch;
      int32_t h_last;
      if (np == 1) {
        const int e = pred_at(0);
        const uint32_t delta = static_cast<uint32_t>(r - (e & kPredRowMask));
        const uint32_t pay0 = delta <= kCodePayload ? delta : kCodeViaK;
        uint32_t pp[2];
        const uint32_t hp0 = load_band(ring, pitch, s, e, lo, V, L, lane, pp);
        h_last = band_row<false>(ps, pp, hp0, pp, 0u, b, pay0, lane, lo, L, gap, own, pitch, ring,
                                 Hg, Cr);
      } else {
        uint32_t mx[2], kx[2] = {0u, 0u};
        uint32_t m0 = load_band(ring, pitch, s, pred_at(0), lo, V, L, lane, mx);
        uint32_t k0 = 0;
        for (int k = 1; k < np; ++k) {
          uint32_t pp[2];
          const uint32_t hk = load_band(ring, pitch, s, pred_at(k), lo, V, L, lane, pp);
          const uint32_t ks = static_cast<uint32_t>(k) * 0x10001u;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const uint32_t m = pk_max(mx[q], pp[q]);
            kx[q] = pk_mad(pk_min_u(pk_sub(m, mx[q]), ps.one), pk_sub(ks, kx[q]), kx[q]);
            mx[q] = m;
          }
          const uint32_t m = pk_max(m0, hk);
          k0 = pk_mad(pk_min_u(pk_sub(m, m0), ps.one), pk_sub(ks, k0), k0);
          m0 = m;
        }
        h_last = band_row<true>(ps, mx, m0, kx, k0, b, 0u, lane, lo, L, gap, own, pitch, ring,
                                Hg, Cr);
      }
      if ((meta & kRowMetaSink) != 0) {
        if (lane == 0) s.sink_score()[r] = h_last;
        if (h_last > best) {
          best = h_last;
          best_row = r;
          ties = 0;
        } else if (h_last == best) {
          ++ties;
        }
      }
      if (own == kSlotGlobal) wait_stores();
      wave_fence();
    }
  }
  *best_score = best;
  *tied = ties > 0 && best > kNegInf;
  return best == kNegInf ? 0 : best_row;
}

// Uniform copy of a pointer (both halves read from the first active lane).
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (T*)((static_cast<uint64_t>(hi) << 32) | lo);
}

enum DpKind : int {
  kDpStrip4,
  kDpStrip8,
  kDpStrip12,
  kDpPacked4,
  kDpPacked8,
  kDpPacked12,
  kDpPacked16,
  kDpGeneric16,
  kDpGeneric32,
  kDpBand4
};

// The DP of one layer as an out-of-line call: its loops get a register
// allocation of their own instead of sharing one with the rest of the window
// loop. Arguments arrive in vector registers and are made uniform again on
// entry. Returns the end-cell row, its score and whether sinks tied.
typedef __attribute__((address_space(3))) uint8_t LByte;

template <int kKind>
__device__ __noinline__ uint64_t dp_layer_call(CParams* params, GByte* base, CLayout* layout,
                                               uint32_t lds_addr, int V, int L) {
  CParams& P = *uniform_ptr(params);
  Slot s = carve(uniform_ptr(base), *uniform_ptr(layout));
  V = uniform(V);
  L = uniform(L);
  // the wave's LDS as a local-memory pointer from the caller (no lookup of
  // the dynamic LDS offset inside the loops)
  uint8_t* lds = (uint8_t*)(LByte*)(size_t)__builtin_amdgcn_readfirstlane(lds_addr);
  int best_score = 0;
  bool tied = false;
  int row;
  if (kKind == kDpStrip4) {
    row = dp_layer_strip<4, false>(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpStrip8) {
    row = dp_layer_strip<8, false>(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpStrip12) {
    row = dp_layer_strip<12, false>(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpPacked4) {
    row = dp_layer_strip<4, true>(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpPacked8) {
    row = dp_layer_strip<8, true>(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpPacked12) {
    row = dp_layer_strip<12, true>(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpPacked16) {
    row = dp_layer_strip<16, true>(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpBand4) {
    row = dp_layer_band(P, s, lds, V, L, &best_score, &tied);
  } else if (kKind == kDpGeneric16) {
    row = dp_layer<int16_t>(P, s, lds, V, L, &best_score, &tied);
  } else {
    row = dp_layer<int32_t>(P, s, lds, V, L, &best_score, &tied);
  }
  return (static_cast<uint64_t>(static_cast<uint32_t>(best_score)) << 32) |
         (static_cast<uint32_t>(row) & 0x7fffffffu) | (tied ? 0x80000000u : 0u);
}

// Graph insertion of one layer (spoa add_alignment), position by position,
// lane-uniform, from the aligned node of every layer position (pos_node,
// -1 for an insertion). New nodes record their anchor rank for the
// incremental order. (Rare: two path nodes share an aligned set. Out of
// line; returns the slot with its updated counters.)
__device__ __noinline__ Slot insert_sequential(Slot s, CCaps& caps, int L, bool* order_ok_out,
                                              int* n_path_out) {
  bool order_ok = true;
  bool aligned_full = false;
  int n_path = 0;
  const int32_t* pos_node = s.path();
  const int before = s.n_nodes;
  int head = -1;
  int prev_w = 0;
  int head_anchor = -1;  // anchor of the previous path node when it is new
  for (int pos = 0; pos < L && !s.overflow; ++pos) {
    const int node = uniform(pos_node[pos]);
    const uint8_t letter = s.seq()[pos];
    int nv;
    if (node == -1) {
      nv = add_node(s, caps, letter);
    } else if (s.base()[node] == letter) {
      nv = node;
    } else {
      int same = -1;
      const int cnt = s.n_aligned()[node];
      for (int k = 0; k < cnt; ++k) {
        const int a = uniform(s.aligned()[node * kAligned + k]);
        if (s.base()[a] == letter) {
          same = a;
          break;
        }
      }
      if (same == -1) {
        nv = add_node(s, caps, letter);
        if (!s.overflow && cnt + 1 > static_cast<int>(kAligned)) {
          aligned_full = true;
          s.overflow = true;
        }
        if (s.overflow) break;
        for (int k = 0; k < cnt; ++k) {
          const int a = uniform(s.aligned()[node * kAligned + k]);
          const int ca = s.n_aligned()[a];
          s.aligned()[nv * kAligned + k] = a;
          s.aligned()[a * kAligned + ca] = nv;
          s.n_aligned()[a] = static_cast<uint8_t>(ca + 1);
        }
        s.aligned()[nv * kAligned + cnt] = node;
        s.n_aligned()[nv] = static_cast<uint8_t>(cnt + 1);
        s.aligned()[node * kAligned + cnt] = nv;
        s.n_aligned()[node] = static_cast<uint8_t>(cnt + 1);
      } else {
        nv = same;
      }
    }
    if (nv >= before) {
      // a new node: placed after its path predecessor's aligned set
      // (mismatch nodes: after the aligned set they join)
      int anchor;
      if (node != -1) {
        anchor = s.node_rank()[node];
        const int cnt = s.n_aligned()[node];
        for (int k = 0; k < cnt; ++k) {
          const int a = uniform(s.aligned()[node * kAligned + k]);
          if (a < before) anchor = max(anchor, s.node_rank()[a]);
        }
      } else if (head == -1) {
        anchor = -1;
      } else if (head >= before) {
        anchor = head_anchor;
      } else {
        anchor = s.node_rank()[head];
        const int cnt = s.n_aligned()[head];
        for (int k = 0; k < cnt; ++k) {
          const int a = uniform(s.aligned()[head * kAligned + k]);
          if (a < before) anchor = max(anchor, s.node_rank()[a]);
        }
      }
      const int k_new = nv - before;
      if (k_new > 0 && anchor < s.anchor()[k_new - 1]) {
        order_ok = false;
      }
      s.anchor()[k_new] = anchor;
      head_anchor = anchor;
    }
    s.lpath()[n_path++] = nv;
    if (head != -1) {
      add_edge(s, caps, head, nv, prev_w + s.qual()[pos]);
    }
    head = nv;
    prev_w = s.qual()[pos];
  }
  s.n_seqs = s.n_seqs + 1;
  *order_ok_out = order_ok;
  *n_path_out = aligned_full ? -1 : n_path;  // -1: an aligned set was full
  return s;
}

// Graph insertion of one layer, one lane per layer position. A global
// alignment covers every layer position once (positions 0..L-1 in order), so
// the path is the whole layer and its graph nodes are distinct. When no two
// of them share an aligned set, every position reads and writes only its own
// node, aligned set, in-edge list and out-edge list, and the result equals
// the position-by-position insertion: node ids and edge ids are handed out
// in position order with wave prefix counts, and the anchor of an insertion
// (the rank bound of the last non-insertion position before it) is a wave
// prefix scan. The same independence lets each pass load what kU chunks of
// 64 positions read before any of their writes, so those loads are in
// flight together. Status 1 when two path nodes share an aligned set
// (nothing in the graph has changed then), 0 when inserted (with the new
// node and edge counts), or a negative status. Out of line, for a register
// allocation of its own.
struct InsertResult {
  int32_t status;
  int32_t n_nodes;
  int32_t n_edges;
  int32_t order_ok;
};

__device__ __noinline__ InsertResult insert_parallel(GByte* slot_base, CLayout* layout,
                                                     uint32_t lseq_addr, int L, int n_nodes,
                                                     int n_edges, int n_seqs, int cap_nodes,
                                                     int cap_edges) {
  constexpr int kU = 4;   // chunks per round in the edge and label passes
  constexpr int kUA = 2;  // in the passes that hold aligned sets
  constexpr int kA = static_cast<int>(kAligned);
  const Slot s = carve(uniform_ptr(slot_base), *uniform_ptr(layout));
  const uint8_t* lseq = (const uint8_t*)(LByte*)(size_t)uniform(static_cast<int>(lseq_addr));
  L = uniform(L);
  n_nodes = uniform(n_nodes);
  n_edges = uniform(n_edges);
  n_seqs = uniform(n_seqs);
  cap_nodes = uniform(cap_nodes);
  cap_edges = uniform(cap_edges);
  InsertResult res;
  res.status = 0;
  res.n_nodes = n_nodes;
  res.n_edges = n_edges;
  res.order_ok = 1;
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const int32_t* pnode = s.path();  // aligned graph node of each layer position (-1: none)
  int32_t* pnv = s.lpath();         // graph node of each layer position after insertion
  uint8_t* stamp = s.in_sub();      // 2 on the path's nodes (the DP no longer needs it)
  uint8_t* base = s.base();
  uint8_t* n_aligned = s.n_aligned();
  int32_t* aligned = s.aligned();
  int32_t* node_rank = s.node_rank();
  for (int p = lane; p < L; p += kWave) {
    const int v = pnode[p];
    if (v != -1) stamp[v] = 2;
  }
  wave_sync();
  // existing node per position (match, or an aligned node with the letter)
  bool conflict = false, full = false;
  int n_new = 0;
  for (int p0 = 0; p0 < L; p0 += kUA * kWave) {
    int v[kUA], cnt[kUA], nv[kUA], letter[kUA], al[kUA][kA];
#pragma unroll
    for (int i = 0; i < kUA; ++i) {
      const int p = p0 + i * kWave + lane;
      v[i] = p < L ? pnode[p] : -1;
      letter[i] = p < L ? static_cast<int>(lseq[p]) : 0;
    }
#pragma unroll
    for (int i = 0; i < kUA; ++i) {
      cnt[i] = v[i] != -1 ? static_cast<int>(n_aligned[v[i]]) : 0;
      nv[i] = v[i] != -1 && static_cast<int>(base[v[i]]) == letter[i] ? v[i] : -1;
#pragma unroll
      for (int k = 0; k < kA; ++k) al[i][k] = v[i] != -1 ? aligned[v[i] * kA + k] : 0;
    }
#pragma unroll
    for (int i = 0; i < kUA; ++i) {
      int st[kA], ba[kA];
#pragma unroll
      for (int k = 0; k < kA; ++k) {
        st[k] = k < cnt[i] ? static_cast<int>(stamp[al[i][k]]) : 0;
        ba[k] = k < cnt[i] ? static_cast<int>(base[al[i][k]]) : -1;
      }
#pragma unroll
      for (int k = 0; k < kA; ++k) {
        conflict |= st[k] == 2;
        if (nv[i] == -1 && ba[k] == letter[i]) nv[i] = al[i][k];
      }
      full |= v[i] != -1 && nv[i] == -1 && cnt[i] + 1 > kA;
      const int p = p0 + i * kWave + lane;
      if (p < L) pnv[p] = nv[i];
      n_new += __popcll(__ballot(p < L && nv[i] == -1));
    }
  }
  if (__ballot(conflict) != 0) {
    res.status = 1;
    return res;
  }
  if (__ballot(full) != 0) {
    res.status = kPoaAlignedFull;  // a column with more distinct bytes than kAligned + 1
    return res;
  }
  if (n_nodes + n_new > cap_nodes) {
    res.status = kPoaNodeOverflow;
    return res;
  }
  wave_sync();
  // new nodes (ids in position order), aligned sets, anchors
  const int before = n_nodes;
  int next_id = before;
  int carry = -1;  // rank bound of the last non-insertion position so far
  for (int p0 = 0; p0 < L; p0 += kUA * kWave) {
    int v[kUA], nv[kUA], cnt[kUA], bound[kUA], al[kUA][kA];
#pragma unroll
    for (int i = 0; i < kUA; ++i) {
      const int p = p0 + i * kWave + lane;
      v[i] = p < L ? pnode[p] : -1;
      nv[i] = p < L ? pnv[p] : 0;
    }
#pragma unroll
    for (int i = 0; i < kUA; ++i) {
      cnt[i] = v[i] != -1 ? static_cast<int>(n_aligned[v[i]]) : 0;
      bound[i] = v[i] != -1 ? node_rank[v[i]] : -1;
#pragma unroll
      for (int k = 0; k < kA; ++k) al[i][k] = v[i] != -1 ? aligned[v[i] * kA + k] : 0;
    }
    int nal[kUA][kA];  // aligned-set sizes of the members (read before any write)
#pragma unroll
    for (int i = 0; i < kUA; ++i) {
      int rk[kA];
#pragma unroll
      for (int k = 0; k < kA; ++k) {
        rk[k] = k < cnt[i] ? node_rank[al[i][k]] : -1;
        nal[i][k] = k < cnt[i] ? static_cast<int>(n_aligned[al[i][k]]) : 0;
      }
#pragma unroll
      for (int k = 0; k < kA; ++k) bound[i] = max(bound[i], rk[k]);  // highest rank in v's set
    }
#pragma unroll
    for (int i = 0; i < kUA; ++i) {
      const int p = p0 + i * kWave + lane;
      const bool in = p < L;
      const bool isnew = in && nv[i] == -1;
      const unsigned long long m = __ballot(isnew);
      if (isnew) nv[i] = next_id + __popcll(m & below);
      next_id += __popcll(m);
      const int incl = wave_prefix_max(v[i] != -1 ? lane : -1);
      const int excl = wave_shift_right(incl, -1);
      const int from_lane = __shfl(bound[i], max(excl, 0), kWave);
      const int prev_bound = excl >= 0 ? from_lane : carry;
      if (isnew) {
        const int id = nv[i];
        s.anchor()[id - before] = v[i] != -1 ? bound[i] : prev_bound;
        base[id] = lseq[p];
        s.in_head()[id] = -1;
        s.out_head()[id] = -1;
        s.in_tail()[id] = -1;
        s.out_tail()[id] = -1;
        n_aligned[id] = 0;
        s.cov()[id] = 0;
        s.last_label()[id] = -1;
        if (v[i] != -1) {
          // the mismatch node joins v's aligned set
          const int c = cnt[i];
#pragma unroll
          for (int k = 0; k < kA; ++k) {
            if (k < c) {
              const int a = al[i][k];
              const int ca = nal[i][k];
              aligned[id * kA + k] = a;
              aligned[a * kA + ca] = id;
              n_aligned[a] = static_cast<uint8_t>(ca + 1);
            }
          }
          aligned[id * kA + c] = v[i];
          n_aligned[id] = static_cast<uint8_t>(c + 1);
          aligned[v[i] * kA + c] = id;
          n_aligned[v[i]] = static_cast<uint8_t>(c + 1);
        }
      }
      if (in) pnv[p] = nv[i];
      const int last = __builtin_amdgcn_readlane(incl, kWave - 1);
      if (last >= 0) carry = __builtin_amdgcn_readlane(bound[i], last);
    }
  }
  res.n_nodes = next_id;
  wave_sync();
  // anchors of the new nodes must be non-decreasing for the merged order
  bool bad = false;
  for (int k = 1 + lane; k < n_new; k += kWave) {
    bad |= s.anchor()[k] < s.anchor()[k - 1];
  }
  res.order_ok = __ballot(bad) == 0 ? 1 : 0;
  // edges between consecutive positions: weight w[p-1] + w[p]
  int32_t* e_tail = s.e_tail();
  int32_t* e_head = s.e_head();
  int32_t* e_weight = s.e_weight();
  int32_t* e_next_in = s.e_next_in();
  int32_t* e_next_out = s.e_next_out();
  int32_t* out_head = s.out_head();
  int32_t* out_tail = s.out_tail();
  int32_t* in_head = s.in_head();
  int32_t* in_tail = s.in_tail();
  for (int p0 = 1; p0 < L; p0 += kU * kWave) {
    int tail[kU], head[kU], w[kU], found[kU], e[kU];
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      const int p = p0 + i * kWave + lane;
      const bool in = p < L;
      tail[i] = in ? pnv[p - 1] : 0;
      head[i] = in ? pnv[p] : 0;
      w[i] = in ? s.qual()[p - 1] + s.qual()[p] : 0;
      found[i] = -1;
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      const int p = p0 + i * kWave + lane;
      e[i] = p < L ? out_head[tail[i]] : -1;
    }
    // out-edge lists of the kU tails in lockstep
    for (;;) {
      bool act = false;
#pragma unroll
      for (int i = 0; i < kU; ++i) act = act || e[i] != -1;
      if (__ballot(act) == 0) break;
      int h[kU], nx[kU];
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        h[i] = -1;
        nx[i] = -1;
        if (e[i] != -1) {
          h[i] = e_head[e[i]];
          nx[i] = e_next_out[e[i]];
        }
      }
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        if (e[i] != -1 && h[i] == head[i]) {
          found[i] = e[i];
          nx[i] = -1;
        }
        e[i] = nx[i];
      }
    }
    int ot[kU], it[kU];
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      const int p = p0 + i * kWave + lane;
      const bool need = p < L && found[i] == -1;
      ot[i] = need ? out_tail[tail[i]] : -1;
      it[i] = need ? in_tail[head[i]] : -1;
      if (found[i] != -1) e_weight[found[i]] = e_weight[found[i]] + w[i];
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      const int p = p0 + i * kWave + lane;
      const bool need = p < L && found[i] == -1;
      const unsigned long long m = __ballot(need);
      if (n_edges + __popcll(m) > cap_edges) {
        res.status = kPoaNodeOverflow;
        return res;
      }
      if (need) {
        const int ne = n_edges + __popcll(m & below);
        e_tail[ne] = tail[i];
        e_head[ne] = head[i];
        e_weight[ne] = w[i];
        e_next_in[ne] = -1;
        e_next_out[ne] = -1;
        if (ot[i] == -1) {
          out_head[tail[i]] = ne;
        } else {
          e_next_out[ot[i]] = ne;
        }
        out_tail[tail[i]] = ne;
        if (it[i] == -1) {
          in_head[head[i]] = ne;
        } else {
          e_next_in[it[i]] = ne;
        }
        in_tail[head[i]] = ne;
      }
      n_edges += __popcll(m);
    }
  }
  res.n_edges = n_edges;
  // every node of the layer has an incident edge of this sequence when L > 1
  if (L > 1) {
    const int label = n_seqs;
    int32_t* last_label = s.last_label();
    int32_t* cov = s.cov();
    for (int p0 = 0; p0 < L; p0 += kU * kWave) {
      int v[kU], ll[kU], cv[kU];
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        const int p = p0 + i * kWave + lane;
        v[i] = p < L ? pnv[p] : -1;
      }
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        ll[i] = v[i] != -1 ? last_label[v[i]] : label;
        cv[i] = v[i] != -1 ? cov[v[i]] : 0;
      }
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        if (ll[i] != label) {
          last_label[v[i]] = label;
          cov[v[i]] = cv[i] + 1;
        }
      }
    }
  }
  wave_sync();
  return res;
}

// Inclusive prefix sum over the 64 lanes (DPP, as wave_prefix_max; 0 is
// the identity, so each move folds into one v_add_u32 with a DPP source).
__device__ __forceinline__ int32_t wave_prefix_sum(int32_t x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
  return x;
}

// DP rows of a layer (row r + 1 is node sub_order[r]): predecessor rows in
// in-edge order (row 0, the virtual start, for rows without one), sink flags,
// and an LDS slot for every row that a later row reads: the lowest slot whose
// previous row is no longer needed (lane k < n_slots holds the last row that
// reads slot k). With kFull every node is in the layer's subgraph and the
// membership checks are skipped. Returns the largest predecessor count, or
// -1 when a row needs a global score row and the slot's pool has none free.
// Each lane takes kU rows per round (64 kU rows) and walks their edge lists
// in lockstep, one edge of every list per step, so that the loads of the
// kU lists are in flight together; the walk keeps the first kP predecessor
// rows of each row, so only rows with more walk their in-edges again. The
// slot of every row is also kept in LDS (one byte per row, in the DP ring,
// which is idle here) for the predecessor entries, which then need no
// reload of the row table just written. Out of line, like the DP, for a
// register allocation of its own.
template <bool kFull>
__device__ __noinline__ int build_rows(GByte* slot_base, CLayout* layout, int V, int n_slots,
                                       uint32_t lds_addr, int lds_bytes) {
  constexpr int kU = 3;  // rows per lane per round
  constexpr int kP = 6;  // predecessor rows kept in registers (99.9% of rows have <= 6)
  const Slot s = carve(uniform_ptr(slot_base), *uniform_ptr(layout));
  V = uniform(V);
  n_slots = uniform(n_slots);
  const bool slots_in_lds = V <= uniform(lds_bytes);
  LByte* lslot = (LByte*)(size_t)__builtin_amdgcn_readfirstlane(lds_addr);
  const int lane = lane_id();
  const int32_t* sub_order = s.sub_order();
  const int32_t* in_head = s.in_head();
  const int32_t* out_head = s.out_head();
  const int32_t* e_next_in = s.e_next_in();
  const int32_t* e_next_out = s.e_next_out();
  const int32_t* e_tail = s.e_tail();
  const int32_t* e_head = s.e_head();
  const uint8_t* in_sub = s.in_sub();
  const uint8_t* base = s.base();
  int32_t* row_of = s.row_of();
  int32_t* row_meta = s.row_meta();
  int32_t* pred_off = s.pred_off();
  int32_t* pred_rows = s.pred_rows();
  uint64_t* row_pred4 = s.row_pred4();
  auto included = [&](int v) { return kFull || in_sub[v] != 0; };
  // predecessor entry of DP row pr: the row and the slot it lives in
  auto pred_entry = [&](int pr) {
    const int sl = slots_in_lds ? static_cast<int>(lslot[pr - 1])
                                : (row_meta[pr - 1] >> kRowMetaSlotShift) & 31;
    return pr | (sl << kPredRowBits);
  };
  int run = 0;
  int max_np = 0;
  int busy = lane < n_slots ? 0 : 0x7fffffff;
  // pooled global score rows: lane k holds the last reader of entry k
  // (pool0) and of entry 64 + k (pool1); entries past the pool never free
  const int hpool = uniform(static_cast<int>(s.L->hpool));
  int pool0 = lane < hpool ? 0 : 0x7fffffff;
  int pool1 = lane + kWave < hpool ? 0 : 0x7fffffff;
  bool pool_full = false;
  for (int r0 = 0; r0 < V; r0 += kU * kWave) {
    int v[kU], ei[kU], eo[kU], np[kU], oc[kU], lu[kU], meta[kU], off[kU];
    int pr[kU][kP];
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      const int r = r0 + i * kWave + lane;
      v[i] = r < V ? sub_order[r] : -1;
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      ei[i] = v[i] >= 0 ? in_head[v[i]] : -1;
      eo[i] = v[i] >= 0 ? out_head[v[i]] : -1;
      meta[i] = v[i] >= 0 ? static_cast<int>(base[v[i]]) : 0;
      np[i] = 0;
      oc[i] = 0;
      lu[i] = 0;
#pragma unroll
      for (int j = 0; j < kP; ++j) pr[i][j] = 0;
    }
    for (;;) {
      bool act = false;
#pragma unroll
      for (int i = 0; i < kU; ++i) act = act || ei[i] != -1 || eo[i] != -1;
      if (__ballot(act) == 0) break;
      int t[kU], h[kU], ni[kU], no[kU];
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        t[i] = -1;
        ni[i] = -1;
        h[i] = -1;
        no[i] = -1;
        if (ei[i] != -1) {
          t[i] = e_tail[ei[i]];
          ni[i] = e_next_in[ei[i]];
        }
        if (eo[i] != -1) {
          h[i] = e_head[eo[i]];
          no[i] = e_next_out[eo[i]];
        }
      }
      int rt[kU], rh[kU];
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        rt[i] = t[i] != -1 && included(t[i]) ? row_of[t[i]] : 0;
        rh[i] = h[i] != -1 && included(h[i]) ? row_of[h[i]] : 0;
      }
#pragma unroll
      for (int i = 0; i < kU; ++i) {
        if (rt[i] > 0) {
#pragma unroll
          for (int j = 0; j < kP; ++j) pr[i][j] = np[i] == j ? rt[i] : pr[i][j];
          ++np[i];
        }
        if (rh[i] > 0) {
          ++oc[i];
          lu[i] = max(lu[i], rh[i]);
        }
        ei[i] = ni[i];
        eo[i] = no[i];
      }
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      if (v[i] >= 0) meta[i] |= (oc[i] == 0 ? kRowMetaSink : 0) | (np[i] <= 1 ? kRowMetaSingle : 0);
      max_np = max(max_np, v[i] >= 0 ? max(np[i], 1) : 0);  // row 0 stands in for none
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      const int c0 = r0 + i * kWave;
      if (c0 >= V) break;
      int my_slot = kSlotNone;
      int my_pool = 0;
      const int n_chunk = uniform(min(kWave, V - c0));
      for (int rr = 0; rr < n_chunk; ++rr) {
        const int l = __builtin_amdgcn_readlane(lu[i], rr);
        if (l > 0) {
          const int row = c0 + rr + 1;
          const unsigned long long free_mask = __ballot(busy < row);
          // the lowest free slot, else the global matrix (then the write
          // goes to lane 63, which holds no slot and stays never free)
          const bool none = free_mask == 0;
          const int sl = none ? kSlotGlobal : __ffsll(static_cast<long long>(free_mask)) - 1;
          busy = lane_write(none ? 0x7fffffff : l, none ? kWave - 1 : sl, busy);
          my_slot = lane_write(sl, rr, my_slot);
          if (none && hpool != 0) {
            // the lowest free pool entry (lane k of pool0 holds entry k, of
            // pool1 entry 64 + k); none free: the window reruns exactly
            const unsigned long long g0 = __ballot(pool0 < row);
            const unsigned long long g1 = __ballot(pool1 < row);
            int g = 0;
            if (g0 != 0) {
              g = __ffsll(static_cast<long long>(g0)) - 1;
              pool0 = lane_write(l, g, pool0);
            } else if (g1 != 0) {
              const int t = __ffsll(static_cast<long long>(g1)) - 1;
              pool1 = lane_write(l, t, pool1);
              g = kWave + t;
            } else {
              pool_full = true;
            }
            my_pool = lane_write(g, rr, my_pool);
          }
        }
      }
      const int r = c0 + lane;
      const int cnt = v[i] >= 0 ? max(np[i], 1) : 0;
      const int x = wave_prefix_sum(cnt);
      off[i] = run + x - cnt;
      if (r < V) {
        row_meta[r] = meta[i] | (my_slot << kRowMetaSlotShift) | (my_pool << kRowMetaGShift);
        pred_off[r] = off[i];
        if (slots_in_lds) lslot[r] = static_cast<uint8_t>(my_slot);
      }
      run += __builtin_amdgcn_readlane(x, kWave - 1);
    }
    // predecessor entries (their slots are all assigned by now: predecessor
    // rows come first)
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      if (v[i] < 0) continue;
      const int r = r0 + i * kWave + lane;
      uint64_t deltas = 0;  // row deltas of the first four predecessors (0: none or too far)
      if (np[i] == 0) {
        pred_rows[off[i]] = 0;
        deltas = static_cast<uint64_t>(r + 1 <= 0xffff ? r + 1 : 0);
      } else if (np[i] <= kP) {
#pragma unroll
        for (int j = 0; j < kP; ++j) {
          if (j < np[i]) {
            pred_rows[off[i] + j] = pred_entry(pr[i][j]);
            const int d = r + 1 - pr[i][j];
            if (d <= 0xffff) deltas |= static_cast<uint64_t>(d) << (16 * j);
          }
        }
      } else {
        int o = off[i];
        int k = 0;
        for (int e = in_head[v[i]]; e != -1; e = e_next_in[e]) {
          const int u = e_tail[e];
          if (included(u)) {
            const int p = row_of[u];
            pred_rows[o++] = pred_entry(p);
            const int d = r + 1 - p;
            if (k < 4 && d <= 0xffff) deltas |= static_cast<uint64_t>(d) << (16 * k);
            ++k;
          }
        }
      }
      row_pred4[r] = deltas;
    }
  }
  if (lane == 0) {
    pred_off[V] = run;
  }
  if (pool_full) return -1;
  return __builtin_amdgcn_readlane(wave_prefix_max(max_np), kWave - 1);
}

typedef __attribute__((address_space(3))) uint32_t LWord;

// Included nodes of a layer that does not span the window (spoa
// Graph::subgraph's walk): the nodes >= begin that reach `end` backwards
// along in-edges and aligned sets through nodes >= begin. The set is a
// reachability closure and the stack bound (every in-edge and aligned entry
// once, plus one) is within stack_cap, so the order of the walk changes
// nothing: here every lane pops one stack entry per round, a visit bit in
// LDS (claimed with an atomic OR) lets each node expand once, and the stack
// sits in LDS above the bits, continuing into the slot's global stack.
// Writes in_sub[0..nn). Returns 0, kPoaNodeOverflow, or 1 when the LDS
// (lds_bytes at lds_addr) cannot hold the bits and 64 stack entries.
__device__ __noinline__ int include_walk(GByte* slot_base, CLayout* layout, uint32_t lds_addr,
                                         int lds_bytes, int begin, int end, int nn,
                                         int stack_cap) {
  const Slot s = carve(uniform_ptr(slot_base), *uniform_ptr(layout));
  lds_addr = uniform(static_cast<int>(lds_addr));
  lds_bytes = uniform(lds_bytes);
  begin = uniform(begin);
  end = uniform(end);
  nn = uniform(nn);
  stack_cap = uniform(stack_cap);
  const int lane = lane_id();
  const int n_words = (nn + 31) >> 5;
  const int bits_bytes = static_cast<int>(align_up(static_cast<uint64_t>(n_words) * 4, 16));
  const int lcap = (lds_bytes - bits_bytes) / 4;
  if (lcap < kWave) return 1;
  LWord* bits = (LWord*)(size_t)lds_addr;
  LWord* lstack = (LWord*)(size_t)(lds_addr + bits_bytes);
  int32_t* gstack = s.stack();
  const int32_t* in_head = s.in_head();
  const int32_t* e_next_in = s.e_next_in();
  const int32_t* e_tail = s.e_tail();
  const int32_t* aligned = s.aligned();
  const uint8_t* n_aligned_of = s.n_aligned();
  for (int w = lane; w < n_words; w += kWave) bits[w] = 0u;
  auto st_load = [&](int i) -> int {
    return i < lcap ? static_cast<int>(lstack[i]) : gstack[i];
  };
  auto st_store = [&](int i, int x) {
    if (i < lcap) {
      lstack[i] = static_cast<uint32_t>(x);
    } else {
      gstack[i] = x;
    }
  };
  auto unseen = [&](int u) {
    return u >= begin && ((bits[u >> 5] >> (u & 31)) & 1u) == 0u;
  };
  const uint64_t below = (1ull << lane) - 1ull;
  if (lane == 0) st_store(0, end);
  int top = 1;
  wave_fence();
  // candidates of one round, pushed in lane order; false on overflow
  auto push = [&](bool c, int x) -> bool {
    const uint64_t m = __ballot(c);
    const int cnt = __popcll(m);
    if (top + cnt > stack_cap) return false;
    if (c) st_store(top + __popcll(m & below), x);
    top += cnt;
    return true;
  };
  // every claim sets a bit that stays set, so more than nn claims cannot
  // happen; the count bounds the rounds even so (rounds without a claim only
  // pop)
  int claims = 0;
  while (top > 0) {
    const int nb = min(top, kWave);
    const int first = top - nb;
    int v = -1;
    if (lane < nb) v = st_load(first + lane);
    top = first;
    bool mine = false;
    if (v >= begin) {
      const uint32_t m = 1u << (v & 31);
      mine = (__atomic_fetch_or(&bits[v >> 5], m, __ATOMIC_RELAXED) & m) == 0u;
    }
    claims += __popcll(__ballot(mine));
    if (claims > nn) return kPoaNodeOverflow;
    int e = -1;
    int na = 0;
    int al[kAligned];
    if (mine) {
      e = in_head[v];
      na = n_aligned_of[v];
#pragma unroll
      for (int k = 0; k < static_cast<int>(kAligned); ++k) al[k] = aligned[v * kAligned + k];
    }
    wave_fence();  // the claims before the visit-bit reads below
    const int max_na = uniform(__builtin_amdgcn_readlane(wave_prefix_max(na), kWave - 1));
#pragma unroll
    for (int k = 0; k < static_cast<int>(kAligned); ++k) {
      if (k >= max_na) break;
      const bool c = k < na && unseen(al[k]);
      if (!push(c, al[k])) return kPoaNodeOverflow;
    }
    while (__ballot(e != -1) != 0) {
      int t = -1;
      int nx = -1;
      if (e != -1) {
        t = e_tail[e];
        nx = e_next_in[e];
      }
      if (!push(e != -1 && unseen(t), t)) return kPoaNodeOverflow;
      e = nx;
    }
    wave_fence();
  }
  uint8_t* in_sub = s.in_sub();
  for (int v = lane; v < nn; v += kWave) {
    in_sub[v] = static_cast<uint8_t>((bits[v >> 5] >> (v & 31)) & 1u);
  }
  return 0;
}

// End cell among tied sinks: spoa takes the first tied sink in its DFS order
// of the (sub)graph (roots in id order; a node's predecessors first; a lead
// node pushes its aligned set and emits it after itself, in aligned-list
// order). A root's DFS reaches exactly the nodes from which it can be
// reached along edges and aligned sets, so the aligned set G of a tied sink
// is first reached in the root iteration r = the smallest id among G and
// everything reachable from G (out-edges and aligned sets). When r is itself
// in G, r leads G and G is emitted as r followed by r's aligned list. Tied
// sinks are then ordered by (r, position after r) without the DFS: this is
// exact for sets of sinks (then r is the smallest member) and for sets whose
// descendants all have larger ids. Returns the first tied sink's row, or -1
// when some tied sink's r lies outside its set or its search outgrows the
// bound (the caller then runs the DFS).
__device__ __forceinline__ int sink_tie_by_sets(const Slot& s, int V, bool full, int best,
                                                int stack_cap) {
  constexpr int kReachCap = 256;  // nodes one search may visit
  const int lane = lane_id();
  const uint8_t* in_sub = s.in_sub();
  const int32_t* sub_order = s.sub_order();
  const int32_t* row_meta = s.row_meta();
  const int32_t* sink_score = s.sink_score();
  const uint8_t* n_aligned = s.n_aligned();
  const int32_t* aligned = s.aligned();
  const int32_t* out_head = s.out_head();
  const int32_t* e_next_out = s.e_next_out();
  const int32_t* e_head = s.e_head();
  int32_t* mark = s.tie_mark();
  int32_t* stack = s.stack();
  if (stack_cap < kReachCap) {
    return -1;
  }
  auto kept = [&](int v) { return full || in_sub[v] != 0; };
  int best_key = 0x7fffffff, best_at = -1, searches = 0;
  for (int r0 = 1; r0 <= V; r0 += kWave) {
    const int row = r0 + lane;
    unsigned long long cand = __ballot(row <= V && (row_meta[row - 1] & kRowMetaSink) != 0 &&
                                       sink_score[row] == best);
    while (cand != 0) {
      // one tied sink at a time (the searches share the visit marks)
      const int l = __ffsll(static_cast<long long>(cand)) - 1;
      cand &= cand - 1;
      const int crow = r0 + l;
      const int v = uniform(sub_order[crow - 1]);
      if (++searches >= 256) return -1;
      const int tag = s.n_seqs << 8 | searches;  // unique per layer and search
      // smallest id reachable from v's aligned set (the set included)
      int top = 0, visited = 0, r = v;
      mark[v] = tag;
      stack[top++] = v;
      while (top > 0) {
        const int x = uniform(stack[--top]);
        r = min(r, x);
        if (++visited > kReachCap) return -1;
        for (int e = uniform(out_head[x]); e != -1; e = uniform(e_next_out[e])) {
          const int h = uniform(e_head[e]);
          if (kept(h) && uniform(mark[h]) != tag) {
            if (top >= kReachCap) return -1;
            mark[h] = tag;
            stack[top++] = h;
          }
        }
        const int na = n_aligned[x];
        for (int k = 0; k < na; ++k) {
          const int a = uniform(aligned[x * kAligned + k]);
          if (kept(a) && uniform(mark[a]) != tag) {
            if (top >= kReachCap) return -1;
            mark[a] = tag;
            stack[top++] = a;
          }
        }
      }
      // r must lead v's set: r == v or r in v's aligned list
      int pos = 0;
      if (r != v) {
        bool member = false;
        const int nv = n_aligned[v];
        for (int k = 0; k < nv; ++k) member |= aligned[v * kAligned + k] == r;
        if (!member) return -1;
        const int nr = n_aligned[r];
        for (int k = 0; k < nr; ++k) {
          const int a = uniform(aligned[r * kAligned + k]);
          if (!kept(a)) continue;
          ++pos;
          if (a == v) break;
        }
      }
      const int key = r * 16 + pos;
      if (key < best_key) {
        best_key = key;
        best_at = crow;
      }
    }
  }
  return best_at;
}

// Heaviest-bundle scores over a topological rank order (spoa's
// traverse_heaviest_bundle scan): a node takes the in-edge of the highest
// weight, a later in-edge of equal weight when its tail scores at least as
// much, and adds that tail's score; nodes without in-edges keep -1. The
// recurrence is sequential in rank order, but its inputs are not: each round
// loads, for 64 ranks at once, up to kE in-edges per node (tail, weight,
// tail rank) and the scores of tails from earlier rounds; the scan then runs
// on lane registers (scores of this round's ranks are written into a lane
// register as they are found). Nodes with more in-edges read the rest from
// memory. Scores are 32-bit: the caller takes this path only when every
// path weight fits (|edge weight| <= 252 per inserted position). Writes
// scores[] and preds[]; returns the first node in rank order with the top
// score, with bit 31 set when another node ties it, or kScanUnordered (and
// nothing usable) when some tail does not precede its node in node_rank.
constexpr uint32_t kScanUnordered = 0xffffffffu;

__device__ __noinline__ uint32_t bundle_scan(GByte* slot_base, CLayout* layout, int N) {
  constexpr int kE = 8;
  const Slot s = carve(uniform_ptr(slot_base), *uniform_ptr(layout));
  N = uniform(N);
  const int lane = lane_id();
  const int32_t* r2n = s.rank_to_node();
  const int32_t* in_head = s.in_head();
  const int32_t* e_next_in = s.e_next_in();
  const int32_t* e_tail = s.e_tail();
  const int32_t* e_weight = s.e_weight();
  const int32_t* node_rank = s.node_rank();
  int64_t* scores = s.scores();
  int32_t* preds = s.preds();
  int max_node = 0;
  int max_score = -1;  // the initial max_node (node 0) scores -1 until it is scanned
  bool tied = false;
  bool unordered = false;
  for (int r0 = 0; r0 < N; r0 += kWave) {
    const int r = r0 + lane;
    const int v = r < N ? r2n[r] : -1;
    int t[kE], w[kE], tr[kE], ts[kE];
    int e = v >= 0 ? in_head[v] : -1;
#pragma unroll
    for (int j = 0; j < kE; ++j) {
      t[j] = -1;
      w[j] = 0;
      if (e != -1) {
        t[j] = e_tail[e];
        w[j] = e_weight[e];
        e = e_next_in[e];
      }
    }
#pragma unroll
    for (int j = 0; j < kE; ++j) tr[j] = t[j] != -1 ? node_rank[t[j]] : r0;
#pragma unroll
    for (int j = 0; j < kE; ++j) {
      ts[j] = t[j] != -1 && tr[j] < r0 ? static_cast<int>(scores[t[j]]) : 0;
    }
    int sc = -1;  // lane k: score of rank r0 + k
    int pr = -1;  // lane k: its predecessor
    const int n = uniform(min(kWave, N - r0));
    for (int k = 0; k < n; ++k) {
      const int vk = __builtin_amdgcn_readlane(v, k);
      int sv = -1, pv = -1, spv = 0;
      auto consider = [&](int tj, int wj, int stj) {
        if (sv < wj || (sv == wj && pv != -1 && spv <= stj)) {
          sv = wj;
          pv = tj;
          spv = stj;
        }
      };
#pragma unroll
      for (int j = 0; j < kE; ++j) {
        const int tj = __builtin_amdgcn_readlane(t[j], k);
        if (tj == -1) break;
        const int trj = __builtin_amdgcn_readlane(tr[j], k);
        unordered |= trj >= r0 + k;
        const int stj = trj >= r0 ? __builtin_amdgcn_readlane(sc, (trj - r0) & (kWave - 1))
                                  : __builtin_amdgcn_readlane(ts[j], k);
        consider(tj, __builtin_amdgcn_readlane(w[j], k), stj);
      }
      for (int f = uniform(__builtin_amdgcn_readlane(e, k)); f != -1; f = uniform(e_next_in[f])) {
        const int tj = uniform(e_tail[f]);
        const int trj = uniform(node_rank[tj]);
        unordered |= trj >= r0 + k;
        const int stj = trj >= r0 ? __builtin_amdgcn_readlane(sc, (trj - r0) & (kWave - 1))
                                  : uniform(static_cast<int>(scores[tj]));
        consider(tj, uniform(e_weight[f]), stj);
      }
      const int score = pv != -1 ? sv + spv : sv;
      sc = lane_write(score, k, sc);
      pr = lane_write(pv, k, pr);
      if (vk == max_node) {
        max_score = score;
      } else if (max_score < score) {
        max_node = vk;
        max_score = score;
        tied = false;
      } else if (max_score == score) {
        tied = true;
      }
    }
    if (v >= 0) {
      scores[v] = sc;
      preds[v] = pr;
    }
  }
  if (unordered) return kScanUnordered;
  return static_cast<uint32_t>(max_node) | (tied ? 0x80000000u : 0u);
}

// ---- the window ----

__device__ __forceinline__ void process_window(CParams& P, Slot& s, uint8_t* lds, CWindow& W,
                               PoaResult& R) {
  const int lane = lane_id();
  CCaps& caps = P.caps;
  const int pitch = static_cast<int>(caps.cols);
  const GByte* arena = global(P.arena);
  const GByte* bb = arena + W.backbone_off;
  const GByte* bq = arena + W.backbone_qual;
  const int n = static_cast<int>(W.backbone_len);
  uint64_t cells = 0;
  int max_rows = 0;
  PhaseClock clk;

  // backbone chain: nodes 0..n-1, edges i-1 -> i (weights q-33), label 0
  if (n > static_cast<int>(caps.nodes) || n - 1 > static_cast<int>(caps.edges)) {
    R.status = kPoaNodeOverflow;
    return;
  }
  for (int i = lane; i < n; i += kWave) {
    s.base()[i] = bb[i];
    s.in_head()[i] = i > 0 ? i - 1 : -1;
    s.in_tail()[i] = i > 0 ? i - 1 : -1;
    s.out_head()[i] = i + 1 < n ? i : -1;
    s.out_tail()[i] = i + 1 < n ? i : -1;
    s.n_aligned()[i] = 0;
    s.cov()[i] = n >= 2 ? 1 : 0;
    s.last_label()[i] = n >= 2 ? 0 : -1;
    s.rank_to_node()[i] = i;
    s.node_rank()[i] = i;
    if (i + 1 < n) {
      s.e_tail()[i] = i;
      s.e_head()[i] = i + 1;
      s.e_weight()[i] = (static_cast<int>(bq[i]) - 33) + (static_cast<int>(bq[i + 1]) - 33);
      s.e_next_in()[i] = -1;
      s.e_next_out()[i] = -1;
    }
  }
  s.n_nodes = n;
  s.n_edges = n > 0 ? n - 1 : 0;
  s.n_seqs = 1;
  // visit marks of the tie resolution: (layer, lane) tags, none set yet
  // (node ids stay below the slot's node capacity, which may be below the
  // window's node bound)
  const int n_marks = min(static_cast<int>(W.node_bound), static_cast<int>(caps.nodes));
  for (int v = lane; v < n_marks; v += kWave) {
    s.tie_mark()[v] = 0;
  }
  wave_sync();
  clk.lap(R, kPhaseSetup);

  for (uint32_t li = 0; li < W.n_layers; ++li) {
    CLayer& Lr = ((CLayer*)P.layers)[W.layer_first + li];
    const int L = static_cast<int>(Lr.len);
    if (L + static_cast<int>(kPoaColPad) > pitch) {
      R.status = kPoaRowOverflow;
      return;
    }
    // layer bytes and weights into the slot
    uint8_t* lds_seq = lds + align_up(static_cast<uint64_t>(P.ring_rows) * pitch * 2, 16);
    for (int j = lane; j < L; j += kWave) {
      const uint8_t c = arena[Lr.seq_off + j];
      s.seq()[j] = c;
      lds_seq[j] = c;
      s.qual()[j] = Lr.qual_off == kNoQuality
                      ? 1
                      : static_cast<int>(arena[Lr.qual_off + j]) - 33;
    }
    const bool full = Lr.begin < W.offset && Lr.end > W.backbone_len - W.offset;
    if (static_cast<int>(Lr.end) >= s.n_nodes || Lr.begin > Lr.end) {
      R.status = kPoaBadInput;
      return;
    }

    // 1. included node set (all nodes for a layer spanning the window)
    const int nn = s.n_nodes;
    int walk_rc = 1;
    if (!full) {
      uint32_t walk_lds = static_cast<uint32_t>((size_t)(LByte*)lds);
      asm volatile("" : "+v"(walk_lds));
      walk_rc = uniform(include_walk(s.p, s.L, walk_lds, static_cast<int>(P.ring_rows) * pitch * 2,
                                     static_cast<int>(Lr.begin), static_cast<int>(Lr.end), nn,
                                     static_cast<int>(caps.stack)));
      if (walk_rc < 0) {
        R.status = walk_rc;
        return;
      }
    }
    if (walk_rc != 0) {
      for (int v = lane; v < nn; v += kWave) {
        s.in_sub()[v] = full ? 1 : 0;
      }
      wave_sync();
      if (!full) {
        int top = 0;
        s.stack()[top++] = static_cast<int>(Lr.end);
        while (top > 0) {
          const int v = uniform(s.stack()[--top]);
          if (s.in_sub()[v] == 0 && v >= static_cast<int>(Lr.begin)) {
            const int na = s.n_aligned()[v];
            int deg = na;
            for (int e = uniform(s.in_head()[v]); e != -1; e = uniform(s.e_next_in()[e])) {
              ++deg;
            }
            if (top + deg > static_cast<int>(caps.stack)) {
              R.status = kPoaNodeOverflow;
              return;
            }
            for (int e = uniform(s.in_head()[v]); e != -1; e = uniform(s.e_next_in()[e])) {
              s.stack()[top++] = s.e_tail()[e];
            }
            for (int k = 0; k < na; ++k) {
              s.stack()[top++] = s.aligned()[v * kAligned + k];
            }
            s.in_sub()[v] = 1;
          }
        }
      }
    }
    wave_sync();

    clk.lap(R, kPhaseInclude);
    // 2. DP row order: the maintained topological order restricted to the
    // included nodes (the DP values and move codes do not depend on which
    // topological order is used; spoa's exact DFS order is computed only
    // where it decides a tie)
    int V = 0;
    // (with the DP row of every included node, row_of)
    if (full) {
      V = nn;
      for (int r = lane; r < V; r += kWave) {
        const int v = s.rank_to_node()[r];
        s.sub_order()[r] = v;
        s.row_of()[v] = r + 1;
      }
    } else {
      for (int r0 = 0; r0 < nn; r0 += kWave) {
        const int r = r0 + lane;
        const int v = r < nn ? s.rank_to_node()[r] : 0;
        const bool keep = r < nn && s.in_sub()[v] != 0;
        const unsigned long long mask = __ballot(keep);
        if (keep) {
          const int at = V + __popcll(mask & ((1ull << lane) - 1ull));
          s.sub_order()[at] = v;
          s.row_of()[v] = at + 1;
        }
        V += __popcll(mask);
      }
    }
    if (V + 1 > static_cast<int>(caps.rows)) {
      R.status = kPoaRowOverflow;
      return;
    }
    max_rows = max(max_rows, V + 1);
    wave_sync();
    clk.lap(R, kPhaseSubSort);

    // score width: 16-bit ring rows when every score of the layer fits
    const int32_t max_abs = max(max(abs(P.match), abs(P.mismatch)), abs(P.gap));
    const bool narrow = static_cast<int64_t>(max_abs) * (V + L + 2) < 32000;
    // (a pass for very long layers may keep no rows in LDS at all)
    const int n_slots = min(kMaxLdsSlots, narrow ? static_cast<int>(P.ring_rows)
                                                 : static_cast<int>(P.ring_rows) / 2);

    bool strip_ok = false;  // the one-pass DP applies (dp_layer_strip)

    // predecessor rows, sink flags, slots
    uint32_t rows_lds = static_cast<uint32_t>((size_t)(LByte*)lds);
    asm volatile("" : "+v"(rows_lds));
    const int ring_bytes = static_cast<int>(P.ring_rows) * pitch * 2;
    const int max_np =
        uniform(full ? build_rows<true>(s.p, s.L, V, n_slots, rows_lds, ring_bytes)
                     : build_rows<false>(s.p, s.L, V, n_slots, rows_lds, ring_bytes));
    wave_sync();
    if (max_np < 0) {
      R.status = kPoaRowOverflow;  // the pooled global rows ran out: rerun exactly
      return;
    }
    strip_ok = narrow && L >= 1 && max_abs <= 4096 && max_np <= 255;
    if (max_np > static_cast<int>(kCodePayload) + 1) {
      // a move code holds the in-edge index in 13 bits
      R.status = kPoaInDegree;
      return;
    }

    clk.lap(R, kPhaseRows);
    // 3. DP over the layer (column 0 folded into each row)
    clk.lap(R, kPhaseColumn0);
    int best_score = 0;
    bool tied = false;
    uint64_t dp;
    CParams* pp = &P;
    CLayout* ly = s.L;
    uint32_t lds_addr = static_cast<uint32_t>((size_t)(LByte*)lds);
    // opaque to the optimizer: otherwise the callee would rebuild the address
    // from the dynamic LDS symbol (a table load) inside its loops
    asm volatile("" : "+v"(lds_addr));
    // packed pairs keep H - j * gap in 16 bits
    const bool packed = strip_ok && static_cast<int64_t>(max_abs) * (V + 2 * L + 2) < 32000;
    // racon -b: band rows for layers longer than the band (approximation)
    const bool banded = packed && P.band != 0 && L > kBand;
    if (banded) {
      dp = dp_layer_call<kDpBand4>(pp, s.p, ly, lds_addr, V, L);
    } else if (packed && L <= 4 * kWave) {
      dp = dp_layer_call<kDpPacked4>(pp, s.p, ly, lds_addr, V, L);
    } else if (packed && L <= 8 * kWave) {
      dp = dp_layer_call<kDpPacked8>(pp, s.p, ly, lds_addr, V, L);
    } else if (packed && L <= 12 * kWave) {
      dp = dp_layer_call<kDpPacked12>(pp, s.p, ly, lds_addr, V, L);
    } else if (packed && L <= 16 * kWave) {
      dp = dp_layer_call<kDpPacked16>(pp, s.p, ly, lds_addr, V, L);
    } else if (strip_ok && L <= 4 * kWave) {
      dp = dp_layer_call<kDpStrip4>(pp, s.p, ly, lds_addr, V, L);
    } else if (strip_ok && L <= 8 * kWave) {
      dp = dp_layer_call<kDpStrip8>(pp, s.p, ly, lds_addr, V, L);
    } else if (strip_ok && L <= 12 * kWave) {
      dp = dp_layer_call<kDpStrip12>(pp, s.p, ly, lds_addr, V, L);
    } else if (narrow) {
      dp = dp_layer_call<kDpGeneric16>(pp, s.p, ly, lds_addr, V, L);
    } else {
      dp = dp_layer_call<kDpGeneric32>(pp, s.p, ly, lds_addr, V, L);
    }
    int best_row = uniform(static_cast<int>(static_cast<uint32_t>(dp) & 0x7fffffffu));
    best_score = uniform(static_cast<int>(dp >> 32));
    tied = uniform(static_cast<int>(static_cast<uint32_t>(dp) >> 31)) != 0;
    wave_sync();
    if (banded && best_score == kNegInf) {
      R.status = kPoaBandMiss;  // no sink reaches column L inside its band
      return;
    }
    if (tied) {
      best_row = uniform(sink_tie_by_sets(s, V, full, best_score, static_cast<int>(caps.stack)));
    }
    if (tied && best_row < 0) {
      // several sinks share the best score: spoa takes the first one in its
      // DFS order of the (sub)graph
      const int cnt = topo_sort(s, full ? nullptr : s.in_sub(), s.path(),
                                static_cast<int>(caps.stack));
      if (s.overflow) {
        R.status = kPoaNodeOverflow;
        return;
      }
      wave_sync();
      int found = -1;
      for (int i0 = 0; i0 < cnt && found < 0; i0 += kWave) {
        const int i = i0 + lane;
        bool hit = false;
        int row = 0;
        if (i < cnt) {
          row = s.row_of()[s.path()[i]];
          hit = (s.row_meta()[row - 1] & kRowMetaSink) != 0 && s.sink_score()[row] == best_score;
        }
        const unsigned long long mask = __ballot(hit);
        if (mask != 0) {
          found = __shfl(row, __ffsll(static_cast<long long>(mask)) - 1, kWave);
        }
      }
      best_row = uniform(found);
    }
    cells += static_cast<uint64_t>(V) * static_cast<uint64_t>(banded ? kBand : L);
    clk.lap(R, kPhaseDp);

    // 4. traceback, from the end cell to (0, 0). The alignment of layer
    // position p (the graph node of a diagonal move into column p + 1, or -1
    // for a horizontal move) goes to pos_node[p]; vertical moves align no
    // layer position. Per round trip the wave loads the 16-row x 8-column
    // block of codes at and before the current cell (the path falls about two
    // rows per column) with the node and the predecessor-delta table of each
    // block row, and every lane turns its two cells into their successor cell
    // in the block (or "leaves the block", or "needs the lists"). The walk
    // through the block is then a chase of successor indices read from lane
    // registers; the move that leaves the block (and any the table left to
    // the predecessor lists) is taken by the full single-step decoder. The
    // block's positions are collected in one register (lane t: position
    // j0 - 1 - t) and stored together.
    {
      const uint16_t* code = s.code();
      const int32_t* sub_order = s.sub_order();
      const int32_t* pred_off = s.pred_off();
      const int32_t* pred_rows = s.pred_rows();
      const uint64_t* row_pred4 = s.row_pred4();
      const int32_t* row_meta = s.row_meta();
      // the packed DP and the band rows keep move kinds for one-predecessor rows
      const bool kinds = banded || (packed && L <= 16 * kWave);
      int32_t* pos_node = s.path();
      constexpr uint32_t kNextExit = 0xff;  // successor: the decoder takes the move
      constexpr uint32_t kDiagBit = 0x100;
      int i = best_row, j = L;
      int steps = 0;
      const int step_cap = uniform(V + L + 1);  // every move lowers i or j
      bool bad = false;
      bool band_miss = false;
      while (i > 0 && !bad) {
        i = uniform(i);
        j = uniform(j);
        const int i0 = i, j0 = j;
        // lane l: row i0 - (l >> 2), columns j0 - 2 (l & 3) - 0..1 (two codes
        // per register); indices outside the matrix are clamped to a valid
        // cell and the value dropped, so the loads issue unconditionally
        const int di_l = lane >> 2;
        const int ri = i0 - di_l;
        const int rc = max(ri, 1);
        const int cj = j0 - 2 * (lane & 3);
        const uint16_t* crow = code + static_cast<int64_t>(rc) * pitch;
        uint32_t c0, c1;
        if (kinds && (row_meta[rc - 1] & kRowMetaSingle) != 0) {
          // a one-predecessor row of the packed DP: its 2-bit move kinds,
          // turned into the codes of moves to in-edge 0
          const uint8_t* krow = reinterpret_cast<const uint8_t*>(crow);
          auto kind_code = [&](int j) -> uint32_t {
            if (j <= 0) return kCodeVert | kCodeViaK;
            const uint32_t k = (static_cast<uint32_t>(krow[(j - 1) >> 2]) >> (2 * ((j - 1) & 3))) & 3u;
            return k == kKindHorz ? kCodeHorz
                                  : ((k == kKindVert ? kCodeVert : kCodeDiag) | kCodeViaK);
          };
          c0 = kind_code(cj);
          c1 = kind_code(cj - 1);
        } else {
          c0 = crow[code_index(max(cj, 0), pitch)];
          c1 = crow[code_index(max(cj - 1, 0), pitch)];
        }
        const uint64_t d4 = row_pred4[rc - 1];  // first four predecessor row deltas of the row
        const int nrow = sub_order[rc - 1];     // graph node of the row
        const uint32_t cw = ri >= 1 ? ((cj >= 0 ? c0 : 0u) | ((cj >= 1 ? c1 : 0u) << 16)) : 0u;
        wait_stores();  // the block is in registers before the walk starts
        // successor of the lane's cell (di_l, dj) as di * 8 + dj, with the
        // diagonal flag; kNextExit when the move leaves the block or row 1..,
        // reaches column 0 other than vertically, or needs the lists
        auto successor = [&](uint32_t c, int dj) -> uint32_t {
          const int kind = static_cast<int>(c & 0xc000u);
          const int pay = static_cast<int>(c & kCodePayload);
          int delta = pay;
          if ((c & kCodeViaK) != 0) {
            delta = pay < 4 ? static_cast<int>((d4 >> (16 * pay)) & 0xffffu) : 0;
          }
          const int jj = j0 - dj;
          if (ri < 1 || jj < 0 || (kind != kCodeVert && jj == 0) || kind == 0xc000) return kNextExit;
          int ndi = di_l, ndj = dj;
          if (kind == kCodeHorz) {
            ndj = dj + 1;
          } else {
            if (delta == 0) return kNextExit;
            ndi = di_l + delta;
            if (kind == kCodeDiag) ndj = dj + 1;
          }
          if (ndi >= 16 || ndj >= 8 || i0 - ndi <= 0) return kNextExit;
          // band rows: a move out of the band is left to the decoder, which
          // stops the walk there
          if (banded && !in_band(i0 - ndi, j0 - ndj, V, L)) return kNextExit;
          return static_cast<uint32_t>(ndi * 8 + ndj) | (kind == kCodeDiag ? kDiagBit : 0u);
        };
        const int dja = 2 * (lane & 3);
        const uint32_t ent = successor(cw & 0xffffu, dja) | (successor(cw >> 16, dja + 1) << 16);
        int out = -1;  // lane t: the node of position j0 - 1 - t
        // moves inside the block (each raises di * 8 + dj, so at most 127)
        int idx = 0;
        for (;;) {
          const uint32_t e = (__builtin_amdgcn_readlane(ent, idx >> 1) >> (16 * (idx & 1))) & 0xffffu;
          if ((e & 0xffu) == kNextExit || steps >= step_cap) break;
          ++steps;
          if ((e & kDiagBit) != 0) {
            const int node = __builtin_amdgcn_readlane(nrow, (idx >> 3) * 4);
            out = lane == (idx & 7) ? node : out;
          }
          idx = static_cast<int>(e & 0xffu);
        }
        i = i0 - (idx >> 3);
        j = j0 - (idx & 7);
        // the move out of the block: the full decoder
        {
          const int dj = idx & 7;
          const uint32_t cp = __builtin_amdgcn_readlane(cw, idx >> 1);
          const int c = static_cast<int>((cp >> (16 * (idx & 1))) & 0xffffu);
          const int kind = c & 0xc000;
          if (++steps > step_cap || (kind != kCodeVert && j == 0)) {
            bad = true;  // only vertical moves leave column 0
          } else if (kind == kCodeHorz) {
            --j;
          } else {
            if (kind == kCodeDiag) {
              const int node = __builtin_amdgcn_readlane(nrow, (idx >> 3) * 4);
              out = lane == dj ? node : out;
              --j;
            }
            const int pay = c & static_cast<int>(kCodePayload);
            int delta = pay;
            if ((c & kCodeViaK) != 0) {
              // predecessor by in-edge index: the row's delta table, else the lists
              delta = 0;
              if (pay < 4) {
                const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(d4), (idx >> 3) * 4);
                const uint32_t hi =
                    __builtin_amdgcn_readlane(static_cast<uint32_t>(d4 >> 32), (idx >> 3) * 4);
                const uint32_t w = pay < 2 ? lo : hi;
                delta = static_cast<int>((w >> (16 * (pay & 1))) & 0xffffu);
              }
              if (delta == 0) {
                delta = i - (uniform(pred_rows[uniform(pred_off[i - 1]) + pay]) & kPredRowMask);
              }
            }
            i -= delta;
            bad = i < 0;
          }
          if (!bad && banded && !in_band(i, j, V, L)) band_miss = bad = true;
        }
        if (lane < j0 - j) {
          pos_node[j0 - 1 - lane] = out;
        }
      }
      if (bad) {
        R.status = band_miss ? kPoaBandMiss : kPoaBadInput;
        return;
      }
      // from the virtual start row only horizontal moves remain
      for (int p = lane; p < j; p += kWave) {
        pos_node[p] = -1;
      }
    }
    wave_sync();

    clk.lap(R, kPhaseTraceback);
    // 5. insert the layer: lane-parallel over layer positions unless two
    // nodes of the path share an aligned set (then position by position)
    bool order_ok = true;
    int n_path = 0;
    const int nodes_before = s.n_nodes;
    {
      uint32_t lseq_addr = static_cast<uint32_t>((size_t)(LByte*)lds_seq);
      asm volatile("" : "+v"(lseq_addr));
      const InsertResult ir =
          insert_parallel(s.p, s.L, lseq_addr, L, s.n_nodes, s.n_edges, s.n_seqs,
                          static_cast<int>(caps.nodes), static_cast<int>(caps.edges));
      const int rc = uniform(ir.status);
      if (rc < 0) {
        R.status = rc;
        return;
      }
      if (rc == 0) {
        s.n_nodes = uniform(ir.n_nodes);
        s.n_edges = uniform(ir.n_edges);
        s.n_seqs = s.n_seqs + 1;
        order_ok = uniform(ir.order_ok) != 0;
        n_path = L;
      } else {
        s = insert_sequential(s, caps, L, &order_ok, &n_path);
      }
    }
    if (s.overflow) {
      R.status = n_path < 0 ? kPoaAlignedFull : kPoaNodeOverflow;
      return;
    }
    wave_sync();
    clk.lap(R, kPhaseInsert);
    // topological order of the grown graph: new nodes merged after their
    // anchors (old nodes keep their relative order), then checked along the
    // inserted path; any violation falls back to the exact DFS order
    {
      const int M = s.n_nodes - nodes_before;
      const int total = s.n_nodes;
      if (order_ok && M > 0) {
        // an old node of rank r moves up by the number of anchors below r
        // (anchors are non-decreasing; searched in LDS when they fit), so a
        // sweep from the top rank down reads every entry before it is
        // overwritten; ranks up to the first anchor stay
        const int32_t* anchor = s.anchor();
        LWord* la = (LWord*)(size_t)static_cast<uint32_t>((size_t)(LByte*)lds);
        const bool anchors_in_lds = M * 4 <= static_cast<int>(P.ring_rows) * pitch * 2;
        if (anchors_in_lds) {
          for (int k = lane; k < M; k += kWave) la[k] = static_cast<uint32_t>(anchor[k]);
          wave_fence();
        }
        const int first_moved = uniform(anchor[0]) + 1;
        // four chunks of 64 ranks per round, loaded before any of their
        // stores (every store lands at or above its own rank, and later
        // rounds read only below this one)
        constexpr int kMU = 4;
        for (int r0 = ((nodes_before - 1) / kWave) * kWave; r0 + kWave > first_moved;
             r0 -= kMU * kWave) {
          int r[kMU], lo[kMU], v[kMU];
          bool moves[kMU];
#pragma unroll
          for (int u = 0; u < kMU; ++u) {
            r[u] = r0 - u * kWave + lane;
            moves[u] = r[u] >= first_moved && r[u] < nodes_before;
            v[u] = moves[u] ? s.rank_to_node()[r[u]] : 0;
          }
#pragma unroll
          for (int u = 0; u < kMU; ++u) {
            int l = 1, hi = M;  // anchor[0] < r for every moved rank
            while (__ballot(moves[u] && l < hi) != 0) {
              if (moves[u] && l < hi) {
                const int mid = (l + hi) >> 1;
                const int a = anchors_in_lds ? static_cast<int>(la[mid]) : anchor[mid];
                if (a < r[u]) l = mid + 1; else hi = mid;
              }
            }
            lo[u] = l;
          }
#pragma unroll
          for (int u = 0; u < kMU; ++u) {
            if (moves[u]) {
              s.rank_to_node()[r[u] + lo[u]] = v[u];
              s.node_rank()[v[u]] = r[u] + lo[u];
            }
          }
          wave_fence();
        }
        wave_sync();
        for (int k = lane; k < M; k += kWave) {
          const int r = anchor[k] + 1 + k;
          s.rank_to_node()[r] = nodes_before + k;
          s.node_rank()[nodes_before + k] = r;
        }
        wave_sync();
      }
      bool bad = !order_ok;
      if (!bad) {
        for (int i = lane; i + 1 < n_path; i += kWave) {
          if (s.node_rank()[s.lpath()[i]] >= s.node_rank()[s.lpath()[i + 1]]) bad = true;
        }
        bad = __ballot(bad) != 0;
      }
      if (bad) {
        topo_sort(s, nullptr, s.rank_to_node(), static_cast<int>(caps.stack));
        if (s.overflow) {
          R.status = kPoaNodeOverflow;
          return;
        }
        wave_sync();
        for (int r = lane; r < total; r += kWave) {
          s.node_rank()[s.rank_to_node()[r]] = r;
        }
      }
    }
    wave_sync();
    clk.lap(R, kPhaseSort);
  }

  clk.lap(R, kPhaseSort);

  // heaviest bundle (lane-uniform). Node scores and predecessors do not
  // depend on which topological order visits the nodes; spoa's DFS order
  // decides only which of several top-scoring nodes comes first and which
  // nodes branch completion revisits. The first pass takes the maintained
  // order; when the top score is shared or its node has out-edges, spoa's
  // exact DFS order of the final graph is computed and the bundle taken
  // again. (With no layers the backbone chain is already in DFS order.)
  const int N = s.n_nodes;
  int max_node = 0;
  // 32-bit scores: every path weight is at most 252 per inserted position
  const bool fast_scan = static_cast<int64_t>(W.node_bound) * 252 < 0x7fffffffll;
  for (int pass = 0;; ++pass) {
    const bool exact = pass > 0 || W.n_layers == 0;
    if (pass > 0) {
      clk.lap(R, kPhaseConsensus);
      topo_sort(s, nullptr, s.rank_to_node(), static_cast<int>(caps.stack));
      if (s.overflow) {
        R.status = kPoaNodeOverflow;
        return;
      }
      wave_sync();
      // ranks of the new order (the scan looks tails up by rank)
      for (int r = lane; r < N; r += kWave) {
        s.node_rank()[s.rank_to_node()[r]] = r;
      }
      wave_sync();
      clk.lap(R, kPhaseSort);
    }
    max_node = 0;
    bool top_tied = false;
    uint32_t mt = kScanUnordered;
    if (fast_scan) {
      mt = static_cast<uint32_t>(uniform(static_cast<int>(bundle_scan(s.p, s.L, N))));
      max_node = static_cast<int>(mt & 0x7fffffffu);
      top_tied = (mt >> 31) != 0;
    }
    if (mt == kScanUnordered) {
    for (int v = lane; v < N; v += kWave) {
      s.scores()[v] = -1;
      s.preds()[v] = -1;
    }
    wave_sync();
    for (int r = 0; r < N; ++r) {
      const int v = uniform(s.rank_to_node()[r]);
      for (int e = uniform(s.in_head()[v]); e != -1; e = uniform(s.e_next_in()[e])) {
        const int t = uniform(s.e_tail()[e]);
        const int64_t w = s.e_weight()[e];
        const int64_t sv = s.scores()[v];
        const int pv = s.preds()[v];
        if (sv < w || (sv == w && pv != -1 && s.scores()[pv] <= s.scores()[t])) {
          s.scores()[v] = w;
          s.preds()[v] = t;
        }
      }
      const int pv = s.preds()[v];
      if (pv != -1) {
        s.scores()[v] = s.scores()[v] + s.scores()[pv];
      }
      if (s.scores()[max_node] < s.scores()[v]) {
        max_node = v;
        top_tied = false;
      } else if (v != max_node && s.scores()[max_node] == s.scores()[v]) {
        top_tied = true;
      }
    }
    }
    wave_sync();
    if (!exact && (top_tied || s.out_head()[max_node] != -1)) continue;
    if (s.out_head()[max_node] != -1) {
      for (int r = lane; r < N; r += kWave) {
        s.node_rank()[s.rank_to_node()[r]] = r;
      }
      wave_sync();
      int guard = 0;
      while (s.out_head()[max_node] != -1) {
        if (++guard > N + 1) {
          R.status = kPoaLoopLimit;
          return;
        }
        // branch completion from the rank of max_node
        const int rank = s.node_rank()[max_node];
        const int v0 = max_node;
        for (int e = uniform(s.out_head()[v0]); e != -1; e = uniform(s.e_next_out()[e])) {
          const int h = uniform(s.e_head()[e]);
          for (int f = uniform(s.in_head()[h]); f != -1; f = uniform(s.e_next_in()[f])) {
            const int t = uniform(s.e_tail()[f]);
            if (t != v0) {
              s.scores()[t] = -1;
            }
          }
        }
        int64_t best_score = 0;
        int best_node = 0;
        for (int r = rank + 1; r < N; ++r) {
          const int u = uniform(s.rank_to_node()[r]);
          s.scores()[u] = -1;
          s.preds()[u] = -1;
          for (int e = uniform(s.in_head()[u]); e != -1; e = uniform(s.e_next_in()[e])) {
            const int t = uniform(s.e_tail()[e]);
            const int64_t st = s.scores()[t];
            if (st == -1) {
              continue;
            }
            const int64_t w = s.e_weight()[e];
            const int64_t su = s.scores()[u];
            const int pu = s.preds()[u];
            if (su < w || (su == w && pu != -1 && s.scores()[pu] <= st)) {
              s.scores()[u] = w;
              s.preds()[u] = t;
            }
          }
          const int pu = s.preds()[u];
          if (pu != -1) {
            s.scores()[u] = s.scores()[u] + s.scores()[pu];
          }
          if (best_score < s.scores()[u]) {
            best_score = s.scores()[u];
            best_node = u;
          }
        }
        max_node = best_node;
      }
    }
    break;
  }
  // path back from max_node
  int plen = 0;
  {
    int v = max_node;
    while (true) {
      s.path()[plen++] = v;
      const int p = uniform(s.preds()[v]);
      if (p == -1 || plen >= N) {
        break;
      }
      v = p;
    }
  }
  wave_sync();
  // coverage and trimming (window.cpp:118-139); path is stored backwards
  int begin = 0, end = plen - 1;
  uint8_t status = kPoaPolished;
  if ((W.flags & 3u) == 3u) {
    const int threshold = static_cast<int>(W.n_layers) / 2;  // (n_seqs - 1) / 2
    auto coverage = [&](int idx) -> int {
      const int v = s.path()[plen - 1 - idx];
      int c = s.cov()[v];
      const int na2 = s.n_aligned()[v];
      for (int k = 0; k < na2; ++k) {
        c += s.cov()[s.aligned()[v * kAligned + k]];
      }
      return c;
    };
    for (; begin < plen; ++begin) {
      if (coverage(begin) >= threshold) break;
    }
    for (; end >= 0; --end) {
      if (coverage(end) >= threshold) break;
    }
    if (begin >= end) {
      status = kPoaPolishedUntrimmed;
      begin = 0;
      end = plen - 1;
    }
  }
  const int out_len = end - begin + 1;
  if (out_len > static_cast<int>(W.out_cap)) {
    R.status = kPoaOutOverflow;
    return;
  }
  for (int i = lane; i < out_len; i += kWave) {
    global(P.out)[W.out_off + i] = s.base()[s.path()[plen - 1 - (begin + i)]];
  }
  clk.lap(R, kPhaseConsensus);
  R.length = static_cast<uint32_t>(out_len);
  R.status = status;
  R.cells = cells;
  R.nodes = static_cast<uint32_t>(N);
  R.max_rows = static_cast<uint32_t>(max_rows);
}

}  // namespace

// Persistent batch kernel: each wave takes windows from the shared queue.
// The parameters stay in device memory (read-only for the whole launch) and
// are read through a restrict pointer: uniform scalar loads, and no private
// copy of the struct (a by-value kernel argument whose address is taken is
// copied to scratch, with its fields held in vector registers).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RH_POA_WAVES_PER_EU))) void RH_POA_SYM(poa_batch_kernel)(
    const PoaParams* __restrict__ params) {
  CParams& P = *(CParams*)params;
  extern __shared__ uint8_t lds[];
  GByte* slot_mem = global(P.scratch) + static_cast<uint64_t>(blockIdx.x) * P.slot_bytes;
  __shared__ uint32_t next;
  while (true) {
    if (threadIdx.x == 0) {
      next = atomicAdd(P.queue_head, 1u);
    }
    __syncthreads();
    const uint32_t q = next;
    __syncthreads();
    if (q >= P.n_order) {
      break;
    }
    const uint32_t w = global(P.order)[q];
    // the heaviest windows (front of the cost-ordered queue) bound the
    // kernel's length: their waves issue ahead of the other waves of their
    // SIMD while both are ready
    if (q < P.n_priority) {
      __builtin_amdgcn_s_setprio(3);
    } else {
      __builtin_amdgcn_s_setprio(0);
    }
    Slot s = carve(slot_mem, P.layout);
    // the window's result accumulates in LDS (phase ticks are read-modify-
    // written by every lane with the same value), not in registers
    __shared__ PoaResult res;
    if (threadIdx.x == 0) {
      res = PoaResult{};
      res.status = kPoaNotRun;
    }
    __syncthreads();
    process_window(P, s, lds, ((CWindow*)P.windows)[w], res);
    __syncthreads();
    if (threadIdx.x == 0) {
      P.results[w] = res;
    }
    __syncthreads();
  }
}

// Plain kernel: one launch per window (single wave, window index given).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RH_POA_WAVES_PER_EU))) void RH_POA_SYM(poa_single_kernel)(
    const PoaParams* __restrict__ params, uint32_t q) {
  CParams& P = *(CParams*)params;
  extern __shared__ uint8_t lds[];
  const uint32_t w = global(P.order)[q];
  Slot s = carve(global(P.scratch), P.layout);
  __shared__ PoaResult res;
  if (threadIdx.x == 0) {
    res = PoaResult{};
    res.status = kPoaNotRun;
  }
  __syncthreads();
  process_window(P, s, lds, ((CWindow*)P.windows)[w], res);
  __syncthreads();
  if (threadIdx.x == 0) {
    P.results[w] = res;
  }
}

// Static-grid kernel: one workgroup per window, no queue and no cost order.
// Block b takes window order[first + b] and scratch slot b; the hardware
// dispatcher starts blocks in index order as CUs free up. Kept as the
// comparison for the persistent kernel's queue.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RH_POA_WAVES_PER_EU))) void RH_POA_SYM(poa_static_kernel)(
    const PoaParams* __restrict__ params, uint32_t first) {
  CParams& P = *(CParams*)params;
  extern __shared__ uint8_t lds[];
  const uint32_t w = global(P.order)[first + blockIdx.x];
  Slot s = carve(global(P.scratch) + static_cast<uint64_t>(blockIdx.x) * P.slot_bytes, P.layout);
  __shared__ PoaResult res;
  if (threadIdx.x == 0) {
    res = PoaResult{};
    res.status = kPoaNotRun;
  }
  __syncthreads();
  process_window(P, s, lds, ((CWindow*)P.windows)[w], res);
  __syncthreads();
  if (threadIdx.x == 0) {
    P.results[w] = res;
  }
}

}  // namespace rh

namespace rh {

// Host launchers (kept in this translation unit with the kernels).
hipError_t RH_POA_SYM(launch_poa_batch)(const PoaParams* d_params, uint32_t lds_bytes,
                                        uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL(RH_POA_SYM(poa_batch_kernel), dim3(grid), dim3(64), lds_bytes, stream,
                     d_params);
  return hipGetLastError();
}

hipError_t RH_POA_SYM(launch_poa_single)(const PoaParams* d_params, uint32_t lds_bytes,
                                         uint32_t q, hipStream_t stream) {
  hipLaunchKernelGGL(RH_POA_SYM(poa_single_kernel), dim3(1), dim3(64), lds_bytes, stream,
                     d_params, q);
  return hipGetLastError();
}

hipError_t RH_POA_SYM(launch_poa_static)(const PoaParams* d_params, uint32_t lds_bytes,
                                         uint32_t first, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL(RH_POA_SYM(poa_static_kernel), dim3(count), dim3(64), lds_bytes, stream,
                     d_params, first);
  return hipGetLastError();
}

hipError_t RH_POA_SYM(poa_occupancy)(int* blocks_per_cu, uint32_t lds_bytes) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                      RH_POA_SYM(poa_batch_kernel), 64,
                                                      lds_bytes);
}

}  // namespace rh
