// Global (NW) edit-distance alignment of racon overlaps on gfx950.
//
// Reproduces the alignment path racon gets from edlib (EDLIB_MODE_NW,
// EDLIB_TASK_PATH; src/overlap.cpp:205-224). The path is defined by a
// recursion on the target: a subproblem (query span, target span, score)
// whose stored columns would take (2*8+4) * ceil(|q|/64) * |t| < 1 MiB is
// traced back directly from its last cell, preferring up (insertion), then
// left (deletion), then the diagonal; a larger one is split at column
// |t|/2 at the first query row x where the forward score D(q[..x], t[..|t|/2))
// plus the reverse score of the rest equals the subproblem score (then x = -1,
// then x = |q|-1), and the two halves are solved in order.
//
// Scores are computed in a diagonal band (Band below) that starts from a
// guess of the distance and doubles until the banded distance fits it; every
// score on an optimal path is then exact, which gives edlib's path: banding
// only drops cells whose score cannot be on an optimal path, and every choice
// above reads cells on optimal paths.
//
// Myers' bit-vector recurrence runs on 64-row blocks, one block per lane, in
// an anti-diagonal pipeline: at step s lane b advances column s - b and hands
// its horizontal carry to lane b + 1 through a lane shuffle. Queries longer
// than 64 blocks are processed in strips of 64 blocks whose bottom carries
// pass through the slot. One wave owns one overlap; a persistent grid takes
// overlaps from a work queue ordered by size.
#include <hip/hip_runtime.h>

#include "align_types.hpp"

namespace rh {

namespace {

constexpr int kWave = 64;
typedef uint64_t Word;

__device__ __forceinline__ int lane_id() { return threadIdx.x; }
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// One Myers/Hyyro block step with horizontal input hin in {-1, 0, +1}.
__device__ __forceinline__ int block_step(Word& P, Word& M, Word eq, int hin) {
  const Word hin_neg = hin < 0 ? Word(1) : Word(0);
  const Word xv = eq | M;
  eq |= hin_neg;
  const Word xh = (((eq & P) + P) ^ P) | eq;
  Word ph = M | ~(xh | P);
  Word mh = P & xh;
  int hout = 0;
  if (ph >> 63) hout = 1;
  if (mh >> 63) hout = -1;
  ph <<= 1;
  mh <<= 1;
  mh |= hin_neg;
  ph |= hin > 0 ? Word(1) : Word(0);
  P = mh | ~(xv | ph);
  M = ph & xv;
  return hout;
}

struct Seg {
  const uint8_t* q;  // query span
  const uint8_t* t;  // target span
  int ql;
  int tl;
  bool rev;  // read both spans from the end
  __device__ uint8_t qc(int i) const { return rev ? q[ql - 1 - i] : q[i]; }
  __device__ uint8_t tc(int j) const { return rev ? t[tl - 1 - j] : t[j]; }
};

// Match words of one block: for each of A, C, G, T the rows holding it (or
// past the query end, pad, which match everything), the rows holding N, and
// pad. A target A/C/G/T byte picks its word with two selects on its 2-bit
// code; other bytes take the N word or an exact byte compare (rare_eq).
struct Peq {
  Word a, c, g, t;  // match words of A, C, G, T (pad included)
  Word n;           // rows holding N
  Word pad;         // rows beyond the query
};

// 2-bit code of A, C, G, T (0..3) from the byte's bits 1 and 2; any byte
// maps to some code, so is_acgt checks the byte against the code's letter.
__device__ __forceinline__ int code2(int c) { return ((c >> 1) ^ (c >> 2)) & 3; }
__device__ __forceinline__ bool is_acgt(int c, int code) {
  return static_cast<int>((0x54474341u >> (8 * code)) & 0xffu) == c;  // "ACGT"
}

__device__ Peq build_peq(const Seg& g, int block) {
  Peq p{0, 0, 0, 0, 0, 0};
  Word lo = 0, hi = 0, acgt = 0;
  const int r0 = block * 64;
  for (int i = 0; i < 64; ++i) {
    const int r = r0 + i;
    const Word bit = Word(1) << i;
    if (r >= g.ql) {
      p.pad |= bit;
      continue;
    }
    const int c = g.qc(r);
    const int code = code2(c);
    if (is_acgt(c, code)) {
      acgt |= bit;
      if (code & 1) lo |= bit;
      if (code & 2) hi |= bit;
    } else if (c == 'N') {
      p.n |= bit;
    }
  }
  p.a = (acgt & ~lo & ~hi) | p.pad;
  p.c = (acgt & lo & ~hi) | p.pad;
  p.g = (acgt & ~lo & hi) | p.pad;
  p.t = (acgt & lo & hi) | p.pad;
  return p;
}

// Match word of an A/C/G/T byte from its 2-bit code.
__device__ __forceinline__ Word acgt_eq(const Peq& p, int code) {
  const Word e01 = (code & 1) ? p.c : p.a;
  const Word e23 = (code & 1) ? p.t : p.g;
  return (code & 2) ? e23 : e01;
}

// Match word for a target byte outside A/C/G/T/N (exact byte compare).
__device__ Word rare_eq(const Seg& g, int block, uint8_t c, Word pad) {
  Word w = pad;
  const int r0 = block * 64;
  for (int i = 0; i < 64; ++i) {
    const int r = r0 + i;
    if (r < g.ql && g.qc(r) == c) w |= Word(1) << i;
  }
  return w;
}

struct SlotA {
  int32_t* col_left;
  int32_t* col_right;
  int8_t* carry;
  Word* store_p;
  Word* store_m;
  int32_t* store_s;
  uint8_t* ops_tmp;
};

// Diagonal band of a (sub)problem: the cells (i, j) with lo <= i - j <= hi.
// A path through (i, j) costs at least |i - j| + |(ql - tl) - (i - j)|, so
// with d >= the subproblem's distance every optimal path lies inside
// band_for(ql, tl, d). Cells outside are not computed: a strip of blocks
// runs only over the columns where its rows meet the band, a strip that
// starts past column 0 begins from upper bounds (vertical deltas of +1 from
// max(row + 1, column + 1)), and a column the strip above did not compute
// feeds it a horizontal delta of +1. Every computed score is then an upper
// bound of the exact one, equal to it on every cell of an optimal path,
// which is all that the traceback and the split search read (a cell they
// look up outside the computed columns reads as kFar).
struct Band {
  int lo, hi;
};
constexpr int32_t kFar = 1 << 28;

__device__ __forceinline__ Band band_for(int ql, int tl, int d) {
  const int delta = ql - tl;
  const int ad = delta < 0 ? -delta : delta;
  const int slack = d >= ad ? (d - ad) / 2 : 0;
  return Band{min(0, delta) - slack, max(0, delta) + slack};
}

// The band of the reversed segment (row ql-1-i, column tl-1-j).
__device__ __forceinline__ Band band_reversed(Band b, int ql, int tl) {
  return Band{(ql - tl) - b.hi, (ql - tl) - b.lo};
}

// Columns [cs, ce] that strip k (rows k * 4096 .. + 64 * nbs - 1) computes.
__device__ __forceinline__ void strip_columns(const Seg& g, Band bd, int k, int nbs, int* cs,
                                              int* ce) {
  const int row_lo = k * kWave * 64;
  const int row_hi = min(g.ql, (k * kWave + nbs) * 64) - 1;
  *cs = max(0, row_lo - bd.hi);
  *ce = min(g.tl - 1, row_hi - bd.lo);
}

// Runs the bit-vector DP of seg over the columns of band bd. With stop >= 0
// (which must be the last column, tl - 1: the only column the recursion asks
// for), writes the scores of that column for rows 0..ql-1 into col, from each
// block's state after its last step (kFar for rows of a strip that does not
// reach it). With store, keeps (P, M, score) of every computed (column,
// block) for the traceback. Returns the score of the last cell (row ql-1,
// column tl-1) when stop < 0. *steps counts the wave's steps.
__device__ int forward(const Seg& g, int stop, int32_t* col, bool store, const SlotA& s, Band bd,
                       uint32_t* steps_done) {
  const int lane = lane_id();
  const int nb = (g.ql + 63) / 64;
  const int strips = (nb + kWave - 1) / kWave;
  int last_score = kFar;
  int prev_cs = 1, prev_ce = 0;  // columns whose bottom carries the strip above stored
  // the last row of the strip above at the column before this strip's first
  // (join_col): this strip's first column continues it by +1 per row
  __shared__ int join_score;
  int join_col = -1;
  for (int k = 0; k < strips; ++k) {
    const int b = k * kWave + lane;
    const int nbs = min(kWave, nb - k * kWave);  // blocks in this strip
    const bool active_lane = lane < nbs;
    int cs, ce;
    strip_columns(g, bd, k, nbs, &cs, &ce);
    cs = uniform(cs);
    ce = uniform(ce);
    // where the next strip starts (its join column is the one before)
    int next_cs = 0, next_ce = -1;
    if (k + 1 < strips) {
      strip_columns(g, bd, k + 1, min(kWave, nb - (k + 1) * kWave), &next_cs, &next_ce);
    }
    const int next_join = uniform(next_cs) - 1;
    if (cs > ce) {
      // the band misses every row of this strip
      if (stop >= 0 && active_lane) {
        for (int i = 0; i < 64; ++i) {
          const int r = b * 64 + i;
          if (r < g.ql) col[r] = kFar;
        }
      }
      prev_cs = 1;
      prev_ce = 0;
      join_col = -1;
      wave_sync();
      continue;
    }
    Peq peq{0, 0, 0, 0, 0, 0};
    if (active_lane) {
      peq = build_peq(g, b);
    }
    Word P = ~Word(0), M = 0;
    // D at the block's last row before column cs, with vertical deltas of +1
    // down the column: exact at column -1; further right, +1 per row from the
    // strip above's last row at column cs - 1 (an upper bound, and a column
    // whose deltas join the carries of the strip above), or from the bound
    // max(row, column) + 1 when the strip above computed nothing
    int score;
    if (cs == 0) {
      score = (b + 1) * 64;
    } else {
      const int top = join_col == cs - 1 ? join_score : max(k * kWave * 64, cs);
      score = top + (lane + 1) * 64;
    }
    wave_sync();  // (join_score read before this strip may write it)
    const int steps = ce + nbs;  // lane 0 runs columns cs..ce, the last lane ends nbs - 1 later
    *steps_done += static_cast<uint32_t>(steps - cs);
    const uint64_t store_base = static_cast<uint64_t>(k) * (g.tl + 63) * 64;
    // Lane b works on column c = st - b, the column lane b - 1 had one step
    // earlier: each lane hands its target byte and its bottom carry (hout + 1
    // in bits 8..9) to the next lane with one DPP wave shift. Lane 0 takes
    // its byte (and, below the first strip, its carry from the strip above)
    // from 64-column chunks that every lane loads one chunk ahead.
    auto t_at = [&](int j) -> int { return j < g.tl ? g.tc(j) : 0; };
    // (a column the strip above did not compute: +1, an upper bound)
    auto carry_at = [&](int j) -> int {
      return k == 0 ? 1 : (j >= prev_cs && j <= prev_ce ? static_cast<int>(s.carry[j]) : 1);
    };
    // (an outer loop per chunk: a loaded chunk is only read in the chunk
    // after the one that issued its load, and no loaded value is carried
    // through the step loop, which would make every step wait for memory)
    const int st_begin = cs & ~(kWave - 1);
    int t_next = t_at(st_begin + lane);
    int c_next = carry_at(st_begin + lane);
    const bool carry_lane = lane == nbs - 1 && k + 1 < strips;
    bool prev_acgt = true;  // (no columns before the first chunk)
    int fwd = 1 << 8;  // (byte, hout + 1) of this lane's last step
    for (int st0 = st_begin; st0 < steps; st0 += kWave) {
     const int t_cur = t_next, c_cur = c_next;
     // this chunk's bytes and carries (loaded a chunk ago) are waited for
     // here, once, before the next chunk's loads go out
     asm volatile("" ::"v"(t_cur), "v"(c_cur));
     // lanes work on columns st0 - 63 .. st0 + 63 in this chunk: when those
     // are all A/C/G/T and every active lane's column is inside the matrix
     // on every step, the steps run without the range and byte checks
     const bool mine_acgt = st0 + lane >= g.tl || is_acgt(t_cur, code2(t_cur));
     const bool chunk_acgt = __all(mine_acgt);
     const int st_end = min(steps, st0 + kWave);
     const bool fast = chunk_acgt && prev_acgt && st0 >= cs + nbs - 1 && st_end - 1 <= ce;
     prev_acgt = chunk_acgt;
     t_next = t_at(st0 + kWave + lane);
     c_next = carry_at(st0 + kWave + lane);
     if (fast) {
      for (int st = st0; st < st_end; ++st) {
       const int c = st - lane;
       const int up = __builtin_amdgcn_update_dpp(0, fwd, 0x138, 0xf, 0xf, false);
       const int t0 = __builtin_amdgcn_readlane(t_cur, st & (kWave - 1));
       const int h0 = __builtin_amdgcn_readlane(c_cur, st & (kWave - 1));
       const uint8_t tc = static_cast<uint8_t>(lane == 0 ? t0 : (up & 0xff));
       const int hin = lane == 0 ? h0 : (up >> 8) - 1;
       // (lanes past the strip compute unused values; they store nothing and
       // feed only lanes past the strip)
       const Word eq = acgt_eq(peq, code2(tc));
       const int hout = block_step(P, M, eq, hin);
       score += hout;
       if (store && active_lane) {
         const uint64_t idx = store_base + static_cast<uint64_t>(st) * nbs + lane;
         s.store_p[idx] = P;
         s.store_m[idx] = M;
         s.store_s[idx] = score;
       }
       if (carry_lane) {
         s.carry[c] = static_cast<int8_t>(hout);
         if (c == next_join) join_score = score;
       }
       fwd = static_cast<int>(tc) | ((hout + 1) << 8);
      }
      continue;
     }
     for (int st = st0; st < st_end; ++st) {
      const int c = st - lane;
      const int up = __builtin_amdgcn_update_dpp(0, fwd, 0x138, 0xf, 0xf, false);
      // lane 0's inputs are uniform reads; selected, not branched to
      const int t0 = __builtin_amdgcn_readlane(t_cur, st & (kWave - 1));
      const int h0 = __builtin_amdgcn_readlane(c_cur, st & (kWave - 1));
      const uint8_t tc = static_cast<uint8_t>(lane == 0 ? t0 : (up & 0xff));
      const int hin = lane == 0 ? h0 : (up >> 8) - 1;
      int hout = 0;
      if (active_lane && c >= cs && c <= ce) {
        // rows equal to the target byte (exact byte equality, as edlib's
        // alphabet without equalities)
        const int code = code2(tc);
        Word eq = acgt_eq(peq, code);
        if (!is_acgt(tc, code)) {
          eq = tc == 'N' ? (peq.n | peq.pad) : rare_eq(g, b, tc, peq.pad);
        }
        hout = block_step(P, M, eq, hin);
        score += hout;
        if (store) {
          const uint64_t idx = store_base + static_cast<uint64_t>(st) * nbs + lane;
          s.store_p[idx] = P;
          s.store_m[idx] = M;
          s.store_s[idx] = score;
        }
        if (carry_lane) {
          s.carry[c] = static_cast<int8_t>(hout);
          if (c == next_join) join_score = score;
        }
      }
      fwd = static_cast<int>(tc) | ((hout + 1) << 8);
     }
    }
    const bool reaches_end = ce == g.tl - 1;
    if (stop >= 0 && active_lane) {
      // this block's rows at the last column
      int v = score;
      for (int i = 63; i >= 0; --i) {
        const int r = b * 64 + i;
        if (r < g.ql) col[r] = reaches_end ? v : kFar;
        if ((P >> i) & 1) v -= 1;
        if ((M >> i) & 1) v += 1;
      }
    }
    // score of the last row (in the last strip's last block)
    if (k == strips - 1 && reaches_end) {
      const int last_block = nb - 1;
      const int last_bit = (g.ql - 1) % 64;
      int v = score;
      for (int i = 63; i > last_bit; --i) {
        if ((P >> i) & 1) v -= 1;
        if ((M >> i) & 1) v += 1;
      }
      last_score = __shfl(v, last_block - k * kWave, kWave);
    }
    prev_cs = cs;
    prev_ce = ce;
    join_col = next_join >= cs && next_join <= ce ? next_join : -1;
    wave_sync();  // carries, stored columns and join_score visible to every lane
  }
  return last_score;
}

// Lane l's 64-bit value, as a uniform value.
__device__ __forceinline__ Word readlane64(Word x, int l) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), l));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x >> 32), l));
  return (static_cast<Word>(hi) << 32) | lo;
}

// Traceback of a stored leaf; writes ops (forward order) at out, returns count.
__device__ int leaf_traceback(const Seg& g, const SlotA& s, uint8_t* out, Band bd) {
  // The walk is uniform (every lane follows it); the stored words it reads
  // come from a window loaded by all lanes in one round trip: lane l holds
  // column cw - l of block bw (the block of row r) and of block bw - 1, so
  // the walk reads them with readlane until it leaves the window (64
  // columns, the two blocks), then reloads. The values and the move order
  // are those of reading every cell's stored words directly.
  const int lane = lane_id();
  const int nb = (g.ql + 63) / 64;
  auto index_of = [&](int b, int c) -> uint64_t {
    const int k = b / kWave;
    const int lb = b % kWave;
    const int nbs = min(kWave, nb - k * kWave);
    return static_cast<uint64_t>(k) * (g.tl + 63) * 64 +
           static_cast<uint64_t>(c + lb) * nbs + lb;
  };
  int r = g.ql - 1, c = g.tl - 1;
  int n = 0;
  uint8_t* tmp = s.ops_tmp;
  while (r >= 0 && c >= 0) {
    const int bw = uniform(r / 64), cw = uniform(c);
    const int col = cw - lane;
    Word pb = 0, mb = 0, pa = 0, ma = 0;
    int sb = 0, sa = 0;
    if (col >= 0) {
      const uint64_t ib = index_of(bw, col);
      pb = s.store_p[ib];
      mb = s.store_m[ib];
      sb = s.store_s[ib];
      if (bw >= 1) {
        const uint64_t ia = index_of(bw - 1, col);
        pa = s.store_p[ia];
        ma = s.store_m[ia];
        sa = s.store_s[ia];
      }
    }
    auto val = [&](int rr, int cc) -> int {
      if (rr < 0) return cc + 1;
      if (cc < 0) return rr + 1;
      {
        // a column the cell's strip did not compute: off every optimal path
        const int k = rr / (64 * kWave);
        int cs, ce;
        strip_columns(g, bd, k, min(kWave, nb - k * kWave), &cs, &ce);
        if (cc < cs || cc > ce) return kFar;
      }
      const int l = cw - cc;
      const bool above = rr / 64 < bw;
      const Word P = readlane64(above ? pa : pb, l);
      const Word M = readlane64(above ? ma : mb, l);
      int v = above ? __builtin_amdgcn_readlane(sa, l) : __builtin_amdgcn_readlane(sb, l);
      const int bit = rr % 64;
      if (bit < 63) {
        const Word mask = ~Word(0) << (bit + 1);
        v -= __popcll(P & mask);
        v += __popcll(M & mask);
      }
      return v;
    };
    while (r >= 0 && c >= 0 && r / 64 >= bw && c >= cw - 62) {
      const int cur = val(r, c);
      const int up = val(r - 1, c);
      if (up + 1 == cur) {
        tmp[n++] = 1;  // insertion
        --r;
        continue;
      }
      const int left = val(r, c - 1);
      if (left + 1 == cur) {
        tmp[n++] = 2;  // deletion
        --c;
        continue;
      }
      const int diag = val(r - 1, c - 1);
      tmp[n++] = diag == cur ? 0 : 3;
      --r;
      --c;
    }
  }
  while (c >= 0) {
    tmp[n++] = 2;
    --c;
  }
  while (r >= 0) {
    tmp[n++] = 1;
    --r;
  }
  wave_sync();
  for (int i = lane_id(); i < n; i += kWave) {
    out[i] = tmp[n - 1 - i];
  }
  wave_sync();
  return n;
}

struct Task {
  int q0, ql, t0, tl, score;
};

__device__ void align_pair(const AlignParams& A, const AlignPair& pr, AlignResult& R) {
  const int lane = lane_id();
  uint8_t* slot = A.scratch + static_cast<uint64_t>(blockIdx.x) * A.slot_bytes;
  SlotA s;
  s.col_left = reinterpret_cast<int32_t*>(slot + A.layout.col_left);
  s.col_right = reinterpret_cast<int32_t*>(slot + A.layout.col_right);
  s.carry = reinterpret_cast<int8_t*>(slot + A.layout.carry);
  s.store_p = reinterpret_cast<Word*>(slot + A.layout.store_p);
  s.store_m = reinterpret_cast<Word*>(slot + A.layout.store_m);
  s.store_s = reinterpret_cast<int32_t*>(slot + A.layout.store_s);
  s.ops_tmp = slot + A.layout.ops_tmp;

  const uint8_t* Q = A.arena + pr.q_off;
  const uint8_t* T = A.arena + pr.t_off;
  uint8_t* out = A.ops + pr.out_off;
  const int QL = static_cast<int>(pr.q_len), TL = static_cast<int>(pr.t_len);
  if (QL > static_cast<int>(A.caps.q) || TL > static_cast<int>(A.caps.t)) {
    R.status = kAlignScratch;
    return;
  }

  auto is_leaf = [](int ql, int tl) {
    const uint64_t nb = (static_cast<uint64_t>(ql) + 63) / 64;
    return 20ull * nb * static_cast<uint64_t>(tl) < kLeafBytes;
  };

  __shared__ Task stack[kAlignStackDepth];
  int top = 0;
  int n_ops = 0;
  int best = -1;

  if (QL == 0 || TL == 0) {
    for (int i = lane; i < QL + TL; i += kWave) out[i] = QL == 0 ? 2 : 1;
    R.n_ops = static_cast<uint32_t>(QL + TL);
    R.edit_distance = QL > TL ? QL : TL;
    R.status = kAlignOk;
    return;
  }
  // The whole pair's distance is not known: its band starts from a guess k
  // (the batch's band width, else |Δ| + max(|q|, |t|) / 5) and doubles until
  // the banded distance is at most k (then it is exact: with k >= the true
  // distance every optimal path lies inside the band) or the band holds the
  // whole matrix. Every subproblem below knows its score, and its band is
  // the tightest one for it.
  const int delta_top = QL > TL ? QL - TL : TL - QL;
  int k_top = A.band0 > 0 ? max(static_cast<int>(A.band0), delta_top)
                          : delta_top + max(QL, TL) / 5 + 64;
  auto full_band = [](Band b, int ql, int tl) { return b.lo <= -(tl - 1) && b.hi >= ql - 1; };
  uint32_t steps = 0;

  if (is_leaf(QL, TL)) {
    Seg g{Q, T, QL, TL, false};
    Band bd;
    while (true) {
      bd = band_for(QL, TL, k_top);
      best = forward(g, -1, nullptr, true, s, bd, &steps);
      if (best <= k_top || full_band(bd, QL, TL)) break;
      k_top *= 2;
    }
    n_ops = leaf_traceback(g, s, out, bd);
    R.n_ops = static_cast<uint32_t>(n_ops);
    R.edit_distance = best;
    R.status = kAlignOk;
    R.pad = steps;
    return;
  }
  // top level: the Hirschberg split also yields the distance
  if (lane == 0) {
    stack[0] = Task{0, QL, 0, TL, -1};
  }
  top = 1;
  wave_sync();
  while (top > 0) {
    const Task tk = stack[top - 1];
    wave_sync();
    --top;
    const int ql = tk.ql, tl = tk.tl;
    if (ql == 0 || tl == 0) {
      for (int i = lane; i < ql + tl; i += kWave) out[n_ops + i] = ql == 0 ? 2 : 1;
      n_ops += ql + tl;
      wave_sync();
      continue;
    }
    if (is_leaf(ql, tl)) {
      Seg g{Q + tk.q0, T + tk.t0, ql, tl, false};
      const Band bd = band_for(ql, tl, tk.score);
      forward(g, -1, nullptr, true, s, bd, &steps);
      n_ops += leaf_traceback(g, s, out + n_ops, bd);
      continue;
    }
    const int lw = tl / 2, rw = tl - lw;
    Seg gl{Q + tk.q0, T + tk.t0, ql, lw, false};
    Seg gr{Q + tk.q0, T + tk.t0 + lw, ql, rw, true};
    int score = tk.score;
    while (true) {
      const Band bd = band_for(ql, tl, score >= 0 ? score : k_top);
      // forward over the left half: scores of column lw-1
      forward(gl, lw - 1, s.col_left, false, s, bd, &steps);
      // reverse over the right half: col_right[r'] = ED of the last r'+1
      // query bytes against t[lw..)
      forward(gr, rw - 1, s.col_right, false, s, band_reversed(bd, ql, tl), &steps);
      wave_sync();
      if (score >= 0) break;
      // right(x) = ED(q[x..], t[lw..]) = col_right[ql-1-x] for x < ql, rw at
      // ql; the distance of the whole pair: min over all split rows
      int m = 0x7fffffff;
      for (int x = lane - 1; x <= ql - 1; x += kWave) {
        const int lv = x < 0 ? lw : s.col_left[x];
        const int rv = x + 1 >= ql ? rw : s.col_right[ql - 2 - x];
        m = min(m, lv + rv);
      }
      for (int d = 1; d < kWave; d <<= 1) m = min(m, __shfl_xor(m, d, kWave));
      m = uniform(m);
      if (m <= k_top || full_band(bd, ql, tl)) {
        score = m;
        best = score;
        break;
      }
      k_top *= 2;
    }
    // first interior split row x in [0, ql-2]
    int split = -2;
    for (int x0 = 0; x0 <= ql - 2 && split == -2; x0 += kWave) {
      const int x = x0 + lane;
      bool hit = false;
      if (x <= ql - 2) {
        hit = s.col_left[x] + s.col_right[ql - 2 - x] == score;
      }
      const unsigned long long mask = __ballot(hit);
      if (mask != 0) {
        split = x0 + __ffsll(static_cast<long long>(mask)) - 1;
      }
    }
    int ls = 0, rs = 0;
    if (split != -2) {
      ls = s.col_left[split];
      rs = s.col_right[ql - 2 - split];
    } else if (lw + s.col_right[ql - 1] == score) {
      split = -1;
      ls = lw;
      rs = s.col_right[ql - 1];
    } else if (s.col_left[ql - 1] + rw == score) {
      split = ql - 1;
      ls = s.col_left[ql - 1];
      rs = rw;
    } else {
      R.status = kAlignNoSplit;
      return;
    }
    split = uniform(split);
    ls = uniform(ls);
    rs = uniform(rs);
    if (top + 2 > static_cast<int>(kAlignStackDepth)) {
      R.status = kAlignStack;
      return;
    }
    wave_sync();
    if (lane == 0) {
      stack[top] = Task{tk.q0 + split + 1, ql - split - 1, tk.t0 + lw, rw, rs};
      stack[top + 1] = Task{tk.q0, split + 1, tk.t0, lw, ls};
    }
    top += 2;
    wave_sync();
  }
  R.n_ops = static_cast<uint32_t>(n_ops);
  R.edit_distance = best;
  R.status = kAlignOk;
  R.pad = steps;
}

}  // namespace

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void align_batch_kernel(
    AlignParams A) {
  __shared__ uint32_t next;
  while (true) {
    if (threadIdx.x == 0) {
      next = atomicAdd(A.queue_head, 1u);
    }
    __syncthreads();
    const uint32_t q = next;
    __syncthreads();
    if (q >= A.n_order) {
      break;
    }
    const uint32_t i = A.order != nullptr ? A.order[q] : q;
    AlignResult res{0, -1, kAlignNotRun, 0};
    align_pair(A, A.pairs[i], res);
    __syncthreads();
    if (threadIdx.x == 0) {
      A.results[i] = res;
    }
    __syncthreads();
  }
}

hipError_t launch_align_batch(const AlignParams& params, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL(align_batch_kernel, dim3(grid), dim3(64), 0, stream, params);
  return hipGetLastError();
}

// Clears the work-queue counter with a kernel store (no memset node, so the
// sequence can sit inside a captured graph).
__global__ void clear_counter_kernel(uint32_t* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

hipError_t launch_align_batch_cleared(const AlignParams& params, uint32_t grid,
                                      hipStream_t stream) {
  hipLaunchKernelGGL(clear_counter_kernel, dim3(1), dim3(64), 0, stream, params.queue_head);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_align_batch(params, grid, stream);
}

// Inclusive scans over the 64 lanes (lane shuffles; small kernels only).
__device__ __forceinline__ int32_t lane_scan_add(int32_t x) {
  const int lane = lane_id() & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int32_t y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}
__device__ __forceinline__ int32_t lane_scan_max(int32_t x) {
  const int lane = lane_id() & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int32_t y = __shfl_up(x, d, kWave);
    if (lane >= d) x = max(x, y);
  }
  return x;
}

// Breaking points of every aligned pair (racon overlap.cpp:226-292, the walk
// over edlib's CIGAR), from the edit operations the batch kernel left in
// device memory. A match or mismatch advances target and query, an insertion
// the query, a deletion the target; the first match of a window and the
// position after its last match are emitted as a pair when the target
// position reaches the window's last base (k * W - 1 for the multiples of W
// inside (t_begin, t_end), then t_end - 1). counts[i] is the number of points
// (2 per window with a match), or 0xffffffff when they would outgrow the
// pair's room.
//
// One wave per pair walks the ops 256 at a time (4 per lane, adjacent bytes
// across lanes): prefix sums of the target and query advances give every
// op's positions, a prefix max the last match so far (matches advance the
// target, so the last one has the largest position), and windows follow
// from target positions: window w covers (end(w - 1), end(w)]. A match whose
// window differs from the previous match's opens that window's entry (first
// match); an op that reaches end(w) closes it when the last match so far lies
// in w. Entry e of a pair is points 4e..4e+3 of its room: windows with a
// match, in order.
__global__ __launch_bounds__(256) void breaking_points_kernel(
    const AlignPair* __restrict__ pairs, const AlignResult* __restrict__ results,
    const uint8_t* __restrict__ ops, const BreakSpan* __restrict__ spans, uint32_t n,
    uint32_t window_length, uint32_t* __restrict__ points, uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (i >= n) return;  // (uniform over the wave)
  const BreakSpan sp = spans[i];
  if (results[i].status != kAlignOk) {
    if (lane == 0) counts[i] = 0xfffffffeu;  // not aligned on the device
    return;
  }
  const uint8_t* op = ops + pairs[i].out_off;
  const uint32_t n_ops = results[i].n_ops;
  const int64_t W = window_length;
  // windows: ends m0 - 1, m0 + W - 1, ... for the nm multiples of W in
  // (t_begin, t_end), then t_end - 1 (window nm)
  const int64_t t_end = sp.t_end;
  const int64_t m0_wide = (static_cast<int64_t>(sp.t_begin) / W + 1) * W;
  const uint32_t nm = m0_wide < t_end ? static_cast<uint32_t>((t_end - 1 - m0_wide) / W + 1) : 0;
  const uint32_t m0 = nm ? static_cast<uint32_t>(m0_wide) : 0;  // < t_end when used
  const uint32_t Wn = window_length;
  // (target positions here are in [0, t_end), so 32-bit arithmetic holds)
  auto window_of = [&](int64_t t) -> uint32_t {
    const uint32_t tt = static_cast<uint32_t>(t);
    const uint32_t w = (nm == 0 || tt < m0) ? 0 : (tt - m0) / Wn + 1;
    return w < nm ? w : nm;
  };
  auto window_end = [&](uint32_t w) -> int64_t {
    return w < nm ? static_cast<int64_t>(m0) + static_cast<int64_t>(w) * W - 1 : t_end - 1;
  };
  uint32_t* out = points + 2 * sp.out;
  const int64_t cap_entries = sp.cap / 2;  // entries that fit the pair's room
  // carried over chunks (uniform)
  int64_t t_base = static_cast<int64_t>(sp.t_begin) - 1;
  int64_t q_base = static_cast<int64_t>(sp.q_first) - 1;
  int64_t last_t = -1, last_q = -1;  // last match so far (t, q after the op)
  int64_t opened = 0;                // entries opened so far
  int64_t closed = 0;                // entries closed so far
  for (uint32_t c0 = 0; c0 < n_ops; c0 += 4 * kWave) {
    uint32_t o[4];
    int32_t dt = 0, dq = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t k = c0 + 4 * lane + u;
      o[u] = k < n_ops ? (op[k] & 3u) : 4u;  // 4: past the end (no op)
      dt += (o[u] != 1u && o[u] != 4u);
      dq += (o[u] != 2u && o[u] != 4u);
    }
    const int32_t st = lane_scan_add(dt), sq = lane_scan_add(dq);
    int64_t t = t_base + (st - dt), q = q_base + (sq - dq);
    // positions after each of the lane's ops, and the lane's last match
    int64_t ta[4], qa[4];
    int32_t lm_t = -1, lm_q = -1;  // relative to t_base / q_base + 1
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (o[u] != 1u && o[u] != 4u) ++t;
      if (o[u] != 2u && o[u] != 4u) ++q;
      ta[u] = t;
      qa[u] = q;
      if (o[u] == 0u || o[u] == 3u) {
        lm_t = static_cast<int32_t>(t - t_base);
        lm_q = static_cast<int32_t>(q - q_base);
      }
    }
    // last match before the lane (exclusive max over lanes, then the carry)
    const int32_t mt = lane_scan_max(lm_t), mq = lane_scan_max(lm_q);
    const int32_t pt = __shfl_up(mt, 1, kWave), pq = __shfl_up(mq, 1, kWave);
    int64_t prev_t = last_t, prev_q = last_q;
    if (lane > 0 && pt >= 0) {
      prev_t = t_base + pt;
      prev_q = q_base + pq;
    }
    // windows of the lane's target advances (-1: none), and of the last
    // match before each op
    int64_t wa[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) wa[u] = (o[u] != 1u && o[u] != 4u) ? window_of(ta[u]) : -1;
    const int64_t pw0 = prev_t >= 0 ? window_of(prev_t) : -1;
    // entries the lane opens, then their numbers
    int32_t n_open = 0;
    {
      int64_t pw = pw0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (o[u] == 0u || o[u] == 3u) {
          if (wa[u] != pw) ++n_open;
          pw = wa[u];
        }
      }
    }
    const int32_t so = lane_scan_add(n_open);
    int64_t e_open = opened + (so - n_open);  // entries opened before this lane
    int32_t n_close = 0;
    int64_t pw = pw0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (o[u] == 0u || o[u] == 3u) {
        if (wa[u] != pw) {
          // first match of window w: entry e_open
          if (e_open < cap_entries) {
            out[4 * e_open] = static_cast<uint32_t>(ta[u]);
            out[4 * e_open + 1] = static_cast<uint32_t>(qa[u]);
          }
          ++e_open;
        }
        prev_t = ta[u];
        prev_q = qa[u];
        pw = wa[u];
      }
      if (wa[u] >= 0) {
        if (ta[u] == window_end(static_cast<uint32_t>(wa[u])) && pw == wa[u]) {
          // the op reaches the window's last base with a match in the window:
          // close its entry (the last one opened)
          const int64_t e = e_open - 1;
          if (e < cap_entries) {
            out[4 * e + 2] = static_cast<uint32_t>(prev_t + 1);
            out[4 * e + 3] = static_cast<uint32_t>(prev_q + 1);
          }
          ++n_close;
        }
      }
    }
    const int32_t sc = lane_scan_add(n_close);
    // carry to the next chunk (from the last lane)
    const int32_t all_t = __shfl(st, kWave - 1, kWave), all_q = __shfl(sq, kWave - 1, kWave);
    const int32_t all_mt = __shfl(mt, kWave - 1, kWave), all_mq = __shfl(mq, kWave - 1, kWave);
    if (all_mt >= 0) {
      last_t = t_base + all_mt;
      last_q = q_base + all_mq;
    }
    t_base += all_t;
    q_base += all_q;
    opened += __shfl(so, kWave - 1, kWave);
    closed += __shfl(sc, kWave - 1, kWave);
  }
  if (lane == 0) {
    counts[i] = 2 * closed > sp.cap ? 0xffffffffu : static_cast<uint32_t>(2 * closed);
  }
}

hipError_t launch_breaking_points(const AlignPair* pairs, const AlignResult* results,
                                  const uint8_t* ops, const BreakSpan* spans, uint32_t n,
                                  uint32_t window_length, uint32_t* points, uint32_t* counts,
                                  hipStream_t stream) {
  // one wave per pair, four pairs per workgroup
  hipLaunchKernelGGL(breaking_points_kernel, dim3((n + 3) / 4), dim3(256), 0, stream, pairs,
                     results, ops, spans, n, window_length, points, counts);
  return hipGetLastError();
}

hipError_t align_occupancy(int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, align_batch_kernel, 64, 0);
}

}  // namespace rh
