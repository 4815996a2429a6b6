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
constexpr int kStrip = 8;                  // DP columns per lane per tile
constexpr int kTile = kWave * kStrip;      // 512 columns per tile
constexpr int32_t kNegInf = -(1 << 29);
constexpr uint16_t kCodeDiag = 0x0000;
constexpr uint16_t kCodeVert = 0x4000;
constexpr uint16_t kCodeHorz = 0x8000;

__device__ __forceinline__ int lane_id() { return threadIdx.x; }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// Views of one slot's arrays.
struct Slot {
  uint8_t* base;
  int32_t* in_head;
  int32_t* out_head;
  int32_t* in_tail;
  int32_t* out_tail;
  uint8_t* n_aligned;
  int32_t* aligned;
  int32_t* cov;
  int32_t* last_label;
  int32_t* e_tail;
  int32_t* e_head;
  int32_t* e_weight;
  int32_t* e_next_in;
  int32_t* e_next_out;
  int32_t* rank_to_node;
  int32_t* node_rank;
  uint8_t* mark;
  uint8_t* lead;
  int32_t* stack;
  uint8_t* in_sub;
  int32_t* row_of;
  int32_t* sub_order;
  int32_t* pred_off;
  int32_t* pred_rows;
  uint8_t* sink;
  int32_t* H;
  uint16_t* code;
  int32_t* aln_node;
  int32_t* aln_pos;
  int64_t* scores;
  int32_t* preds;
  int32_t* path;
  uint8_t* seq;
  int32_t* qual;  // per-base weights of the current layer

  int32_t n_nodes;
  int32_t n_edges;
  int32_t n_seqs;
  bool overflow;
};

__device__ Slot carve(uint8_t* p, const PoaSlotLayout& l) {
  Slot s;
  s.base = p + l.base;
  s.in_head = reinterpret_cast<int32_t*>(p + l.in_head);
  s.out_head = reinterpret_cast<int32_t*>(p + l.out_head);
  s.in_tail = reinterpret_cast<int32_t*>(p + l.in_tail);
  s.out_tail = reinterpret_cast<int32_t*>(p + l.out_tail);
  s.n_aligned = p + l.n_aligned;
  s.aligned = reinterpret_cast<int32_t*>(p + l.aligned);
  s.cov = reinterpret_cast<int32_t*>(p + l.cov);
  s.last_label = reinterpret_cast<int32_t*>(p + l.last_label);
  s.e_tail = reinterpret_cast<int32_t*>(p + l.e_tail);
  s.e_head = reinterpret_cast<int32_t*>(p + l.e_head);
  s.e_weight = reinterpret_cast<int32_t*>(p + l.e_weight);
  s.e_next_in = reinterpret_cast<int32_t*>(p + l.e_next_in);
  s.e_next_out = reinterpret_cast<int32_t*>(p + l.e_next_out);
  s.rank_to_node = reinterpret_cast<int32_t*>(p + l.rank_to_node);
  s.node_rank = reinterpret_cast<int32_t*>(p + l.node_rank);
  s.mark = p + l.mark;
  s.lead = p + l.lead;
  s.stack = reinterpret_cast<int32_t*>(p + l.stack);
  s.in_sub = p + l.in_sub;
  s.row_of = reinterpret_cast<int32_t*>(p + l.row_of);
  s.sub_order = reinterpret_cast<int32_t*>(p + l.sub_order);
  s.pred_off = reinterpret_cast<int32_t*>(p + l.pred_off);
  s.pred_rows = reinterpret_cast<int32_t*>(p + l.pred_rows);
  s.sink = p + l.sink;
  s.H = reinterpret_cast<int32_t*>(p + l.H);
  s.code = reinterpret_cast<uint16_t*>(p + l.code);
  s.aln_node = reinterpret_cast<int32_t*>(p + l.aln_node);
  s.aln_pos = reinterpret_cast<int32_t*>(p + l.aln_pos);
  s.scores = reinterpret_cast<int64_t*>(p + l.scores);
  s.preds = reinterpret_cast<int32_t*>(p + l.preds);
  s.path = reinterpret_cast<int32_t*>(p + l.path);
  s.seq = p + l.seq;
  s.qual = reinterpret_cast<int32_t*>(p + l.qual);
  s.n_nodes = 0;
  s.n_edges = 0;
  s.n_seqs = 0;
  s.overflow = false;
  return s;
}

// ---- graph construction (lane-uniform) ----

__device__ void touch(Slot& s, int node, int label) {
  if (s.last_label[node] != label) {
    s.last_label[node] = label;
    s.cov[node] = s.cov[node] + 1;
  }
}

__device__ int add_node(Slot& s, const PoaSlotCaps& caps, uint8_t base) {
  if (s.n_nodes >= static_cast<int>(caps.nodes)) {
    s.overflow = true;
    return 0;
  }
  const int v = s.n_nodes++;
  s.base[v] = base;
  s.in_head[v] = -1;
  s.out_head[v] = -1;
  s.in_tail[v] = -1;
  s.out_tail[v] = -1;
  s.n_aligned[v] = 0;
  s.cov[v] = 0;
  s.last_label[v] = -1;
  return v;
}

__device__ void add_edge(Slot& s, const PoaSlotCaps& caps, int tail, int head, int weight) {
  for (int e = s.out_head[tail]; e != -1; e = s.e_next_out[e]) {
    if (s.e_head[e] == head) {
      s.e_weight[e] = s.e_weight[e] + weight;
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
  s.e_tail[e] = tail;
  s.e_head[e] = head;
  s.e_weight[e] = weight;
  s.e_next_in[e] = -1;
  s.e_next_out[e] = -1;
  if (s.out_tail[tail] == -1) {
    s.out_head[tail] = e;
  } else {
    s.e_next_out[s.out_tail[tail]] = e;
  }
  s.out_tail[tail] = e;
  if (s.in_tail[head] == -1) {
    s.in_head[head] = e;
  } else {
    s.e_next_in[s.in_tail[head]] = e;
  }
  s.in_tail[head] = e;
  touch(s, tail, s.n_seqs);
  touch(s, head, s.n_seqs);
}

// Chain of new nodes for seq[begin, end); returns the first node or -1.
__device__ int add_chain(Slot& s, const PoaSlotCaps& caps, int begin, int end) {
  if (begin >= end) {
    return -1;
  }
  const int first = add_node(s, caps, s.seq[begin]);
  for (int i = begin + 1; i < end && !s.overflow; ++i) {
    const int v = add_node(s, caps, s.seq[i]);
    add_edge(s, caps, v - 1, v, s.qual[i - 1] + s.qual[i]);
  }
  return first;
}

// DFS topological order over nodes with keep[v] (all nodes when keep is
// null); writes node ids in rank order to out, returns how many.
__device__ int topo_sort(Slot& s, const uint8_t* keep, int32_t* out, int stack_cap) {
  const int n = s.n_nodes;
  for (int v = lane_id(); v < n; v += kWave) {
    s.mark[v] = 0;
    s.lead[v] = 1;
  }
  wave_sync();
  int count = 0;
  for (int root = 0; root < n; ++root) {
    if (keep != nullptr && keep[root] == 0) {
      continue;
    }
    if (s.mark[root] != 0) {
      continue;
    }
    int top = 0;
    s.stack[top++] = root;
    while (top > 0) {
      const int v = uniform(s.stack[top - 1]);
      bool ready = true;
      if (s.mark[v] != 2) {
        for (int e = uniform(s.in_head[v]); e != -1; e = uniform(s.e_next_in[e])) {
          const int u = uniform(s.e_tail[e]);
          if (keep != nullptr && keep[u] == 0) {
            continue;
          }
          if (s.mark[u] != 2) {
            if (top >= stack_cap) {
              s.overflow = true;
              return count;
            }
            s.stack[top++] = u;
            ready = false;
          }
        }
        if (s.lead[v] != 0) {
          const int na = s.n_aligned[v];
          for (int k = 0; k < na; ++k) {
            const int a = uniform(s.aligned[v * kMaxAligned + k]);
            if (keep != nullptr && keep[a] == 0) {
              continue;
            }
            if (s.mark[a] != 2) {
              if (top >= stack_cap) {
                s.overflow = true;
                return count;
              }
              s.stack[top++] = a;
              s.lead[a] = 0;
              ready = false;
            }
          }
        }
        if (ready) {
          s.mark[v] = 2;
          if (s.lead[v] != 0) {
            if (count + 1 + s.n_aligned[v] > n) {
              s.overflow = true;
              return count;
            }
            out[count++] = v;
            const int na = s.n_aligned[v];
            for (int k = 0; k < na; ++k) {
              const int a = uniform(s.aligned[v * kMaxAligned + k]);
              if (keep == nullptr || keep[a] != 0) {
                out[count++] = a;
              }
            }
          }
        } else {
          s.mark[v] = 1;
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

__device__ __forceinline__ int32_t wave_prefix_max(int32_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int32_t y = __shfl_up(x, d, kWave);
    if (lane >= d) {
      x = max(x, y);
    }
  }
  return x;
}

// One DP row: row r (1-based) of node base b with predecessor rows
// preds[0..np). H rows have pitch `pitch`; column j of a row is at index j.
__device__ void dp_row(const Slot& s, int r, uint8_t b, const int32_t* preds, int np,
                       int L, int pitch, int32_t match, int32_t mismatch, int32_t gap) {
  const int lane = lane_id();
  int32_t* Hr = s.H + static_cast<int64_t>(r) * pitch;
  uint16_t* Cr = s.code + static_cast<int64_t>(r) * pitch;
  int32_t carry = Hr[0];  // column 0, set by the column pass (u_0 = H[r][0])
  for (int t0 = 1; t0 <= L; t0 += kTile) {
    const int j0 = t0 + lane * kStrip;
    int32_t sc[kStrip];
#pragma unroll
    for (int u = 0; u < kStrip; ++u) {
      const int j = j0 + u;
      sc[u] = (j <= L && s.seq[j - 1] == b) ? match : mismatch;
    }
    int32_t t[kStrip];
#pragma unroll
    for (int u = 0; u < kStrip; ++u) {
      t[u] = kNegInf;
    }
    for (int k = 0; k < np; ++k) {
      const int32_t* Hp = s.H + static_cast<int64_t>(preds[k]) * pitch;
      int32_t left = j0 - 1 <= L ? Hp[j0 - 1] : kNegInf;
#pragma unroll
      for (int u = 0; u < kStrip; ++u) {
        const int j = j0 + u;
        const int32_t up = j <= L ? Hp[j] : kNegInf;
        t[u] = max(t[u], max(left + sc[u], up + gap));
        left = up;
      }
    }
    // horizontal chain: H[j] = j*g + max_{k<=j} (t_k - k*g)
    int32_t pm = kNegInf;
    int32_t v[kStrip];
#pragma unroll
    for (int u = 0; u < kStrip; ++u) {
      const int j = j0 + u;
      v[u] = j <= L ? t[u] - j * gap : kNegInf;
      pm = max(pm, v[u]);
      v[u] = pm;
    }
    int32_t incl = wave_prefix_max(pm);
    int32_t excl = __shfl_up(incl, 1, kWave);
    if (lane == 0) {
      excl = kNegInf;
    }
    const int32_t in_carry = max(excl, carry - (t0 - 1) * gap);
    int32_t h[kStrip];
#pragma unroll
    for (int u = 0; u < kStrip; ++u) {
      const int j = j0 + u;
      h[u] = max(v[u], in_carry) + j * gap;
    }
    // store H
#pragma unroll
    for (int u = 0; u < kStrip; ++u) {
      if (j0 + u <= L) {
        Hr[j0 + u] = h[u];
      }
    }
    // left neighbour of the strip's first column
    int32_t hleft = __shfl_up(h[kStrip - 1], 1, kWave);
    if (lane == 0) {
      hleft = carry;
    }
    // move codes: first diagonal predecessor, else first vertical, else left
    uint16_t cd[kStrip];
#pragma unroll
    for (int u = 0; u < kStrip; ++u) {
      cd[u] = 0xffff;
    }
    for (int k = 0; k < np; ++k) {
      const int32_t* Hp = s.H + static_cast<int64_t>(preds[k]) * pitch;
      int32_t left = j0 - 1 <= L ? Hp[j0 - 1] : kNegInf;
#pragma unroll
      for (int u = 0; u < kStrip; ++u) {
        const int j = j0 + u;
        const int32_t up = j <= L ? Hp[j] : kNegInf;
        if (cd[u] == 0xffff && h[u] == left + sc[u]) {
          cd[u] = kCodeDiag | static_cast<uint16_t>(k);
        }
        left = up;
      }
    }
    for (int k = 0; k < np; ++k) {
      const int32_t* Hp = s.H + static_cast<int64_t>(preds[k]) * pitch;
#pragma unroll
      for (int u = 0; u < kStrip; ++u) {
        const int j = j0 + u;
        if (j <= L && cd[u] == 0xffff && h[u] == Hp[j] + gap) {
          cd[u] = kCodeVert | static_cast<uint16_t>(k);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kStrip; ++u) {
      const int j = j0 + u;
      if (j <= L) {
        const int32_t hl = u == 0 ? hleft : h[u - 1];
        if (cd[u] == 0xffff && h[u] == hl + gap) {
          cd[u] = kCodeHorz;
        }
        Cr[j] = cd[u];
      }
    }
    // carry into the next tile: the value of column t0 + kTile - 1
    const int last_lane_col = min(L, t0 + kTile - 1);
    const int owner = (last_lane_col - t0) / kStrip;
    const int last_u = (last_lane_col - t0) % kStrip;
    int32_t last_val = h[0];
#pragma unroll
    for (int u = 1; u < kStrip; ++u) {
      if (u == last_u) {
        last_val = h[u];
      }
    }
    carry = __shfl(last_val, owner, kWave);
  }
}

// ---- the window ----

__device__ void process_window(const PoaParams& P, Slot& s, const PoaWindow& W,
                               PoaResult& R) {
  const int lane = lane_id();
  const PoaSlotCaps& caps = P.caps;
  const int pitch = static_cast<int>(caps.cols);
  const uint8_t* bb = P.arena + W.backbone_off;
  const uint8_t* bq = P.arena + W.backbone_qual;
  const int n = static_cast<int>(W.backbone_len);
  uint64_t cells = 0;
  int max_rows = 0;

  // backbone chain: nodes 0..n-1, edges i-1 -> i (weights q-33), label 0
  if (n > static_cast<int>(caps.nodes) || n - 1 > static_cast<int>(caps.edges)) {
    R.status = kPoaNodeOverflow;
    return;
  }
  for (int i = lane; i < n; i += kWave) {
    s.base[i] = bb[i];
    s.in_head[i] = i > 0 ? i - 1 : -1;
    s.in_tail[i] = i > 0 ? i - 1 : -1;
    s.out_head[i] = i + 1 < n ? i : -1;
    s.out_tail[i] = i + 1 < n ? i : -1;
    s.n_aligned[i] = 0;
    s.cov[i] = n >= 2 ? 1 : 0;
    s.last_label[i] = n >= 2 ? 0 : -1;
    s.rank_to_node[i] = i;
    if (i + 1 < n) {
      s.e_tail[i] = i;
      s.e_head[i] = i + 1;
      s.e_weight[i] = (static_cast<int>(bq[i]) - 33) + (static_cast<int>(bq[i + 1]) - 33);
      s.e_next_in[i] = -1;
      s.e_next_out[i] = -1;
    }
  }
  s.n_nodes = n;
  s.n_edges = n > 0 ? n - 1 : 0;
  s.n_seqs = 1;
  wave_sync();

  for (uint32_t li = 0; li < W.n_layers; ++li) {
    const PoaLayer& Lr = P.layers[W.layer_first + li];
    const int L = static_cast<int>(Lr.len);
    if (L + 1 > pitch) {
      R.status = kPoaRowOverflow;
      return;
    }
    // layer bytes and weights into the slot
    for (int j = lane; j < L; j += kWave) {
      s.seq[j] = P.arena[Lr.seq_off + j];
      s.qual[j] = Lr.qual_off == kNoQuality
                      ? 1
                      : static_cast<int>(P.arena[Lr.qual_off + j]) - 33;
    }
    const bool full = Lr.begin < W.offset && Lr.end > W.backbone_len - W.offset;
    if (static_cast<int>(Lr.end) >= s.n_nodes || Lr.begin > Lr.end) {
      R.status = kPoaBadInput;
      return;
    }

    // 1. included node set
    const int nn = s.n_nodes;
    for (int v = lane; v < nn; v += kWave) {
      s.in_sub[v] = full ? 1 : 0;
    }
    wave_sync();
    if (!full) {
      int top = 0;
      s.stack[top++] = static_cast<int>(Lr.end);
      while (top > 0) {
        const int v = uniform(s.stack[--top]);
        if (s.in_sub[v] == 0 && v >= static_cast<int>(Lr.begin)) {
          const int na = s.n_aligned[v];
          int deg = na;
          for (int e = uniform(s.in_head[v]); e != -1; e = uniform(s.e_next_in[e])) {
            ++deg;
          }
          if (top + deg > static_cast<int>(caps.stack)) {
            R.status = kPoaNodeOverflow;
            return;
          }
          for (int e = uniform(s.in_head[v]); e != -1; e = uniform(s.e_next_in[e])) {
            s.stack[top++] = s.e_tail[e];
          }
          for (int k = 0; k < na; ++k) {
            s.stack[top++] = s.aligned[v * kMaxAligned + k];
          }
          s.in_sub[v] = 1;
        }
      }
    }

    // 2. topological order of the included nodes
    int V;
    if (full) {
      V = nn;
      for (int r = lane; r < V; r += kWave) {
        s.sub_order[r] = s.rank_to_node[r];
      }
    } else {
      V = topo_sort(s, s.in_sub, s.sub_order, static_cast<int>(caps.stack));
      if (s.overflow) {
        R.status = kPoaNodeOverflow;
        return;
      }
    }
    if (V + 1 > static_cast<int>(caps.rows)) {
      R.status = kPoaRowOverflow;
      return;
    }
    max_rows = max(max_rows, V + 1);
    wave_sync();

    // DP row of every included node; predecessor rows and sink flags
    for (int r = lane; r < V; r += kWave) {
      s.row_of[s.sub_order[r]] = r + 1;
    }
    wave_sync();
    // predecessor counts per row (lane-parallel), then offsets
    {
      int run = 0;
      for (int r0 = 0; r0 < V; r0 += kWave) {
        const int r = r0 + lane;
        int cnt = 0;
        if (r < V) {
          const int v = s.sub_order[r];
          for (int e = s.in_head[v]; e != -1; e = s.e_next_in[e]) {
            if (s.in_sub[s.e_tail[e]] != 0) {
              ++cnt;
            }
          }
          if (cnt == 0) {
            cnt = 1;  // the virtual start row
          }
          int out_cnt = 0;
          for (int e = s.out_head[v]; e != -1; e = s.e_next_out[e]) {
            if (s.in_sub[s.e_head[e]] != 0) {
              ++out_cnt;
            }
          }
          s.sink[r] = out_cnt == 0 ? 1 : 0;
        }
        // inclusive scan of cnt
        int x = cnt;
        for (int d = 1; d < kWave; d <<= 1) {
          int y = __shfl_up(x, d, kWave);
          if (lane >= d) x += y;
        }
        if (r < V) {
          s.pred_off[r] = run + x - cnt;
        }
        run += __shfl(x, kWave - 1, kWave);
      }
      if (lane == 0) {
        s.pred_off[V] = run;
      }
      wave_sync();
      for (int r = lane; r < V; r += kWave) {
        const int v = s.sub_order[r];
        int o = s.pred_off[r];
        bool any = false;
        for (int e = s.in_head[v]; e != -1; e = s.e_next_in[e]) {
          const int u = s.e_tail[e];
          if (s.in_sub[u] != 0) {
            s.pred_rows[o++] = s.row_of[u];
            any = true;
          }
        }
        if (!any) {
          s.pred_rows[o] = 0;
        }
      }
    }
    wave_sync();

    // 3. DP. Row 0 and column 0 first.
    for (int j = lane; j <= L; j += kWave) {
      s.H[j] = j * P.gap;
    }
    wave_sync();
    // column 0: H[r][0] = max_p H[p][0] + g, code = first such p (vertical)
    for (int r0 = 0; r0 < V; r0 += 1) {
      const int r = r0 + 1;
      const int b = uniform(s.pred_off[r0]);
      const int e = uniform(s.pred_off[r0 + 1]);
      int32_t best = kNegInf;
      for (int k = b; k < e; ++k) {
        best = max(best, s.H[static_cast<int64_t>(s.pred_rows[k]) * pitch]);
      }
      best += P.gap;
      int first = 0;
      for (int k = b; k < e; ++k) {
        if (s.H[static_cast<int64_t>(s.pred_rows[k]) * pitch] + P.gap == best) {
          first = k - b;
          break;
        }
      }
      s.H[static_cast<int64_t>(r) * pitch] = best;
      s.code[static_cast<int64_t>(r) * pitch] = kCodeVert | static_cast<uint16_t>(first);
    }
    wave_sync();
    for (int r = 1; r <= V; ++r) {
      const int v = uniform(s.sub_order[r - 1]);
      const int b = uniform(s.pred_off[r - 1]);
      const int e = uniform(s.pred_off[r]);
      dp_row(s, r, s.base[v], s.pred_rows + b, e - b, L, pitch, P.match, P.mismatch, P.gap);
      wave_sync();
    }
    cells += static_cast<uint64_t>(V) * static_cast<uint64_t>(L);

    // end cell: first sink row with the highest last-column score
    int32_t best = kNegInf;
    int best_row = 0x7fffffff;
    for (int r = lane + 1; r <= V; r += kWave) {
      if (s.sink[r - 1] != 0) {
        const int32_t h = s.H[static_cast<int64_t>(r) * pitch + L];
        if (h > best) {
          best = h;
          best_row = r;
        }
      }
    }
    for (int d = 1; d < kWave; d <<= 1) {
      const int32_t ob = __shfl_xor(best, d, kWave);
      const int orow = __shfl_xor(best_row, d, kWave);
      if (ob > best || (ob == best && orow < best_row)) {
        best = ob;
        best_row = orow;
      }
    }
    best_row = uniform(best_row);

    // 4. traceback (lane-uniform), pairs collected backwards
    int na = 0;
    {
      int i = best_row, j = L;
      const int aln_cap = static_cast<int>(caps.nodes + caps.cols);
      while (!(i == 0 && j == 0)) {
        if (na >= aln_cap || j < 0 || i < 0) {
          R.status = kPoaBadInput;
          return;
        }
        if (i == 0) {
          s.aln_node[na] = -1;
          s.aln_pos[na] = j - 1;
          ++na;
          --j;
          continue;
        }
        const int c = uniform(s.code[static_cast<int64_t>(i) * pitch + j]);
        const int kind = c & 0xc000;
        const int k = c & 0x3fff;
        const int node = uniform(s.sub_order[i - 1]);
        if (kind == kCodeDiag) {
          s.aln_node[na] = node;
          s.aln_pos[na] = j - 1;
          ++na;
          i = uniform(s.pred_rows[s.pred_off[i - 1] + k]);
          --j;
        } else if (kind == kCodeVert) {
          s.aln_node[na] = node;
          s.aln_pos[na] = -1;
          ++na;
          i = uniform(s.pred_rows[s.pred_off[i - 1] + k]);
        } else {
          s.aln_node[na] = -1;
          s.aln_pos[na] = j - 1;
          ++na;
          --j;
        }
      }
    }

    // 5. insert the layer (alignment walked forwards = pairs backwards)
    {
      int first_pos = -1, last_pos = -1;
      for (int q = na - 1; q >= 0; --q) {
        if (s.aln_pos[q] != -1) {
          first_pos = s.aln_pos[q];
          break;
        }
      }
      for (int q = 0; q < na; ++q) {
        if (s.aln_pos[q] != -1) {
          last_pos = s.aln_pos[q];
          break;
        }
      }
      const int before = s.n_nodes;
      add_chain(s, caps, 0, first_pos);
      int head = before == s.n_nodes ? -1 : s.n_nodes - 1;
      const int tail = add_chain(s, caps, last_pos + 1, L);
      int prev_w = head == -1 ? 0 : s.qual[first_pos - 1];
      for (int q = na - 1; q >= 0 && !s.overflow; --q) {
        const int pos = uniform(s.aln_pos[q]);
        if (pos == -1) {
          continue;
        }
        const int node = uniform(s.aln_node[q]);
        const uint8_t letter = s.seq[pos];
        int nv;
        if (node == -1) {
          nv = add_node(s, caps, letter);
        } else if (s.base[node] == letter) {
          nv = node;
        } else {
          int same = -1;
          const int cnt = s.n_aligned[node];
          for (int k = 0; k < cnt; ++k) {
            const int a = uniform(s.aligned[node * kMaxAligned + k]);
            if (s.base[a] == letter) {
              same = a;
              break;
            }
          }
          if (same == -1) {
            nv = add_node(s, caps, letter);
            if (s.overflow || cnt + 1 > static_cast<int>(kMaxAligned)) {
              s.overflow = true;
              break;
            }
            for (int k = 0; k < cnt; ++k) {
              const int a = uniform(s.aligned[node * kMaxAligned + k]);
              const int ca = s.n_aligned[a];
              s.aligned[nv * kMaxAligned + k] = a;
              s.aligned[a * kMaxAligned + ca] = nv;
              s.n_aligned[a] = static_cast<uint8_t>(ca + 1);
            }
            s.aligned[nv * kMaxAligned + cnt] = node;
            s.n_aligned[nv] = static_cast<uint8_t>(cnt + 1);
            s.aligned[node * kMaxAligned + cnt] = nv;
            s.n_aligned[node] = static_cast<uint8_t>(cnt + 1);
          } else {
            nv = same;
          }
        }
        if (head != -1) {
          add_edge(s, caps, head, nv, prev_w + s.qual[pos]);
        }
        head = nv;
        prev_w = s.qual[pos];
      }
      if (tail != -1 && !s.overflow) {
        add_edge(s, caps, head, tail, prev_w + s.qual[last_pos + 1]);
      }
      s.n_seqs = s.n_seqs + 1;
    }
    if (s.overflow) {
      R.status = kPoaNodeOverflow;
      return;
    }
    wave_sync();
    // new topological order of the whole graph
    topo_sort(s, nullptr, s.rank_to_node, static_cast<int>(caps.stack));
    if (s.overflow) {
      R.status = kPoaNodeOverflow;
      return;
    }
    wave_sync();
  }

  // heaviest bundle (lane-uniform)
  const int N = s.n_nodes;
  for (int v = lane; v < N; v += kWave) {
    s.scores[v] = -1;
    s.preds[v] = -1;
  }
  wave_sync();
  int max_node = 0;
  for (int r = 0; r < N; ++r) {
    const int v = uniform(s.rank_to_node[r]);
    for (int e = uniform(s.in_head[v]); e != -1; e = uniform(s.e_next_in[e])) {
      const int t = uniform(s.e_tail[e]);
      const int64_t w = s.e_weight[e];
      const int64_t sv = s.scores[v];
      const int pv = s.preds[v];
      if (sv < w || (sv == w && pv != -1 && s.scores[pv] <= s.scores[t])) {
        s.scores[v] = w;
        s.preds[v] = t;
      }
    }
    const int pv = s.preds[v];
    if (pv != -1) {
      s.scores[v] = s.scores[v] + s.scores[pv];
    }
    if (s.scores[max_node] < s.scores[v]) {
      max_node = v;
    }
  }
  if (s.out_head[max_node] != -1) {
    for (int r = lane; r < N; r += kWave) {
      s.node_rank[s.rank_to_node[r]] = r;
    }
    wave_sync();
    int guard = 0;
    while (s.out_head[max_node] != -1) {
      if (++guard > N + 1) {
        R.status = kPoaLoopLimit;
        return;
      }
      // branch completion from the rank of max_node
      const int rank = s.node_rank[max_node];
      const int v0 = max_node;
      for (int e = uniform(s.out_head[v0]); e != -1; e = uniform(s.e_next_out[e])) {
        const int h = uniform(s.e_head[e]);
        for (int f = uniform(s.in_head[h]); f != -1; f = uniform(s.e_next_in[f])) {
          const int t = uniform(s.e_tail[f]);
          if (t != v0) {
            s.scores[t] = -1;
          }
        }
      }
      int64_t best_score = 0;
      int best_node = 0;
      for (int r = rank + 1; r < N; ++r) {
        const int u = uniform(s.rank_to_node[r]);
        s.scores[u] = -1;
        s.preds[u] = -1;
        for (int e = uniform(s.in_head[u]); e != -1; e = uniform(s.e_next_in[e])) {
          const int t = uniform(s.e_tail[e]);
          const int64_t st = s.scores[t];
          if (st == -1) {
            continue;
          }
          const int64_t w = s.e_weight[e];
          const int64_t su = s.scores[u];
          const int pu = s.preds[u];
          if (su < w || (su == w && pu != -1 && s.scores[pu] <= st)) {
            s.scores[u] = w;
            s.preds[u] = t;
          }
        }
        const int pu = s.preds[u];
        if (pu != -1) {
          s.scores[u] = s.scores[u] + s.scores[pu];
        }
        if (best_score < s.scores[u]) {
          best_score = s.scores[u];
          best_node = u;
        }
      }
      max_node = best_node;
    }
  }
  // path back from max_node
  int plen = 0;
  {
    int v = max_node;
    while (true) {
      s.path[plen++] = v;
      const int p = uniform(s.preds[v]);
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
      const int v = s.path[plen - 1 - idx];
      int c = s.cov[v];
      const int na2 = s.n_aligned[v];
      for (int k = 0; k < na2; ++k) {
        c += s.cov[s.aligned[v * kMaxAligned + k]];
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
    P.out[W.out_off + i] = s.base[s.path[plen - 1 - (begin + i)]];
  }
  R.length = static_cast<uint32_t>(out_len);
  R.status = status;
  R.cells = cells;
  R.nodes = static_cast<uint32_t>(N);
  R.max_rows = static_cast<uint32_t>(max_rows);
}

}  // namespace

// Persistent batch kernel: each wave takes windows from the shared queue.
__global__ __launch_bounds__(64) void poa_batch_kernel(PoaParams P) {
  uint8_t* slot_mem = P.scratch + static_cast<uint64_t>(blockIdx.x) * P.slot_bytes;
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
    const uint32_t w = P.order[q];
    Slot s = carve(slot_mem, P.layout);
    PoaResult res{0, kPoaNotRun, 0, 0, 0};
    process_window(P, s, P.windows[w], res);
    __syncthreads();
    if (threadIdx.x == 0) {
      P.results[w] = res;
    }
    __syncthreads();
  }
}

// Plain kernel: one launch per window (single wave, window index given).
__global__ __launch_bounds__(64) void poa_single_kernel(PoaParams P, uint32_t q) {
  const uint32_t w = P.order[q];
  Slot s = carve(P.scratch, P.layout);
  PoaResult res{0, kPoaNotRun, 0, 0, 0};
  process_window(P, s, P.windows[w], res);
  __syncthreads();
  if (threadIdx.x == 0) {
    P.results[w] = res;
  }
}

}  // namespace rh

namespace rh {

// Host launchers (kept in this translation unit with the kernels).
hipError_t launch_poa_batch(const PoaParams& params, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL(poa_batch_kernel, dim3(grid), dim3(64), 0, stream, params);
  return hipGetLastError();
}

hipError_t launch_poa_single(const PoaParams& params, uint32_t q, hipStream_t stream) {
  hipLaunchKernelGGL(poa_single_kernel, dim3(1), dim3(64), 0, stream, params, q);
  return hipGetLastError();
}

hipError_t poa_occupancy(int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, poa_batch_kernel, 64, 0);
}

}  // namespace rh
