// ORACLE (test infrastructure only). See edlib_cpu.hpp for scope and sources.
#include "edlib_cpu.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>

namespace oracle {

namespace {

typedef uint64_t Word;
constexpr int kWordSize = 64;
constexpr Word kWord1 = 1;
constexpr Word kHighBit = kWord1 << (kWordSize - 1);

int ceil_div(int x, int y) { return x % y ? x / y + 1 : x / y; }

struct Block {
  Word P = 0;
  Word M = 0;
  int score = 0;
};

// Column data kept for the path: per (column, block) P, M and the score of
// the block's last cell, plus the band of blocks computed in each column.
struct ColumnStore {
  std::vector<Word> Ps, Ms;
  std::vector<int> scores, first_blocks, last_blocks;
  ColumnStore(int max_blocks, int columns)
      : Ps(static_cast<size_t>(max_blocks) * columns),
        Ms(static_cast<size_t>(max_blocks) * columns),
        scores(static_cast<size_t>(max_blocks) * columns),
        first_blocks(columns), last_blocks(columns) {}
};

// Match bit-vectors: bit r of word (symbol, block) is set when query row
// block*64 + r equals symbol; rows past the query end match everything.
std::vector<Word> build_peq(int alphabet_size, const unsigned char* query,
                            int query_length) {
  const int blocks = ceil_div(query_length, kWordSize);
  std::vector<Word> peq(static_cast<size_t>(alphabet_size + 1) * blocks, 0);
  for (int s = 0; s <= alphabet_size; ++s) {
    for (int b = 0; b < blocks; ++b) {
      Word w = 0;
      if (s == alphabet_size) {
        w = ~Word(0);
      } else {
        for (int r = (b + 1) * kWordSize - 1; r >= b * kWordSize; --r) {
          w <<= 1;
          if (r >= query_length || query[r] == s) {
            w |= 1;
          }
        }
      }
      peq[static_cast<size_t>(s) * blocks + b] = w;
    }
  }
  return peq;
}

// One 64-row block advanced by one column (Myers 1999 / Hyyro 2003 step with
// horizontal input hin in {-1, 0, +1}); returns hout.
inline int advance_block(Word Pv, Word Mv, Word Eq, int hin, Word& Pout, Word& Mout) {
  const Word hin_neg = static_cast<Word>(hin < 0);
  const Word Xv = Eq | Mv;
  Eq |= hin_neg;
  const Word Xh = (((Eq & Pv) + Pv) ^ Pv) | Eq;
  Word Ph = Mv | ~(Xh | Pv);
  Word Mh = Pv & Xh;
  int hout = 0;
  if (Ph & kHighBit) hout = 1;
  if (Mh & kHighBit) hout = -1;
  Ph <<= 1;
  Mh <<= 1;
  Mh |= hin_neg;
  Ph |= static_cast<Word>(hin > 0);
  Pout = Mh | ~(Xv | Ph);
  Mout = Ph & Xv;
  return hout;
}

// Cell values of a block, index 0 = last row, 63 = first row.
void block_values(const Block& b, int* out) {
  int score = b.score;
  Word mask = kHighBit;
  for (int i = 0; i < kWordSize - 1; ++i) {
    out[i] = score;
    if (b.P & mask) score--;
    if (b.M & mask) score++;
    mask >>= 1;
  }
  out[kWordSize - 1] = score;
}

// Banded NW columns (edlib's myersCalcEditDistanceNW). With store != null,
// every column is kept; with stop_column >= 0 only that column is kept (in
// the store's column 0) and the scan ends there.
void nw_columns(const Word* peq, int W, int max_blocks, int query_length,
                const unsigned char* target, int target_length, int k,
                int* best_score, int* position, ColumnStore* store,
                bool keep_all, int stop_column) {
  const int kStrongReduce = 2048;
  *best_score = *position = -1;
  if (k < std::abs(target_length - query_length)) {
    return;
  }
  k = std::min(k, std::max(query_length, target_length));

  int first_block = 0;
  int last_block = std::min(max_blocks,
                            ceil_div(std::min(k, (k + query_length - target_length) / 2) + 1,
                                     kWordSize)) - 1;
  std::vector<Block> blocks(max_blocks);
  for (int b = 0; b <= last_block; ++b) {
    blocks[b].score = (b + 1) * kWordSize;
    blocks[b].P = ~Word(0);
    blocks[b].M = 0;
  }

  int values[kWordSize];
  for (int c = 0; c < target_length; ++c) {
    const Word* peq_c = peq + static_cast<size_t>(target[c]) * max_blocks;

    int hout = 1;
    for (int b = first_block; b <= last_block; ++b) {
      hout = advance_block(blocks[b].P, blocks[b].M, peq_c[b], hout, blocks[b].P,
                           blocks[b].M);
      blocks[b].score += hout;
    }
    int bl = last_block;

    k = std::min(k, blocks[bl].score +
                        std::max(target_length - c - 1,
                                 query_length - ((1 + last_block) * kWordSize - 1) - 1) +
                        (last_block == max_blocks - 1 ? W : 0));

    // one more block below the band, when it may hold cells within k
    if (last_block + 1 < max_blocks &&
        !((last_block + 1) * kWordSize - 1 >
          k - blocks[bl].score + 2 * kWordSize - 2 - target_length + c + query_length)) {
      last_block++;
      bl++;
      blocks[bl].P = ~Word(0);
      blocks[bl].M = 0;
      int new_hout = advance_block(blocks[bl].P, blocks[bl].M, peq_c[last_block], hout,
                                   blocks[bl].P, blocks[bl].M);
      blocks[bl].score = blocks[bl - 1].score - hout + kWordSize + new_hout;
      hout = new_hout;
    }

    // drop last blocks that left the band
    while (last_block >= first_block &&
           (blocks[bl].score >= k + kWordSize ||
            ((last_block + 1) * kWordSize - 1 >
             k - blocks[bl].score + 2 * kWordSize - 2 - target_length + c + query_length + 1))) {
      last_block--;
      bl--;
    }
    // drop first blocks that left the band
    while (first_block <= last_block &&
           (blocks[first_block].score >= k + kWordSize ||
            ((first_block + 1) * kWordSize - 1 <
             blocks[first_block].score - k - target_length + query_length + c))) {
      first_block++;
    }

    // every kStrongReduce columns, a cell-exact band check
    if (c % kStrongReduce == 0) {
      while (last_block >= first_block) {
        block_values(blocks[last_block], values);
        int num_cells = last_block == max_blocks - 1 ? kWordSize - W : kWordSize;
        int r = last_block * kWordSize + num_cells - 1;
        bool reduce = true;
        for (int i = kWordSize - num_cells; i < kWordSize; ++i) {
          if (values[i] <= k && r <= k - values[i] - target_length + c + query_length + 1) {
            reduce = false;
            break;
          }
          r--;
        }
        if (!reduce) break;
        last_block--;
      }
      while (first_block <= last_block) {
        block_values(blocks[first_block], values);
        int num_cells = first_block == max_blocks - 1 ? kWordSize - W : kWordSize;
        int r = first_block * kWordSize + num_cells - 1;
        bool reduce = true;
        for (int i = kWordSize - num_cells; i < kWordSize; ++i) {
          if (values[i] <= k && r >= values[i] - k - target_length + c + query_length) {
            reduce = false;
            break;
          }
          r--;
        }
        if (!reduce) break;
        first_block++;
      }
    }

    if (last_block < first_block) {
      *best_score = *position = -1;
      return;
    }

    if (keep_all) {
      for (int b = first_block; b <= last_block; ++b) {
        const size_t idx = static_cast<size_t>(max_blocks) * c + b;
        store->Ps[idx] = blocks[b].P;
        store->Ms[idx] = blocks[b].M;
        store->scores[idx] = blocks[b].score;
      }
      store->first_blocks[c] = first_block;
      store->last_blocks[c] = last_block;
    }
    if (c == stop_column) {
      for (int b = first_block; b <= last_block; ++b) {
        store->Ps[b] = blocks[b].P;
        store->Ms[b] = blocks[b].M;
        store->scores[b] = blocks[b].score;
      }
      store->first_blocks[0] = first_block;
      store->last_blocks[0] = last_block;
      *best_score = -1;
      *position = stop_column;
      return;
    }
  }

  if (last_block == max_blocks - 1) {
    block_values(blocks[last_block], values);
    int best = values[W];
    if (best <= k) {
      *best_score = best;
      *position = target_length - 1;
    }
  }
}

// edlib's obtainAlignmentTraceback over the stored band.
int traceback(int query_length, int target_length, int best_score,
              const ColumnStore& d, std::vector<uint8_t>& ops) {
  const int max_blocks = ceil_div(query_length, kWordSize);
  const int W = max_blocks * kWordSize - query_length;
  auto P = [&](int c, int b) { return d.Ps[static_cast<size_t>(c) * max_blocks + b]; };
  auto M = [&](int c, int b) { return d.Ms[static_cast<size_t>(c) * max_blocks + b]; };
  auto S = [&](int c, int b) { return d.scores[static_cast<size_t>(c) * max_blocks + b]; };
  auto in_band = [&](int c, int b) {
    return b >= d.first_blocks[c] && b <= d.last_blocks[c];
  };

  ops.clear();
  int c = target_length - 1;
  int b = max_blocks - 1;
  int curr_score = best_score;
  int l_score = -1, u_score = -1, ul_score = -1;
  Word curr_p = P(c, b), curr_m = M(c, b);
  bool left_block = c > 0 && in_band(c - 1, b);
  Word l_p = 0, l_m = 0;
  if (left_block) {
    l_p = P(c - 1, b);
    l_m = M(c - 1, b);
  }
  curr_p <<= W;
  curr_m <<= W;
  int block_pos = kWordSize - W - 1;

  while (true) {
    if (c == 0) {
      left_block = true;
      l_score = b * kWordSize + block_pos + 1;
      ul_score = l_score - 1;
    }
    if (l_score == -1 && left_block) {
      l_score = S(c - 1, b);
      for (int i = 0; i < kWordSize - block_pos - 1; ++i) {
        if (l_p & kHighBit) l_score--;
        if (l_m & kHighBit) l_score++;
        l_p <<= 1;
        l_m <<= 1;
      }
    }
    if (ul_score == -1) {
      if (l_score != -1) {
        ul_score = l_score;
        if (l_p & kHighBit) ul_score--;
        if (l_m & kHighBit) ul_score++;
      } else if (c > 0 && in_band(c - 1, b - 1)) {
        ul_score = S(c - 1, b - 1);
      }
    }
    if (u_score == -1) {
      u_score = curr_score;
      if (curr_p & kHighBit) u_score--;
      if (curr_m & kHighBit) u_score++;
      curr_p <<= 1;
      curr_m <<= 1;
    }

    if (u_score != -1 && u_score + 1 == curr_score) {  // up: insertion
      curr_score = u_score;
      l_score = ul_score;
      u_score = ul_score = -1;
      if (block_pos == 0) {
        if (b == 0) {
          ops.push_back(kEdInsert);
          for (int i = 0; i < c + 1; ++i) ops.push_back(kEdDelete);
          break;
        }
        block_pos = kWordSize - 1;
        b--;
        curr_p = P(c, b);
        curr_m = M(c, b);
        if (c > 0 && in_band(c - 1, b)) {
          left_block = true;
          l_p = P(c - 1, b);
          l_m = M(c - 1, b);
        } else {
          left_block = false;
        }
      } else {
        block_pos--;
        l_p <<= 1;
        l_m <<= 1;
      }
      ops.push_back(kEdInsert);
    } else if (l_score != -1 && l_score + 1 == curr_score) {  // left: deletion
      curr_score = l_score;
      u_score = ul_score;
      l_score = ul_score = -1;
      c--;
      if (c == -1) {
        ops.push_back(kEdDelete);
        int num_up = b * kWordSize + block_pos + 1;
        for (int i = 0; i < num_up; ++i) ops.push_back(kEdInsert);
        break;
      }
      curr_p = l_p;
      curr_m = l_m;
      if (c > 0 && in_band(c - 1, b)) {
        left_block = true;
        l_p = P(c - 1, b);
        l_m = M(c - 1, b);
      } else if (c == 0) {
        left_block = true;
        l_score = b * kWordSize + block_pos + 1;
        ul_score = l_score - 1;
      } else {
        left_block = false;
      }
      ops.push_back(kEdDelete);
    } else if (ul_score != -1) {  // diagonal: (mis)match
      const uint8_t code = ul_score == curr_score ? kEdMatch : kEdMismatch;
      curr_score = ul_score;
      u_score = l_score = ul_score = -1;
      c--;
      if (c == -1) {
        ops.push_back(code);
        int num_up = b * kWordSize + block_pos;
        for (int i = 0; i < num_up; ++i) ops.push_back(kEdInsert);
        break;
      }
      if (block_pos == 0) {
        if (b == 0) {
          ops.push_back(code);
          for (int i = 0; i < c + 1; ++i) ops.push_back(kEdDelete);
          break;
        }
        block_pos = kWordSize - 1;
        b--;
        curr_p = P(c, b);
        curr_m = M(c, b);
      } else {
        block_pos--;
        curr_p = l_p;
        curr_m = l_m;
        curr_p <<= 1;
        curr_m <<= 1;
      }
      if (c > 0 && in_band(c - 1, b)) {
        left_block = true;
        l_p = P(c - 1, b);
        l_m = M(c - 1, b);
      } else if (c == 0) {
        left_block = true;
        l_score = b * kWordSize + block_pos + 1;
        ul_score = l_score - 1;
      } else {
        left_block = false;
      }
      ops.push_back(code);
    } else {
      break;
    }
  }
  std::reverse(ops.begin(), ops.end());
  return 0;
}

int obtain_alignment(const unsigned char* query, const unsigned char* r_query,
                     int query_length, const unsigned char* target,
                     const unsigned char* r_target, int target_length,
                     int alphabet_size, int best_score, std::vector<uint8_t>& ops);

// edlib's obtainAlignmentHirschberg.
int hirschberg(const unsigned char* query, const unsigned char* r_query,
               int query_length, const unsigned char* target,
               const unsigned char* r_target, int target_length, int alphabet_size,
               int best_score, std::vector<uint8_t>& ops) {
  const int max_blocks = ceil_div(query_length, kWordSize);
  const int W = max_blocks * kWordSize - query_length;
  std::vector<Word> peq = build_peq(alphabet_size, query, query_length);
  std::vector<Word> r_peq = build_peq(alphabet_size, r_query, query_length);

  const int left_width = target_length / 2;
  const int right_width = target_length - left_width;

  int score_, left_end, right_end;
  ColumnStore left(max_blocks, 1), right(max_blocks, 1);
  nw_columns(peq.data(), W, max_blocks, query_length, target, target_length,
             best_score, &score_, &left_end, &left, false, left_width - 1);
  nw_columns(r_peq.data(), W, max_blocks, query_length, r_target, target_length,
             best_score, &score_, &right_end, &right, false, right_width - 1);
  if (left_end != left_width - 1 || right_end != right_width - 1) {
    return 1;
  }

  // left column values, first row first
  const int lf = left.first_blocks[0], ll = left.last_blocks[0];
  int left_len = (ll - lf + 1) * kWordSize;
  std::vector<int> scores_left(left_len);
  for (int b = lf; b <= ll; ++b) {
    Block blk{left.Ps[b], left.Ms[b], left.scores[b]};
    int v[kWordSize];
    block_values(blk, v);
    for (int i = 0; i < kWordSize; ++i) {
      scores_left[(b - lf) * kWordSize + i] = v[kWordSize - 1 - i];
    }
  }
  const int left_start = lf * kWordSize;
  if (ll == max_blocks - 1) {
    left_len -= W;
  }

  // right column values, reversed back into query order
  const int rf = right.first_blocks[0], rl = right.last_blocks[0];
  int right_len = (rl - rf + 1) * kWordSize;
  std::vector<int> scores_right_all(right_len);
  for (int b = rf; b <= rl; ++b) {
    Block blk{right.Ps[b], right.Ms[b], right.scores[b]};
    int v[kWordSize];
    block_values(blk, v);
    for (int i = 0; i < kWordSize; ++i) {
      scores_right_all[(rl - b) * kWordSize + i] = v[i];
    }
  }
  int right_start = query_length - (rl + 1) * kWordSize;
  int right_offset = 0;
  if (right_start < 0) {
    right_offset = W;
    right_start += W;
    right_len -= W;
  }
  const int* scores_right = scores_right_all.data() + right_offset;

  const int q_from = std::max(left_start, right_start - 1);
  const int q_to = std::min(left_start + left_len - 1, right_start + right_len - 2);
  int left_score = -1, right_score = -1;
  int split = -1;
  bool found = false;
  for (int q = q_from; q <= q_to; ++q) {
    left_score = scores_left[q - left_start];
    right_score = scores_right[q + 1 - right_start];
    if (left_score + right_score == best_score) {
      split = q;
      found = true;
      break;
    }
  }
  if (!found && left_start == 0 && right_start == 0) {
    left_score = left_width;
    right_score = scores_right[0];
    if (left_score + right_score == best_score) {
      split = -1;
      found = true;
    }
  }
  if (!found && left_start + left_len == query_length &&
      right_start + right_len == query_length) {
    left_score = scores_left[left_len - 1];
    right_score = right_width;
    if (left_score + right_score == best_score) {
      split = query_length - 1;
      found = true;
    }
  }
  if (!found) {
    return 1;
  }

  const int ul_height = split + 1;
  const int lr_height = query_length - ul_height;
  std::vector<uint8_t> ul_ops, lr_ops;
  int s1 = obtain_alignment(query, r_query + lr_height, ul_height, target,
                            r_target + right_width, left_width, alphabet_size,
                            left_score, ul_ops);
  int s2 = obtain_alignment(query + ul_height, r_query, lr_height, target + left_width,
                            r_target, right_width, alphabet_size, right_score, lr_ops);
  if (s1 || s2) {
    return 1;
  }
  ops = std::move(ul_ops);
  ops.insert(ops.end(), lr_ops.begin(), lr_ops.end());
  return 0;
}

int obtain_alignment(const unsigned char* query, const unsigned char* r_query,
                     int query_length, const unsigned char* target,
                     const unsigned char* r_target, int target_length,
                     int alphabet_size, int best_score, std::vector<uint8_t>& ops) {
  if (query_length == 0 || target_length == 0) {
    ops.assign(query_length + target_length, query_length == 0 ? kEdDelete : kEdInsert);
    return 0;
  }
  const int max_blocks = ceil_div(query_length, kWordSize);
  const int W = max_blocks * kWordSize - query_length;
  const long long data_size = (2ll * sizeof(Word) + sizeof(int)) * max_blocks * target_length;
  if (data_size < 1024 * 1024) {
    std::vector<Word> peq = build_peq(alphabet_size, query, query_length);
    ColumnStore store(max_blocks, target_length);
    int score_, end_;
    nw_columns(peq.data(), W, max_blocks, query_length, target, target_length,
               best_score, &score_, &end_, &store, true, -1);
    return traceback(query_length, target_length, best_score, store, ops);
  }
  return hirschberg(query, r_query, query_length, target, r_target, target_length,
                    alphabet_size, best_score, ops);
}

// Maps bytes to alphabet indices in order of first appearance (query, then
// target).
int transform(const char* q, int ql, const char* t, int tl,
              std::vector<unsigned char>& qo, std::vector<unsigned char>& to) {
  int map[256];
  std::fill(map, map + 256, -1);
  int n = 0;
  qo.resize(ql);
  to.resize(tl);
  for (int i = 0; i < ql; ++i) {
    unsigned char c = static_cast<unsigned char>(q[i]);
    if (map[c] < 0) map[c] = n++;
    qo[i] = static_cast<unsigned char>(map[c]);
  }
  for (int i = 0; i < tl; ++i) {
    unsigned char c = static_cast<unsigned char>(t[i]);
    if (map[c] < 0) map[c] = n++;
    to[i] = static_cast<unsigned char>(map[c]);
  }
  return n;
}

// ---- exact (unbanded) version of the same recursion ----

// D(query[0..r], target[0..c]) for the last column, rows -1..ql-1.
void exact_last_column(const unsigned char* q, int ql, const unsigned char* t, int tl,
                       std::vector<int>& col) {
  col.resize(ql + 1);
  for (int r = 0; r <= ql; ++r) col[r] = r;  // column -1: D[r-1][-1] = r
  for (int c = 0; c < tl; ++c) {
    int diag = col[0];
    col[0] = c + 1;
    for (int r = 1; r <= ql; ++r) {
      int up = col[r - 1];
      int left = col[r];
      int v = std::min(std::min(up, left) + 1, diag + (q[r - 1] == t[c] ? 0 : 1));
      diag = left;
      col[r] = v;
    }
  }
}

int exact_obtain(const unsigned char* q, int ql, const unsigned char* t, int tl,
                 int best, std::vector<uint8_t>& ops);

int exact_traceback(const unsigned char* q, int ql, const unsigned char* t, int tl,
                    int best, std::vector<uint8_t>& ops) {
  // full matrix with boundary row/column at index 0
  const int W = tl + 1;
  std::vector<int> D(static_cast<size_t>(ql + 1) * W);
  for (int c = 0; c <= tl; ++c) D[c] = c;
  for (int r = 1; r <= ql; ++r) {
    D[static_cast<size_t>(r) * W] = r;
    for (int c = 1; c <= tl; ++c) {
      int v = std::min(D[static_cast<size_t>(r - 1) * W + c], D[static_cast<size_t>(r) * W + c - 1]) + 1;
      v = std::min(v, D[static_cast<size_t>(r - 1) * W + c - 1] + (q[r - 1] == t[c - 1] ? 0 : 1));
      D[static_cast<size_t>(r) * W + c] = v;
    }
  }
  if (D[static_cast<size_t>(ql) * W + tl] != best) {
    return 1;
  }
  ops.clear();
  int r = ql, c = tl;  // 1-based cell (row r-1, column c-1)
  while (r > 0 && c > 0) {
    const int cur = D[static_cast<size_t>(r) * W + c];
    if (D[static_cast<size_t>(r - 1) * W + c] + 1 == cur) {
      ops.push_back(kEdInsert);
      r--;
    } else if (D[static_cast<size_t>(r) * W + c - 1] + 1 == cur) {
      ops.push_back(kEdDelete);
      c--;
    } else {
      ops.push_back(D[static_cast<size_t>(r - 1) * W + c - 1] == cur ? kEdMatch : kEdMismatch);
      r--;
      c--;
    }
  }
  while (c > 0) {
    ops.push_back(kEdDelete);
    c--;
  }
  while (r > 0) {
    ops.push_back(kEdInsert);
    r--;
  }
  std::reverse(ops.begin(), ops.end());
  return 0;
}

int exact_hirschberg(const unsigned char* q, int ql, const unsigned char* t, int tl,
                     int best, std::vector<uint8_t>& ops) {
  const int lw = tl / 2, rw = tl - lw;
  std::vector<int> left, right_rev;
  exact_last_column(q, ql, t, lw, left);  // left[r+1] = D(q[0..r], t[0..lw-1])
  std::vector<unsigned char> rq(q, q + ql), rt(t + lw, t + tl);
  std::reverse(rq.begin(), rq.end());
  std::reverse(rt.begin(), rt.end());
  exact_last_column(rq.data(), ql, rt.data(), rw, right_rev);
  // right(x) = ED(q[x..], t[lw..]) = right_rev[ql - x]
  auto right = [&](int x) { return right_rev[ql - x]; };
  int split = -2, ls = -1, rs = -1;
  for (int x = 0; x <= ql - 2; ++x) {
    if (left[x + 1] + right(x + 1) == best) {
      split = x;
      ls = left[x + 1];
      rs = right(x + 1);
      break;
    }
  }
  if (split == -2 && lw + right(0) == best) {
    split = -1;
    ls = lw;
    rs = right(0);
  }
  if (split == -2 && left[ql] + rw == best) {
    split = ql - 1;
    ls = left[ql];
    rs = rw;
  }
  if (split == -2) {
    return 1;
  }
  std::vector<uint8_t> a, b;
  if (exact_obtain(q, split + 1, t, lw, ls, a) ||
      exact_obtain(q + split + 1, ql - split - 1, t + lw, rw, rs, b)) {
    return 1;
  }
  ops = std::move(a);
  ops.insert(ops.end(), b.begin(), b.end());
  return 0;
}

int exact_obtain(const unsigned char* q, int ql, const unsigned char* t, int tl,
                 int best, std::vector<uint8_t>& ops) {
  if (ql == 0 || tl == 0) {
    ops.assign(ql + tl, ql == 0 ? kEdDelete : kEdInsert);
    return 0;
  }
  const long long blocks = ceil_div(ql, kWordSize);
  if ((2ll * sizeof(Word) + sizeof(int)) * blocks * tl < 1024 * 1024) {
    return exact_traceback(q, ql, t, tl, best, ops);
  }
  return exact_hirschberg(q, ql, t, tl, best, ops);
}

}  // namespace

EdlibResult edlib_align_nw(const char* query_s, int32_t query_length,
                           const char* target_s, int32_t target_length, bool find_path) {
  EdlibResult result;
  std::vector<unsigned char> query, target;
  const int alphabet_size = transform(query_s, query_length, target_s, target_length,
                                      query, target);
  if (query_length == 0 || target_length == 0) {
    result.edit_distance = std::max(query_length, target_length);
    if (find_path) {
      result.ops.assign(query_length + target_length,
                        query_length == 0 ? kEdDelete : kEdInsert);
    }
    return result;
  }
  const int max_blocks = ceil_div(query_length, kWordSize);
  const int W = max_blocks * kWordSize - query_length;
  std::vector<Word> peq = build_peq(alphabet_size, query.data(), query_length);
  int k = kWordSize;
  int position = -1;
  do {
    nw_columns(peq.data(), W, max_blocks, query_length, target.data(), target_length,
               k, &result.edit_distance, &position, nullptr, false, -1);
    k *= 2;
  } while (result.edit_distance == -1);

  if (find_path) {
    std::vector<unsigned char> r_query(query.rbegin(), query.rend());
    std::vector<unsigned char> r_target(target.rbegin(), target.rend());
    result.status = obtain_alignment(query.data(), r_query.data(), query_length,
                                     target.data(), r_target.data(), target_length,
                                     alphabet_size, result.edit_distance, result.ops);
  }
  return result;
}

EdlibResult edlib_align_nw_exact(const char* query_s, int32_t query_length,
                                 const char* target_s, int32_t target_length) {
  EdlibResult result;
  std::vector<unsigned char> query(query_s, query_s + query_length);
  std::vector<unsigned char> target(target_s, target_s + target_length);
  if (query_length == 0 || target_length == 0) {
    result.edit_distance = std::max(query_length, target_length);
    result.ops.assign(query_length + target_length,
                      query_length == 0 ? kEdDelete : kEdInsert);
    return result;
  }
  std::vector<int> col;
  exact_last_column(query.data(), query_length, target.data(), target_length, col);
  result.edit_distance = col[query_length];
  result.status = exact_obtain(query.data(), query_length, target.data(), target_length,
                               result.edit_distance, result.ops);
  return result;
}

int32_t edit_distance(const char* query, int32_t query_length, const char* target,
                      int32_t target_length) {
  return edlib_align_nw(query, query_length, target, target_length, false).edit_distance;
}

}  // namespace oracle
