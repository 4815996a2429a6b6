#include "overlap.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "sequence.hpp"

namespace racon {

Overlap::Overlap(uint64_t a_id, uint64_t b_id, double, uint32_t,
                 uint32_t a_rc, uint32_t a_begin, uint32_t a_end,
                 uint32_t a_length, uint32_t b_rc, uint32_t b_begin,
                 uint32_t b_end, uint32_t b_length)
    : q_id_(a_id - 1), q_begin_(a_begin), q_end_(a_end), q_length_(a_length),
      t_id_(b_id - 1), t_begin_(b_begin), t_end_(b_end), t_length_(b_length),
      strand_(a_rc ^ b_rc) {
  length_ = std::max(q_end_ - q_begin_, t_end_ - t_begin_);
  error_ = 1 - std::min(q_end_ - q_begin_, t_end_ - t_begin_) /
                   static_cast<double>(length_);
}

Overlap::Overlap(const char* q_name, uint32_t q_name_length, uint32_t q_length,
                 uint32_t q_begin, uint32_t q_end, char orientation,
                 const char* t_name, uint32_t t_name_length,
                 uint32_t t_length, uint32_t t_begin, uint32_t t_end,
                 uint32_t, uint32_t, uint32_t)
    : q_name_(q_name, q_name_length), q_begin_(q_begin), q_end_(q_end),
      q_length_(q_length), t_name_(t_name, t_name_length), t_begin_(t_begin),
      t_end_(t_end), t_length_(t_length), strand_(orientation == '-') {
  length_ = std::max(q_end_ - q_begin_, t_end_ - t_begin_);
  error_ = 1 - std::min(q_end_ - q_begin_, t_end_ - t_begin_) /
                   static_cast<double>(length_);
}

Overlap::Overlap(const char* q_name, uint32_t q_name_length, uint32_t flag,
                 const char* t_name, uint32_t t_name_length, uint32_t t_begin,
                 uint32_t, const char* cigar, uint32_t cigar_length)
    : q_name_(q_name, q_name_length), t_name_(t_name, t_name_length),
      t_begin_(t_begin - 1), strand_(flag & 0x10), cigar_(cigar, cigar_length),
      is_valid_(!(flag & 0x4)) {
  if (cigar_.size() < 2 && is_valid_) {
    fprintf(stderr, "[Racon::Overlap::Overlap] error: "
                    "missing alignment from SAM object!\n");
    exit(1);
  }
  // a leading soft/hard clip gives the query begin (overlap.cpp:60-69)
  for (uint32_t i = 0; i < cigar_.size(); ++i) {
    const char c = cigar_[i];
    if (c == 'S' || c == 'H') {
      q_begin_ = atoi(&cigar_[0]);
      break;
    } else if (c == 'M' || c == '=' || c == 'I' || c == 'D' || c == 'N' ||
               c == 'P' || c == 'X') {
      break;
    }
  }
  uint32_t q_alignment_length = 0, q_clip_length = 0, t_alignment_length = 0;
  for (uint32_t i = 0, j = 0; i < cigar_.size(); ++i) {
    const char c = cigar_[i];
    if (c == 'M' || c == '=' || c == 'X') {
      uint32_t n = atoi(&cigar_[j]);
      j = i + 1;
      q_alignment_length += n;
      t_alignment_length += n;
    } else if (c == 'I') {
      q_alignment_length += atoi(&cigar_[j]);
      j = i + 1;
    } else if (c == 'D' || c == 'N') {
      t_alignment_length += atoi(&cigar_[j]);
      j = i + 1;
    } else if (c == 'S' || c == 'H') {
      q_clip_length += atoi(&cigar_[j]);
      j = i + 1;
    } else if (c == 'P') {
      j = i + 1;
    }
  }
  q_end_ = q_begin_ + q_alignment_length;
  q_length_ = q_clip_length + q_alignment_length;
  if (strand_) {
    uint32_t tmp = q_begin_;
    q_begin_ = q_length_ - q_end_;
    q_end_ = q_length_ - tmp;
  }
  t_end_ = t_begin_ + t_alignment_length;
  length_ = std::max(q_alignment_length, t_alignment_length);
  error_ = 1 - std::min(q_alignment_length, t_alignment_length) /
                   static_cast<double>(length_);
}

template <typename K>
static bool lookup_id(const std::unordered_map<K, uint64_t>& map, const K& key,
                      uint64_t& id) {
  auto it = map.find(key);
  if (it == map.end()) {
    return false;
  }
  id = it->second;
  return true;
}

void Overlap::transmute(const std::vector<std::unique_ptr<Sequence>>& sequences,
                        const std::unordered_map<std::string, uint64_t>& name_to_id,
                        const std::unordered_map<uint64_t, uint64_t>& id_to_id) {
  // overlap.cpp:129-177
  if (!is_valid_ || is_transmuted_) {
    return;
  }
  if (!q_name_.empty()) {
    if (!lookup_id(name_to_id, q_name_ + "q", q_id_)) {
      is_valid_ = false;
      return;
    }
    std::string().swap(q_name_);
  } else if (!lookup_id(id_to_id, q_id_ << 1 | 0, q_id_)) {
    is_valid_ = false;
    return;
  }
  if (q_length_ != sequences[q_id_]->data().size()) {
    fprintf(stderr, "[racon::Overlap::transmute] error: "
                    "unequal lengths in sequence and overlap file for sequence %s!\n",
            sequences[q_id_]->name().c_str());
    exit(1);
  }
  if (!t_name_.empty()) {
    if (!lookup_id(name_to_id, t_name_ + "t", t_id_)) {
      is_valid_ = false;
      return;
    }
    std::string().swap(t_name_);
  } else if (!lookup_id(id_to_id, t_id_ << 1 | 1, t_id_)) {
    is_valid_ = false;
    return;
  }
  if (t_length_ != 0 && t_length_ != sequences[t_id_]->data().size()) {
    fprintf(stderr, "[racon::Overlap::transmute] error: "
                    "unequal lengths in target and overlap file for target %s!\n",
            sequences[t_id_]->name().c_str());
    exit(1);
  }
  t_length_ = sequences[t_id_]->data().size();
  is_transmuted_ = true;
}

void Overlap::check_transmuted() const {
  if (!is_transmuted_) {
    fprintf(stderr, "[racon::Overlap::find_breaking_points] error: "
                    "overlap is not transmuted!\n");
    exit(1);
  }
}

bool Overlap::needs_alignment() const {
  return breaking_points_.empty() && cigar_.empty();
}

void Overlap::alignment_input(const std::vector<std::unique_ptr<Sequence>>& sequences,
                              const char** q, uint32_t* q_length,
                              const char** t, uint32_t* t_length) const {
  *q = !strand_ ? &(sequences[q_id_]->data()[q_begin_])
                : &(sequences[q_id_]->reverse_complement()[q_length_ - q_end_]);
  *q_length = q_end_ - q_begin_;
  *t = &(sequences[t_id_]->data()[t_begin_]);
  *t_length = t_end_ - t_begin_;
}

std::vector<CigarRun> parse_cigar_runs(const std::string& cigar) {
  std::vector<CigarRun> runs;
  for (uint32_t i = 0, j = 0; i < cigar.size(); ++i) {
    const char c = cigar[i];
    if (c >= '0' && c <= '9') {
      continue;
    }
    runs.push_back({c, static_cast<uint32_t>(atoi(&cigar[j]))});
    j = i + 1;
  }
  return runs;
}

std::string ops_to_cigar(const uint8_t* ops, uint64_t n_ops) {
  static const char kOpChar[4] = {'M', 'I', 'D', 'M'};
  std::string cigar;
  uint64_t i = 0;
  while (i < n_ops) {
    const char c = kOpChar[ops[i] & 3];
    uint64_t k = i + 1;
    while (k < n_ops && kOpChar[ops[k] & 3] == c) {
      ++k;
    }
    cigar += std::to_string(k - i);
    cigar += c;
    i = k;
  }
  return cigar;
}

void Overlap::find_breaking_points_from_runs(const std::vector<CigarRun>& runs,
                                             uint32_t window_length) {
  // overlap.cpp:226-292
  std::vector<int32_t> window_ends;
  for (uint32_t i = 0; i < t_end_; i += window_length) {
    if (i > t_begin_) {
      window_ends.emplace_back(i - 1);
    }
  }
  window_ends.emplace_back(t_end_ - 1);

  uint32_t w = 0;
  bool found_first_match = false;
  std::pair<uint32_t, uint32_t> first_match = {0, 0}, last_match = {0, 0};

  int32_t q_ptr = (strand_ ? (q_length_ - q_end_) : q_begin_) - 1;
  int32_t t_ptr = t_begin_ - 1;

  auto window_end_reached = [&]() {
    if (w < window_ends.size() && t_ptr == window_ends[w]) {
      if (found_first_match) {
        breaking_points_.emplace_back(first_match);
        breaking_points_.emplace_back(last_match);
      }
      found_first_match = false;
      ++w;
    }
  };

  for (const auto& run : runs) {
    if (run.op == 'M' || run.op == '=' || run.op == 'X') {
      for (uint32_t k = 0; k < run.length; ++k) {
        ++q_ptr;
        ++t_ptr;
        if (!found_first_match) {
          found_first_match = true;
          first_match.first = t_ptr;
          first_match.second = q_ptr;
        }
        last_match.first = t_ptr + 1;
        last_match.second = q_ptr + 1;
        window_end_reached();
      }
    } else if (run.op == 'I') {
      q_ptr += run.length;
    } else if (run.op == 'D' || run.op == 'N') {
      for (uint32_t k = 0; k < run.length; ++k) {
        ++t_ptr;
        window_end_reached();
      }
    }
  }
}

void Overlap::breaking_point_span(uint32_t* t_begin, uint32_t* t_end, uint32_t* q_first) const {
  // the origins of the walk in find_breaking_points_from_runs
  *t_begin = t_begin_;
  *t_end = t_end_;
  *q_first = strand_ ? (q_length_ - q_end_) : q_begin_;
}

void Overlap::set_breaking_points(const uint32_t* points, uint64_t n_points) {
  check_transmuted();
  breaking_points_.clear();
  breaking_points_.reserve(n_points);
  for (uint64_t k = 0; k < n_points; ++k) {
    breaking_points_.emplace_back(points[2 * k], points[2 * k + 1]);
  }
  std::string().swap(cigar_);
}

void Overlap::find_breaking_points_from_cigar(uint32_t window_length) {
  check_transmuted();
  if (!breaking_points_.empty()) {
    return;
  }
  find_breaking_points_from_runs(parse_cigar_runs(cigar_), window_length);
  std::string().swap(cigar_);
}

void Overlap::find_breaking_points_from_ops(const uint8_t* ops, uint64_t n_ops,
                                            uint32_t window_length) {
  check_transmuted();
  if (!breaking_points_.empty()) {
    return;
  }
  // run-length view of the op path; match and mismatch both walk as 'M'
  std::vector<CigarRun> runs;
  static const char kOpChar[4] = {'M', 'I', 'D', 'M'};
  for (uint64_t i = 0; i < n_ops;) {
    const char c = kOpChar[ops[i] & 3];
    uint64_t k = i + 1;
    while (k < n_ops && kOpChar[ops[k] & 3] == c) {
      ++k;
    }
    runs.push_back({c, static_cast<uint32_t>(k - i)});
    i = k;
  }
  find_breaking_points_from_runs(runs, window_length);
  std::string().swap(cigar_);
}

}  // namespace racon
