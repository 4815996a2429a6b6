// Overlap record (MHAP/PAF/SAM) and its breaking points.
//
// Fields, constructors and transmute follow racon src/overlap.hpp:32-124 and
// src/overlap.cpp:15-177. The alignment of an overlap (racon's
// Overlap::align_overlaps, overlap.cpp:205-224, an edlib NW path) is not done
// here: the polishing engine aligns batches of overlaps and hands back the
// alignment as edit operations, which find_breaking_points_from_ops walks the
// same way racon walks the CIGAR (overlap.cpp:226-292).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace racon {

class Sequence;

// One run of a CIGAR: an operation character and its length.
struct CigarRun {
  char op;
  uint32_t length;
};

// Edit operation codes of an alignment path (edlib's encoding): 0 match,
// 1 insertion (query base, no target base), 2 deletion (target base, no
// query base), 3 mismatch.
enum : uint8_t { kOpMatch = 0, kOpInsert = 1, kOpDelete = 2, kOpMismatch = 3 };

class Overlap {
 public:
  // MHAP (overlap.cpp:15-27)
  Overlap(uint64_t a_id, uint64_t b_id, double accuracy, uint32_t minmers,
          uint32_t a_rc, uint32_t a_begin, uint32_t a_end, uint32_t a_length,
          uint32_t b_rc, uint32_t b_begin, uint32_t b_end, uint32_t b_length);
  // PAF (overlap.cpp:29-42)
  Overlap(const char* q_name, uint32_t q_name_length, uint32_t q_length,
          uint32_t q_begin, uint32_t q_end, char orientation,
          const char* t_name, uint32_t t_name_length, uint32_t t_length,
          uint32_t t_begin, uint32_t t_end, uint32_t matching_bases,
          uint32_t overlap_length, uint32_t mapping_quality);
  // SAM (overlap.cpp:44-108)
  Overlap(const char* q_name, uint32_t q_name_length, uint32_t flag,
          const char* t_name, uint32_t t_name_length, uint32_t t_begin,
          uint32_t mapping_quality, const char* cigar, uint32_t cigar_length);

  Overlap(const Overlap&) = delete;
  Overlap& operator=(const Overlap&) = delete;

  uint64_t q_id() const { return q_id_; }
  uint64_t t_id() const { return t_id_; }
  uint32_t strand() const { return strand_; }
  bool is_valid() const { return is_valid_; }
  uint32_t length() const { return length_; }
  double error() const { return error_; }
  const std::string& cigar() const { return cigar_; }
  const std::vector<std::pair<uint32_t, uint32_t>>& breaking_points() const {
    return breaking_points_;
  }

  void transmute(const std::vector<std::unique_ptr<Sequence>>& sequences,
                 const std::unordered_map<std::string, uint64_t>& name_to_id,
                 const std::unordered_map<uint64_t, uint64_t>& id_to_id);

  // True when the overlap still needs a pairwise alignment (no CIGAR was
  // given by the overlap file and no breaking points exist yet):
  // overlap.cpp:188-198.
  bool needs_alignment() const;

  // Query and target spans that racon hands to edlib (overlap.cpp:193-197):
  // the query is the read span, reverse-complemented for a '-' overlap; the
  // target is the contig span.
  void alignment_input(const std::vector<std::unique_ptr<Sequence>>& sequences,
                       const char** q, uint32_t* q_length, const char** t,
                       uint32_t* t_length) const;

  // Breaking points from an alignment path given as edit operations
  // (kOpMatch/kOpInsert/kOpDelete/kOpMismatch), equivalent to racon walking
  // the edlib CIGAR (overlap.cpp:205-216 + 226-292).
  void find_breaking_points_from_ops(const uint8_t* ops, uint64_t n_ops,
                                     uint32_t window_length);

  // Target span and query start of the breaking-point walk
  // (overlap.cpp:228-240), for an engine that computes the points itself.
  void breaking_point_span(uint32_t* t_begin, uint32_t* t_end, uint32_t* q_first) const;

  // Breaking points computed elsewhere: n_points (target, query) pairs.
  void set_breaking_points(const uint32_t* points, uint64_t n_points);

  // Breaking points from the SAM CIGAR held by the overlap (overlap.cpp:200).
  void find_breaking_points_from_cigar(uint32_t window_length);

  void set_cigar(std::string cigar) { cigar_ = std::move(cigar); }

  // Breaking points from an explicit run list (the walk shared by both).
  void find_breaking_points_from_runs(const std::vector<CigarRun>& runs,
                                      uint32_t window_length);

  void check_transmuted() const;

 private:
  std::string q_name_;
  uint64_t q_id_ = 0;
  uint32_t q_begin_ = 0;
  uint32_t q_end_ = 0;
  uint32_t q_length_ = 0;

  std::string t_name_;
  uint64_t t_id_ = 0;
  uint32_t t_begin_ = 0;
  uint32_t t_end_ = 0;
  uint32_t t_length_ = 0;

  uint32_t strand_ = 0;
  uint32_t length_ = 0;
  double error_ = 0;
  std::string cigar_;

  bool is_valid_ = true;
  bool is_transmuted_ = false;
  std::vector<std::pair<uint32_t, uint32_t>> breaking_points_;
};

// Parses the run list of a CIGAR the way racon's loops do (atoi at the start
// of each run): overlap.cpp:72-92 and 244-291.
std::vector<CigarRun> parse_cigar_runs(const std::string& cigar);

// Run-length encodes edit operations into a standard CIGAR (M for match and
// mismatch), as edlibAlignmentToCigar(EDLIB_CIGAR_STANDARD) does.
std::string ops_to_cigar(const uint8_t* ops, uint64_t n_ops);

}  // namespace racon
