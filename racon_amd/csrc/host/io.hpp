// Readers for FASTA/FASTQ sequences and MHAP/PAF/SAM overlaps, plain or
// gzip-compressed (the role of rvaser/bioparser in racon,
// polisher.cpp:83-133). Record names are the header up to the first
// whitespace.
#pragma once

#include <memory>
#include <string>
#include <vector>

namespace racon {

class Sequence;
class Overlap;
class ThreadPool;

enum class SequenceFormat { kFasta, kFastq };
enum class OverlapFormat { kMhap, kPaf, kSam };

// Extension checks of createPolisher (polisher.cpp:76-133).
bool is_suffix(const std::string& src, const std::string& suffix);
bool sequence_format_of(const std::string& path, SequenceFormat* format);
bool overlap_format_of(const std::string& path, OverlapFormat* format);

// The bytes of a whole input file: a plain file is mapped read-only (no
// copy, no zero-fill of a buffer: 15 ms instead of 310 ms for a 487 MB
// FASTQ), a gzip file is inflated into memory. Exits with a racon-style
// message on failure.
class FileText {
 public:
  explicit FileText(const std::string& path);
  ~FileText();
  FileText(const FileText&) = delete;
  FileText& operator=(const FileText&) = delete;
  const char* data() const { return data_; }
  size_t size() const { return size_; }

 private:
  std::string inflated_;
  void* map_ = nullptr;
  size_t map_size_ = 0;
  const char* data_ = "";
  size_t size_ = 0;
};

// Appends the file's records to dst; with a pool, the Sequence objects are
// built on its threads (same objects, same order).
void parse_sequences(const std::string& path, SequenceFormat format,
                     std::vector<std::unique_ptr<Sequence>>& dst, ThreadPool* pool = nullptr);
// Appends the file's overlaps to dst in file order; with a pool, the lines
// are parsed on its threads in ranges.
void parse_overlaps(const std::string& path, OverlapFormat format,
                    std::vector<std::unique_ptr<Overlap>>& dst, ThreadPool* pool = nullptr);

}  // namespace racon
