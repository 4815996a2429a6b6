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

// Reads the whole file (gzip or plain); exits with racon-style message on
// failure.
std::string read_file(const std::string& path);

// Appends the file's records to dst; with a pool, the Sequence objects are
// built on its threads (same objects, same order).
void parse_sequences(const std::string& path, SequenceFormat format,
                     std::vector<std::unique_ptr<Sequence>>& dst, ThreadPool* pool = nullptr);
void parse_overlaps(const std::string& path, OverlapFormat format,
                    std::vector<std::unique_ptr<Overlap>>& dst);

}  // namespace racon
