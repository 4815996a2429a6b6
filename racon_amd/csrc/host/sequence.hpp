// Sequence record (FASTA/FASTQ) of the racon host path.
//
// Behaviour follows racon src/sequence.hpp:28-74 and src/sequence.cpp:13-100:
// data is upper-cased on load, a FASTQ quality string whose every byte is '!'
// is dropped, the reverse complement maps only A/C/G/T (other bytes are kept)
// and carries a reversed quality string.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

namespace racon {

class Sequence;
std::unique_ptr<Sequence> createSequence(const std::string& name,
                                         const std::string& data);

class Sequence {
 public:
  // FASTA record (sequence.cpp:19-28)
  Sequence(const char* name, uint32_t name_length, const char* data,
           uint32_t data_length);
  // FASTQ record (sequence.cpp:30-42)
  Sequence(const char* name, uint32_t name_length, const char* data,
           uint32_t data_length, const char* quality, uint32_t quality_length);
  // polished output record (sequence.cpp:44-47)
  Sequence(const std::string& name, const std::string& data);

  Sequence(const Sequence&) = delete;
  Sequence& operator=(const Sequence&) = delete;

  const std::string& name() const { return name_; }
  const std::string& data() const { return data_; }
  const std::string& reverse_complement() const { return reverse_complement_; }
  const std::string& quality() const { return quality_; }
  const std::string& reverse_quality() const { return reverse_quality_; }

  void create_reverse_complement();
  void transmute(bool has_name, bool has_data, bool has_reverse_data);

 private:
  std::string name_;
  std::string data_;
  std::string reverse_complement_;
  std::string quality_;
  std::string reverse_quality_;
};

}  // namespace racon
