#include "sequence.hpp"

#include <cctype>

namespace racon {

std::unique_ptr<Sequence> createSequence(const std::string& name,
                                         const std::string& data) {
  return std::unique_ptr<Sequence>(new Sequence(name, data));
}

namespace {

// toupper of the "C" locale (the only one racon runs in) as a byte table
struct UpperTable {
  char map[256];
  UpperTable() {
    for (int c = 0; c < 256; ++c) {
      map[c] = static_cast<char>(c >= 'a' && c <= 'z' ? c - 'a' + 'A' : c);
    }
  }
};
const UpperTable kUpper;

}  // namespace

Sequence::Sequence(const char* name, uint32_t name_length, const char* data,
                   uint32_t data_length)
    : name_(name, name_length) {
  // sequence.cpp:26: upper case
  data_.resize(data_length);
  char* out = &data_[0];
  for (uint32_t i = 0; i < data_length; ++i) {
    out[i] = kUpper.map[static_cast<unsigned char>(data[i])];
  }
}

Sequence::Sequence(const char* name, uint32_t name_length, const char* data,
                   uint32_t data_length, const char* quality,
                   uint32_t quality_length)
    : Sequence(name, name_length, data, data_length) {
  // sequence.cpp:34-41: keep the quality only if some byte is above '!'
  uint32_t sum = 0;
  for (uint32_t i = 0; i < quality_length; ++i) {
    sum += static_cast<uint32_t>(quality[i] - '!');
  }
  if (sum > 0) {
    quality_.assign(quality, quality_length);
  }
}

Sequence::Sequence(const std::string& name, const std::string& data)
    : name_(name), data_(data) {}

void Sequence::create_reverse_complement() {
  if (!reverse_complement_.empty()) {
    return;
  }
  const size_t n = data_.size();
  reverse_complement_.resize(n);
  for (size_t i = 0; i < n; ++i) {
    char c = data_[n - 1 - i];
    switch (c) {
      case 'A': c = 'T'; break;
      case 'T': c = 'A'; break;
      case 'C': c = 'G'; break;
      case 'G': c = 'C'; break;
      default: break;
    }
    reverse_complement_[i] = c;
  }
  reverse_quality_.assign(quality_.rbegin(), quality_.rend());
}

void Sequence::transmute(bool has_name, bool has_data, bool has_reverse_data) {
  // sequence.cpp:86-100
  if (!has_name) {
    std::string().swap(name_);
  }
  if (has_reverse_data) {
    create_reverse_complement();
  }
  if (!has_data) {
    std::string().swap(data_);
    std::string().swap(quality_);
  }
}

}  // namespace racon
