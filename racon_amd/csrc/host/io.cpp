#include "io.hpp"

#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "overlap.hpp"
#include "sequence.hpp"

namespace racon {

bool is_suffix(const std::string& src, const std::string& suffix) {
  if (src.size() < suffix.size()) {
    return false;
  }
  return src.compare(src.size() - suffix.size(), suffix.size(), suffix) == 0;
}

bool sequence_format_of(const std::string& path, SequenceFormat* format) {
  for (const char* ext : {".fasta", ".fasta.gz", ".fna", ".fna.gz", ".fa", ".fa.gz"}) {
    if (is_suffix(path, ext)) {
      *format = SequenceFormat::kFasta;
      return true;
    }
  }
  for (const char* ext : {".fastq", ".fastq.gz", ".fq", ".fq.gz"}) {
    if (is_suffix(path, ext)) {
      *format = SequenceFormat::kFastq;
      return true;
    }
  }
  return false;
}

bool overlap_format_of(const std::string& path, OverlapFormat* format) {
  if (is_suffix(path, ".mhap") || is_suffix(path, ".mhap.gz")) {
    *format = OverlapFormat::kMhap;
  } else if (is_suffix(path, ".paf") || is_suffix(path, ".paf.gz")) {
    *format = OverlapFormat::kPaf;
  } else if (is_suffix(path, ".sam") || is_suffix(path, ".sam.gz")) {
    *format = OverlapFormat::kSam;
  } else {
    return false;
  }
  return true;
}

std::string read_file(const std::string& path) {
  gzFile f = gzopen(path.c_str(), "rb");
  if (f == nullptr) {
    fprintf(stderr, "[racon::io] error: unable to open file %s!\n", path.c_str());
    exit(1);
  }
  gzbuffer(f, 1 << 20);
  std::string out;
  std::vector<char> buf(1 << 22);
  while (true) {
    int n = gzread(f, buf.data(), static_cast<unsigned>(buf.size()));
    if (n < 0) {
      fprintf(stderr, "[racon::io] error: unable to read file %s!\n", path.c_str());
      exit(1);
    }
    if (n == 0) {
      break;
    }
    out.append(buf.data(), n);
  }
  gzclose(f);
  return out;
}

namespace {

// Line cursor over an in-memory file.
struct Lines {
  const std::string& text;
  size_t pos = 0;

  explicit Lines(const std::string& t) : text(t) {}

  // Next line without the newline and trailing whitespace; false at end.
  bool next(const char** begin, size_t* length) {
    if (pos >= text.size()) {
      return false;
    }
    size_t end = text.find('\n', pos);
    if (end == std::string::npos) {
      end = text.size();
    }
    size_t stop = end;
    while (stop > pos && (text[stop - 1] == '\r' || text[stop - 1] == ' ' ||
                          text[stop - 1] == '\t')) {
      --stop;
    }
    *begin = text.data() + pos;
    *length = stop - pos;
    pos = end + 1;
    return true;
  }
};

size_t name_length(const char* s, size_t n) {
  size_t i = 0;
  while (i < n && s[i] != ' ' && s[i] != '\t') {
    ++i;
  }
  return i;
}

// Splits a line on the given delimiter set into at most max_fields fields.
std::vector<std::pair<const char*, size_t>> split(const char* s, size_t n,
                                                  const char* delims) {
  std::vector<std::pair<const char*, size_t>> fields;
  size_t i = 0;
  while (i <= n) {
    size_t j = i;
    while (j < n && strchr(delims, s[j]) == nullptr) {
      ++j;
    }
    fields.emplace_back(s + i, j - i);
    if (j >= n) {
      break;
    }
    i = j + 1;
    if (delims[0] == ' ') {  // whitespace separated: collapse runs
      while (i < n && strchr(delims, s[i]) != nullptr) {
        ++i;
      }
    }
  }
  return fields;
}

uint64_t to_u64(const std::pair<const char*, size_t>& f) {
  uint64_t v = 0;
  for (size_t i = 0; i < f.second; ++i) {
    char c = f.first[i];
    if (c < '0' || c > '9') {
      break;
    }
    v = v * 10 + static_cast<uint64_t>(c - '0');
  }
  return v;
}

double to_double(const std::pair<const char*, size_t>& f) {
  return atof(std::string(f.first, f.second).c_str());
}

void malformed(const std::string& path) {
  fprintf(stderr, "[racon::io] error: invalid file format %s!\n", path.c_str());
  exit(1);
}

}  // namespace

void parse_sequences(const std::string& path, SequenceFormat format,
                     std::vector<std::unique_ptr<Sequence>>& dst) {
  std::string text = read_file(path);
  Lines lines(text);
  const char* line;
  size_t length;
  if (format == SequenceFormat::kFasta) {
    std::string name, data;
    bool open = false;
    auto flush = [&]() {
      if (open) {
        dst.emplace_back(new Sequence(name.data(), name.size(), data.data(),
                                      data.size()));
      }
    };
    while (lines.next(&line, &length)) {
      if (length > 0 && line[0] == '>') {
        flush();
        name.assign(line + 1, name_length(line + 1, length - 1));
        data.clear();
        open = true;
      } else if (open) {
        data.append(line, length);
      } else if (length > 0) {
        malformed(path);
      }
    }
    flush();
  } else {
    std::string name, data, quality;
    while (lines.next(&line, &length)) {
      if (length == 0) {
        continue;
      }
      if (line[0] != '@') {
        malformed(path);
      }
      name.assign(line + 1, name_length(line + 1, length - 1));
      data.clear();
      quality.clear();
      bool plus = false;
      while (lines.next(&line, &length)) {
        if (length > 0 && line[0] == '+') {
          plus = true;
          break;
        }
        data.append(line, length);
      }
      if (!plus) {
        malformed(path);
      }
      while (quality.size() < data.size() && lines.next(&line, &length)) {
        quality.append(line, length);
      }
      if (quality.size() != data.size()) {
        malformed(path);
      }
      dst.emplace_back(new Sequence(name.data(), name.size(), data.data(),
                                    data.size(), quality.data(), quality.size()));
    }
  }
}

void parse_overlaps(const std::string& path, OverlapFormat format,
                    std::vector<std::unique_ptr<Overlap>>& dst) {
  std::string text = read_file(path);
  Lines lines(text);
  const char* line;
  size_t length;
  while (lines.next(&line, &length)) {
    if (length == 0) {
      continue;
    }
    if (format == OverlapFormat::kMhap) {
      auto f = split(line, length, " \t");
      if (f.size() < 12) {
        malformed(path);
      }
      dst.emplace_back(new Overlap(
          to_u64(f[0]), to_u64(f[1]), to_double(f[2]), to_u64(f[3]),
          to_u64(f[4]), to_u64(f[5]), to_u64(f[6]), to_u64(f[7]), to_u64(f[8]),
          to_u64(f[9]), to_u64(f[10]), to_u64(f[11])));
    } else if (format == OverlapFormat::kPaf) {
      auto f = split(line, length, "\t");
      if (f.size() < 12) {
        malformed(path);
      }
      dst.emplace_back(new Overlap(
          f[0].first, f[0].second, to_u64(f[1]), to_u64(f[2]), to_u64(f[3]),
          f[4].second > 0 ? f[4].first[0] : '+', f[5].first, f[5].second,
          to_u64(f[6]), to_u64(f[7]), to_u64(f[8]), to_u64(f[9]),
          to_u64(f[10]), to_u64(f[11])));
    } else {
      if (line[0] == '@') {
        continue;
      }
      auto f = split(line, length, "\t");
      if (f.size() < 11) {
        malformed(path);
      }
      dst.emplace_back(new Overlap(f[0].first, f[0].second, to_u64(f[1]),
                                   f[2].first, f[2].second, to_u64(f[3]),
                                   to_u64(f[4]), f[5].first, f[5].second));
    }
  }
}

}  // namespace racon
