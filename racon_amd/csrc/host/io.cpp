#include "io.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "overlap.hpp"
#include "sequence.hpp"
#include "thread_pool.hpp"

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

namespace {

std::string inflate_file(const std::string& path) {
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

}  // namespace

FileText::FileText(const std::string& path) {
  const int fd = open(path.c_str(), O_RDONLY);
  if (fd < 0) {
    fprintf(stderr, "[racon::io] error: unable to open file %s!\n", path.c_str());
    exit(1);
  }
  struct stat st {};
  unsigned char magic[2] = {0, 0};
  const bool sized = fstat(fd, &st) == 0 && S_ISREG(st.st_mode);
  const bool gz = pread(fd, magic, 2, 0) == 2 && magic[0] == 0x1f && magic[1] == 0x8b;
  if (sized && !gz && st.st_size > 0) {
    void* m = mmap(nullptr, static_cast<size_t>(st.st_size), PROT_READ,
                   MAP_PRIVATE | MAP_POPULATE, fd, 0);
    if (m != MAP_FAILED) {
      map_ = m;
      map_size_ = static_cast<size_t>(st.st_size);
      data_ = static_cast<const char*>(m);
      size_ = map_size_;
    }
  }
  close(fd);
  if (map_ == nullptr && !(sized && !gz && st.st_size == 0)) {
    inflated_ = inflate_file(path);  // gzip (or a stream that cannot be mapped)
    data_ = inflated_.data();
    size_ = inflated_.size();
  }
}

FileText::~FileText() {
  if (map_ != nullptr) munmap(map_, map_size_);
}

namespace {

// Line cursor over an in-memory file.
struct Lines {
  const char* text;
  size_t size;
  size_t pos = 0;

  explicit Lines(const FileText& t) : text(t.data()), size(t.size()) {}
  Lines(const char* t, size_t n) : text(t), size(n) {}

  // Next line without the newline and trailing whitespace; false at end.
  bool next(const char** begin, size_t* length) {
    if (pos >= size) {
      return false;
    }
    const void* nl = memchr(text + pos, '\n', size - pos);
    const size_t end = nl != nullptr ? static_cast<size_t>(static_cast<const char*>(nl) - text)
                                     : size;
    size_t stop = end;
    while (stop > pos && (text[stop - 1] == '\r' || text[stop - 1] == ' ' ||
                          text[stop - 1] == '\t')) {
      --stop;
    }
    *begin = text + pos;
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

namespace {

// A parsed record before its Sequence is built: the name and the line spans
// of its data (and quality), as [first, last) into a flat span list.
struct Span {
  const char* p;
  size_t n;
};
struct Record {
  Span name;
  uint32_t d0, d1, q0, q1;
};

// Bytes of spans [a, b) as one buffer: the line itself when there is one,
// else joined into `buf`.
Span joined(const std::vector<Span>& spans, uint32_t a, uint32_t b, std::string& buf) {
  if (b == a + 1) return spans[a];
  buf.clear();
  for (uint32_t i = a; i < b; ++i) buf.append(spans[i].p, spans[i].n);
  return Span{buf.data(), buf.size()};
}

}  // namespace

// One serial pass over the lines finds the records (and every malformed
// one, in file order, as before); the Sequence objects (upper-casing,
// copies, the quality check) are then built in parallel on `pool`.
void parse_sequences(const std::string& path, SequenceFormat format,
                     std::vector<std::unique_ptr<Sequence>>& dst, ThreadPool* pool) {
  FileText text(path);
  Lines lines(text);
  const char* line;
  size_t length;
  std::vector<Record> recs;
  std::vector<Span> spans;
  const bool fastq = format != SequenceFormat::kFasta;
  if (!fastq) {
    bool open = false;
    while (lines.next(&line, &length)) {
      if (length > 0 && line[0] == '>') {
        if (open) recs.back().d1 = static_cast<uint32_t>(spans.size());
        recs.push_back(Record{Span{line + 1, name_length(line + 1, length - 1)},
                              static_cast<uint32_t>(spans.size()), 0, 0, 0});
        open = true;
      } else if (open) {
        spans.push_back(Span{line, length});
      } else if (length > 0) {
        malformed(path);
      }
    }
    if (open) recs.back().d1 = static_cast<uint32_t>(spans.size());
  } else {
    while (lines.next(&line, &length)) {
      if (length == 0) {
        continue;
      }
      if (line[0] != '@') {
        malformed(path);
      }
      Record r{Span{line + 1, name_length(line + 1, length - 1)},
               static_cast<uint32_t>(spans.size()), 0, 0, 0};
      size_t data_size = 0;
      bool plus = false;
      while (lines.next(&line, &length)) {
        if (length > 0 && line[0] == '+') {
          plus = true;
          break;
        }
        spans.push_back(Span{line, length});
        data_size += length;
      }
      if (!plus) {
        malformed(path);
      }
      r.d1 = r.q0 = static_cast<uint32_t>(spans.size());
      size_t quality_size = 0;
      while (quality_size < data_size && lines.next(&line, &length)) {
        spans.push_back(Span{line, length});
        quality_size += length;
      }
      if (quality_size != data_size) {
        malformed(path);
      }
      r.q1 = static_cast<uint32_t>(spans.size());
      recs.push_back(r);
    }
  }

  const size_t base = dst.size();
  dst.resize(base + recs.size());
  auto build = [&](size_t begin, size_t end) {
    std::string dbuf, qbuf;
    for (size_t i = begin; i < end; ++i) {
      const Record& r = recs[i];
      const Span d = r.d1 > r.d0 ? joined(spans, r.d0, r.d1, dbuf) : Span{"", 0};
      if (fastq) {
        const Span q = r.q1 > r.q0 ? joined(spans, r.q0, r.q1, qbuf) : Span{"", 0};
        dst[base + i].reset(new Sequence(r.name.p, r.name.n, d.p, d.n, q.p, q.n));
      } else {
        dst[base + i].reset(new Sequence(r.name.p, r.name.n, d.p, d.n));
      }
    }
  };
  const size_t workers = pool != nullptr ? pool->size() : 1;
  if (workers <= 1 || recs.size() < 64) {
    build(0, recs.size());
    return;
  }
  const size_t chunk = (recs.size() + 4 * workers - 1) / (4 * workers);
  std::vector<std::future<void>> done;
  for (size_t b = 0; b < recs.size(); b += chunk) {
    done.emplace_back(pool->submit(build, b, std::min(recs.size(), b + chunk)));
  }
  for (auto& f : done) f.wait();
}

void parse_overlaps(const std::string& path, OverlapFormat format,
                    std::vector<std::unique_ptr<Overlap>>& dst, ThreadPool* pool) {
  FileText text(path);
  // one line's overlap (none for an empty or header line); false: malformed
  auto parse_line = [format](const char* line, size_t length,
                             std::unique_ptr<Overlap>* out) -> bool {
    if (length == 0) {
      return true;
    }
    if (format == OverlapFormat::kMhap) {
      auto f = split(line, length, " \t");
      if (f.size() < 12) {
        return false;
      }
      out->reset(new Overlap(
          to_u64(f[0]), to_u64(f[1]), to_double(f[2]), to_u64(f[3]),
          to_u64(f[4]), to_u64(f[5]), to_u64(f[6]), to_u64(f[7]), to_u64(f[8]),
          to_u64(f[9]), to_u64(f[10]), to_u64(f[11])));
    } else if (format == OverlapFormat::kPaf) {
      auto f = split(line, length, "\t");
      if (f.size() < 12) {
        return false;
      }
      out->reset(new Overlap(
          f[0].first, f[0].second, to_u64(f[1]), to_u64(f[2]), to_u64(f[3]),
          f[4].second > 0 ? f[4].first[0] : '+', f[5].first, f[5].second,
          to_u64(f[6]), to_u64(f[7]), to_u64(f[8]), to_u64(f[9]),
          to_u64(f[10]), to_u64(f[11])));
    } else {
      if (line[0] == '@') {
        return true;
      }
      auto f = split(line, length, "\t");
      if (f.size() < 11) {
        return false;
      }
      out->reset(new Overlap(f[0].first, f[0].second, to_u64(f[1]),
                             f[2].first, f[2].second, to_u64(f[3]),
                             to_u64(f[4]), f[5].first, f[5].second));
    }
    return true;
  };
  // the text cut into ranges at line starts, parsed on the pool; the
  // overlaps are then appended range by range (file order)
  const size_t workers = pool != nullptr ? pool->size() : 1;
  const size_t ranges = workers <= 1 || text.size() < (1u << 20) ? 1 : 4 * workers;
  std::vector<size_t> cuts{0};
  for (size_t k = 1; k < ranges; ++k) {
    const size_t at = std::max(cuts.back(), text.size() / ranges * k);
    const void* nl = at < text.size() ? memchr(text.data() + at, '\n', text.size() - at) : nullptr;
    const size_t cut = nl != nullptr
                           ? static_cast<size_t>(static_cast<const char*>(nl) - text.data()) + 1
                           : text.size();
    if (cut > cuts.back() && cut < text.size()) cuts.push_back(cut);
  }
  cuts.push_back(text.size());
  std::vector<std::vector<std::unique_ptr<Overlap>>> parts(cuts.size() - 1);
  std::vector<uint8_t> bad(parts.size(), 0);
  auto run = [&](size_t r) {
    Lines lines(text.data() + cuts[r], cuts[r + 1] - cuts[r]);
    const char* line;
    size_t length;
    while (lines.next(&line, &length)) {
      std::unique_ptr<Overlap> o;
      if (!parse_line(line, length, &o)) {
        bad[r] = 1;
        return;
      }
      if (o != nullptr) parts[r].emplace_back(std::move(o));
    }
  };
  if (parts.size() == 1) {
    run(0);
  } else {
    std::vector<std::future<void>> done;
    for (size_t r = 0; r < parts.size(); ++r) done.emplace_back(pool->submit(run, r));
    for (auto& f : done) f.wait();
  }
  size_t total = dst.size();
  for (size_t r = 0; r < parts.size(); ++r) {
    if (bad[r]) {
      malformed(path);
    }
    total += parts[r].size();
  }
  dst.reserve(total);
  for (auto& part : parts) {
    for (auto& o : part) dst.emplace_back(std::move(o));
  }
}

}  // namespace racon
