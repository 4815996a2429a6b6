#include "polisher.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <future>
#include <unordered_map>

#include "logger.hpp"
#include "overlap.hpp"
#include "sequence.hpp"
#include "thread_pool.hpp"
#include "window.hpp"

namespace racon {

namespace {

// polisher.cpp:28-53: compacts nulls out of src[begin..], keeping order.
template <class T>
uint64_t shrinkToFit(std::vector<std::unique_ptr<T>>& src, uint64_t begin) {
  uint64_t i = begin;
  for (uint64_t j = begin; i < src.size(); ++i) {
    if (src[i] != nullptr) {
      continue;
    }
    j = std::max(j, i);
    while (j < src.size() && src[j] == nullptr) {
      ++j;
    }
    if (j >= src.size()) {
      break;
    } else if (i != j) {
      src[i].swap(src[j]);
    }
  }
  uint64_t num_deletions = src.size() - i;
  if (i < src.size()) {
    src.resize(i);
  }
  return num_deletions;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

std::unique_ptr<Polisher> createPolisher(
    const std::string& sequences_path, const std::string& overlaps_path,
    const std::string& target_path, PolisherType type, uint32_t window_length,
    double quality_threshold, double error_threshold, bool trim, int8_t match,
    int8_t mismatch, int8_t gap, uint32_t num_threads, uint32_t cuda_batches,
    bool cuda_banded_alignment, uint32_t cudaaligner_batches,
    uint32_t cudaaligner_band_width, void* poa_server) {
  // polisher.cpp:62-133 (messages kept verbatim)
  if (type != PolisherType::kC && type != PolisherType::kF) {
    fprintf(stderr, "[racon::createPolisher] error: invalid polisher type!\n");
    exit(1);
  }
  if (window_length == 0) {
    fprintf(stderr, "[racon::createPolisher] error: invalid window length!\n");
    exit(1);
  }
  SequenceFormat sformat, tformat;
  OverlapFormat oformat;
  if (!sequence_format_of(sequences_path, &sformat)) {
    fprintf(stderr, "[racon::createPolisher] error: "
                    "file %s has unsupported format extension (valid extensions: "
                    ".fasta, .fasta.gz, .fna, .fna.gz, .fa, .fa.gz, .fastq, .fastq.gz, "
                    ".fq, .fq.gz)!\n",
            sequences_path.c_str());
    exit(1);
  }
  if (!overlap_format_of(overlaps_path, &oformat)) {
    fprintf(stderr, "[racon::createPolisher] error: "
                    "file %s has unsupported format extension (valid extensions: "
                    ".mhap, .mhap.gz, .paf, .paf.gz, .sam, .sam.gz)!\n",
            overlaps_path.c_str());
    exit(1);
  }
  if (!sequence_format_of(target_path, &tformat)) {
    fprintf(stderr, "[racon::createPolisher] error: "
                    "file %s has unsupported format extension (valid extensions: "
                    ".fasta, .fasta.gz, .fna, .fna.gz, .fa, .fa.gz, .fastq, .fastq.gz, "
                    ".fq, .fq.gz)!\n",
            target_path.c_str());
    exit(1);
  }

  EngineConfig config;
  config.match = match;
  config.mismatch = mismatch;
  config.gap = gap;
  config.num_threads = num_threads;
  config.window_length = window_length;
  config.poa_batches = cuda_batches;
  config.aligner_batches = cudaaligner_batches;
  config.banded = cuda_banded_alignment;
  config.aligner_band_width = cudaaligner_band_width;
  config.poa_server = poa_server;

  return std::unique_ptr<Polisher>(new Polisher(
      sequences_path, sformat, overlaps_path, oformat, target_path, tformat, type,
      window_length, quality_threshold, error_threshold, trim, num_threads,
      createEngine(config)));
}

Polisher::Polisher(const std::string& sequences_path,
                   SequenceFormat sequences_format,
                   const std::string& overlaps_path, OverlapFormat overlaps_format,
                   const std::string& target_path, SequenceFormat target_format,
                   PolisherType type, uint32_t window_length,
                   double quality_threshold, double error_threshold, bool trim,
                   uint32_t num_threads, std::unique_ptr<PolishEngine> engine)
    : sequences_path_(sequences_path), sequences_format_(sequences_format),
      overlaps_path_(overlaps_path), overlaps_format_(overlaps_format),
      target_path_(target_path), target_format_(target_format), type_(type),
      quality_threshold_(quality_threshold), error_threshold_(error_threshold),
      trim_(trim), dummy_quality_(window_length, '!'),
      window_length_(window_length),
      thread_pool_(new ThreadPool(num_threads)), logger_(new Logger()),
      engine_(std::move(engine)) {}

Polisher::~Polisher() { logger_->total("[racon::Polisher::] total ="); }

void Polisher::initialize() {
  // polisher.cpp:192-460
  if (!windows_.empty()) {
    fprintf(stderr, "[racon::Polisher::initialize] warning: "
                    "object already initialized!\n");
    return;
  }
  auto t_load = std::chrono::steady_clock::now();
  logger_->log();

  parse_sequences(target_path_, target_format_, sequences_, thread_pool_.get());
  uint64_t targets_size = sequences_.size();
  if (targets_size == 0) {
    fprintf(stderr, "[racon::Polisher::initialize] error: "
                    "empty target sequences set!\n");
    exit(1);
  }

  std::unordered_map<std::string, uint64_t> name_to_id;
  std::unordered_map<uint64_t, uint64_t> id_to_id;
  for (uint64_t i = 0; i < targets_size; ++i) {
    name_to_id[sequences_[i]->name() + "t"] = i;
    id_to_id[i << 1 | 1] = i;
  }

  std::vector<bool> has_name(targets_size, true);
  std::vector<bool> has_data(targets_size, true);
  std::vector<bool> has_reverse_data(targets_size, false);

  logger_->log("[racon::Polisher::initialize] loaded target sequences");
  logger_->log();

  uint64_t sequences_size = 0, total_sequences_length = 0;
  {
    uint64_t l = sequences_.size();
    parse_sequences(sequences_path_, sequences_format_, sequences_, thread_pool_.get());
    uint64_t n = 0;
    for (uint64_t i = l; i < sequences_.size(); ++i, ++sequences_size) {
      total_sequences_length += sequences_[i]->data().size();
      auto it = name_to_id.find(sequences_[i]->name() + "t");
      if (it != name_to_id.end()) {
        if (sequences_[i]->data().size() != sequences_[it->second]->data().size() ||
            sequences_[i]->quality().size() !=
                sequences_[it->second]->quality().size()) {
          fprintf(stderr, "[racon::Polisher::initialize] error: "
                          "duplicate sequence %s with unequal data\n",
                  sequences_[i]->name().c_str());
          exit(1);
        }
        name_to_id[sequences_[i]->name() + "q"] = it->second;
        id_to_id[sequences_size << 1 | 0] = it->second;
        sequences_[i].reset();
        ++n;
      } else {
        name_to_id[sequences_[i]->name() + "q"] = i - n;
        id_to_id[sequences_size << 1 | 0] = i - n;
      }
    }
    shrinkToFit(sequences_, l);
  }

  if (sequences_size == 0) {
    fprintf(stderr, "[racon::Polisher::initialize] error: empty sequences set!\n");
    exit(1);
  }

  has_name.resize(sequences_.size(), false);
  has_data.resize(sequences_.size(), false);
  has_reverse_data.resize(sequences_.size(), false);

  WindowType window_type =
      static_cast<double>(total_sequences_length) / sequences_size <= 1000
          ? WindowType::kNGS
          : WindowType::kTGS;

  logger_->log("[racon::Polisher::initialize] loaded sequences");
  logger_->log();

  std::vector<std::unique_ptr<Overlap>> overlaps;

  auto remove_invalid_overlaps = [&](uint64_t begin, uint64_t end) -> void {
    // polisher.cpp:284-308
    for (uint64_t i = begin; i < end; ++i) {
      if (overlaps[i] == nullptr) {
        continue;
      }
      if (overlaps[i]->error() > error_threshold_ ||
          overlaps[i]->q_id() == overlaps[i]->t_id()) {
        overlaps[i].reset();
        continue;
      }
      if (type_ == PolisherType::kC) {
        for (uint64_t j = i + 1; j < end; ++j) {
          if (overlaps[j] == nullptr) {
            continue;
          }
          if (overlaps[i]->length() > overlaps[j]->length()) {
            overlaps[j].reset();
          } else {
            overlaps[i].reset();
            break;
          }
        }
      }
    }
  };

  {
    // polisher.cpp:310-355 with the whole file as one chunk
    parse_overlaps(overlaps_path_, overlaps_format_, overlaps, thread_pool_.get());
    uint64_t c = 0;
    for (uint64_t i = 0; i < overlaps.size(); ++i) {
      overlaps[i]->transmute(sequences_, name_to_id, id_to_id);
      if (!overlaps[i]->is_valid()) {
        overlaps[i].reset();
        continue;
      }
      while (overlaps[c] == nullptr) {
        ++c;
      }
      if (overlaps[c]->q_id() != overlaps[i]->q_id()) {
        remove_invalid_overlaps(c, i);
        c = i;
      }
    }
    remove_invalid_overlaps(c, overlaps.size());
    for (uint64_t i = 0; i < overlaps.size(); ++i) {
      if (overlaps[i] == nullptr) {
        continue;
      }
      if (overlaps[i]->strand()) {
        has_reverse_data[overlaps[i]->q_id()] = true;
      } else {
        has_data[overlaps[i]->q_id()] = true;
      }
    }
    shrinkToFit(overlaps, 0);
  }

  std::unordered_map<std::string, uint64_t>().swap(name_to_id);
  std::unordered_map<uint64_t, uint64_t>().swap(id_to_id);

  if (overlaps.empty()) {
    fprintf(stderr, "[racon::Polisher::initialize] error: empty overlap set!\n");
    exit(1);
  }

  logger_->log("[racon::Polisher::initialize] loaded overlaps");
  logger_->log();

  {
    std::vector<std::future<void>> futures;
    for (uint64_t i = 0; i < sequences_.size(); ++i) {
      futures.emplace_back(thread_pool_->submit(
          [&](uint64_t j) -> void {
            sequences_[j]->transmute(has_name[j], has_data[j], has_reverse_data[j]);
          },
          i));
    }
    for (const auto& it : futures) {
      it.wait();
    }
  }
  timings_.load = seconds_since(t_load);

  auto t_align = std::chrono::steady_clock::now();
  std::vector<LayerRecord> device_layers;
  const bool on_device = align_layers_on_device(overlaps, targets_size, &device_layers);
  if (!on_device) {
    find_overlap_breaking_points(overlaps);
  }
  timings_.align = seconds_since(t_align);

  logger_->log();
  auto t_windows = std::chrono::steady_clock::now();

  // device layers come grouped by window: window w's are records
  // [layer_at[w], layer_at[w + 1])
  std::vector<uint64_t> layer_at;
  if (on_device) {
    uint64_t n_windows = 0;
    for (uint64_t i = 0; i < targets_size; ++i) {
      n_windows += (sequences_[i]->data().size() + window_length_ - 1) / window_length_;
    }
    layer_at.assign(n_windows + 1, 0);
    for (const LayerRecord& l : device_layers) ++layer_at[l.window + 1];
    for (uint64_t w = 0; w < n_windows; ++w) layer_at[w + 1] += layer_at[w];
  }

  // polisher.cpp:384-457
  std::vector<uint64_t> id_to_first_window_id(targets_size + 1, 0);
  for (uint64_t i = 0; i < targets_size; ++i) {
    uint32_t k = 0;
    for (uint32_t j = 0; j < sequences_[i]->data().size(); j += window_length_, ++k) {
      uint32_t length = std::min(j + window_length_,
                                 static_cast<uint32_t>(sequences_[i]->data().size())) - j;
      const uint64_t w = windows_.size();
      windows_.emplace_back(createWindow(
          i, k, window_type, &(sequences_[i]->data()[j]), length,
          sequences_[i]->quality().empty() ? &(dummy_quality_[0])
                                           : &(sequences_[i]->quality()[j]),
          length,
          on_device ? static_cast<uint32_t>(layer_at[w + 1] - layer_at[w]) : 0));
    }
    id_to_first_window_id[i + 1] = id_to_first_window_id[i] + k;
  }

  targets_coverages_.resize(targets_size, 0);

  if (on_device) {
    // the engine applied the filters and grouped the layers by window, in
    // overlap order: each window's layers are added by one worker
    const double create_s = seconds_since(t_windows);
    for (uint64_t i = 0; i < overlaps.size(); ++i) {
      ++targets_coverages_[overlaps[i]->t_id()];
    }
    // each overlap's bytes in its aligned orientation
    std::vector<const char*> data_at(overlaps.size()), quality_at(overlaps.size());
    auto bases = [&](uint64_t begin, uint64_t end) {
      for (uint64_t i = begin; i < end; ++i) {
        const Overlap& o = *overlaps[i];
        const auto& sequence = sequences_[o.q_id()];
        data_at[i] = o.strand() ? sequence->reverse_complement().data() : sequence->data().data();
        const std::string& q = o.strand() ? sequence->reverse_quality() : sequence->quality();
        quality_at[i] = q.empty() ? nullptr : q.data();
      }
    };
    auto attach = [&](uint64_t begin, uint64_t end) {  // windows [begin, end)
      for (uint64_t r = layer_at[begin]; r < layer_at[end]; ++r) {
        const LayerRecord& l = device_layers[r];
        const uint32_t data_length = l.q_end - l.q_begin;
        const char* quality = quality_at[l.overlap];
        windows_[l.window]->add_layer(data_at[l.overlap] + l.q_begin, data_length,
                                      quality == nullptr ? nullptr : quality + l.q_begin,
                                      quality == nullptr ? 0 : data_length, l.begin, l.end);
      }
    };
    const uint64_t n = device_layers.size();
    const uint64_t workers = thread_pool_->size();
    auto fan_out = [&](const std::function<void(uint64_t, uint64_t)>& f, uint64_t items) {
      if (workers <= 1 || items < 4096) {
        f(0, items);
        return;
      }
      const uint64_t step = (items + 4 * workers - 1) / (4 * workers);
      std::vector<std::future<void>> done;
      for (uint64_t b = 0; b < items; b += step) {
        done.emplace_back(thread_pool_->submit(f, b, std::min(items, b + step)));
      }
      for (auto& d : done) d.wait();
    };
    fan_out(bases, overlaps.size());
    fan_out(attach, windows_.size());
    const double attach_s = seconds_since(t_windows) - create_s;
    for (auto& o : overlaps) o.reset();
    timings_.windows = seconds_since(t_windows);
    fprintf(stderr,
            "[racon_amd::windows] %lu windows: created in %.1f ms, %lu device layers added in "
            "%.1f ms, overlaps released in %.1f ms\n",
            static_cast<unsigned long>(windows_.size()), 1e3 * create_s,
            static_cast<unsigned long>(n), 1e3 * attach_s,
            1e3 * (timings_.windows - create_s - attach_s));
    timings_.num_windows = windows_.size();
    logger_->log("[racon::Polisher::initialize] transformed data into windows");
    return;
  }

  // the two layer filters of each breaking-point pair (span length, mean
  // quality), decided on the pool; the layers are then added in overlap order
  std::vector<std::vector<uint8_t>> passes(overlaps.size());
  auto decide = [&](uint64_t begin, uint64_t end) {
    for (uint64_t i = begin; i < end; ++i) {
      const auto& sequence = sequences_[overlaps[i]->q_id()];
      const auto& breaking_points = overlaps[i]->breaking_points();
      auto& pass = passes[i];
      pass.assign(breaking_points.size() / 2, 0);
      for (uint32_t j = 0; j < breaking_points.size(); j += 2) {
        if (breaking_points[j + 1].second - breaking_points[j].second <
            0.02 * window_length_) {
          continue;
        }
        if (!sequence->quality().empty() || !sequence->reverse_quality().empty()) {
          const auto& quality = overlaps[i]->strand() ? sequence->reverse_quality()
                                                      : sequence->quality();
          double average_quality = 0;
          for (uint32_t k = breaking_points[j].second;
               k < breaking_points[j + 1].second; ++k) {
            average_quality += static_cast<uint32_t>(quality[k]) - 33;
          }
          average_quality /= breaking_points[j + 1].second - breaking_points[j].second;
          if (average_quality < quality_threshold_) {
            continue;
          }
        }
        pass[j / 2] = 1;
      }
    }
  };
  {
    const uint64_t n = overlaps.size();
    const uint64_t workers = thread_pool_->size();
    if (workers <= 1 || n < 256) {
      decide(0, n);
    } else {
      const uint64_t chunk = (n + 4 * workers - 1) / (4 * workers);
      std::vector<std::future<void>> done;
      for (uint64_t b = 0; b < n; b += chunk) {
        done.emplace_back(thread_pool_->submit(decide, b, std::min(n, b + chunk)));
      }
      for (auto& f : done) f.wait();
    }
  }

  for (uint64_t i = 0; i < overlaps.size(); ++i) {
    ++targets_coverages_[overlaps[i]->t_id()];

    const auto& sequence = sequences_[overlaps[i]->q_id()];
    const auto& breaking_points = overlaps[i]->breaking_points();

    for (uint32_t j = 0; j < breaking_points.size(); j += 2) {
      if (passes[i][j / 2] == 0) {
        continue;
      }

      uint64_t window_id = id_to_first_window_id[overlaps[i]->t_id()] +
                           breaking_points[j].first / window_length_;
      uint32_t window_start = (breaking_points[j].first / window_length_) *
                              window_length_;

      const char* data =
          overlaps[i]->strand()
              ? &(sequence->reverse_complement()[breaking_points[j].second])
              : &(sequence->data()[breaking_points[j].second]);
      uint32_t data_length = breaking_points[j + 1].second - breaking_points[j].second;

      const char* quality =
          overlaps[i]->strand()
              ? (sequence->reverse_quality().empty()
                     ? nullptr
                     : &(sequence->reverse_quality()[breaking_points[j].second]))
              : (sequence->quality().empty()
                     ? nullptr
                     : &(sequence->quality()[breaking_points[j].second]));
      uint32_t quality_length = quality == nullptr ? 0 : data_length;

      windows_[window_id]->add_layer(data, data_length, quality, quality_length,
                                     breaking_points[j].first - window_start,
                                     breaking_points[j + 1].first - window_start - 1);
    }
    overlaps[i].reset();
  }
  timings_.windows = seconds_since(t_windows);
  timings_.num_windows = windows_.size();

  logger_->log("[racon::Polisher::initialize] transformed data into windows");
}

void Polisher::dump_windows() const {
  // "RHW1", u64 n, then per window: u32 length, u8 tgs, 3 pad bytes,
  // backbone, backbone quality, u32 layers, and per layer (insertion order):
  // u32 length, u32 begin, u32 end, u8 has_quality, bytes, [quality]
  FILE* f = fopen(dump_windows_path_.c_str(), "wb");
  if (f == nullptr) {
    fprintf(stderr, "[racon::Polisher] error: unable to write %s\n",
            dump_windows_path_.c_str());
    exit(1);
  }
  const uint64_t n = windows_.size();
  fwrite("RHW1", 1, 4, f);
  fwrite(&n, 8, 1, f);
  const uint8_t pad[3] = {0, 0, 0};
  for (const auto& w : windows_) {
    const uint32_t len = w->sequence(0).second;
    const uint8_t tgs = w->type() == WindowType::kTGS ? 1 : 0;
    fwrite(&len, 4, 1, f);
    fwrite(&tgs, 1, 1, f);
    fwrite(pad, 1, 3, f);
    fwrite(w->sequence(0).first, 1, len, f);
    fwrite(w->quality(0).first, 1, len, f);
    const uint32_t layers = w->size() - 1;
    fwrite(&layers, 4, 1, f);
    for (uint32_t k = 1; k < w->size(); ++k) {
      const uint32_t l = w->sequence(k).second;
      const uint32_t b = w->position(k).first, e = w->position(k).second;
      const uint8_t hq = w->quality(k).first != nullptr ? 1 : 0;
      fwrite(&l, 4, 1, f);
      fwrite(&b, 4, 1, f);
      fwrite(&e, 4, 1, f);
      fwrite(&hq, 1, 1, f);
      fwrite(w->sequence(k).first, 1, l, f);
      if (hq) {
        fwrite(w->quality(k).first, 1, l, f);
      }
    }
  }
  fclose(f);
}

bool Polisher::align_layers_on_device(std::vector<std::unique_ptr<Overlap>>& overlaps,
                                      uint64_t targets_size,
                                      std::vector<LayerRecord>* layers) {
  if (!engine_->device_layers() || overlaps.empty() || overlaps.size() >= (1ull << 32)) {
    return false;
  }
  for (const auto& o : overlaps) {
    if (!o->needs_alignment()) return false;
  }
  const auto t_inputs = std::chrono::steady_clock::now();
  // window ids as the window loop below gives them (polisher.cpp:384-398)
  std::vector<uint64_t> first_window(targets_size + 1, 0);
  for (uint64_t i = 0; i < targets_size; ++i) {
    const uint64_t length = sequences_[i]->data().size();
    first_window[i + 1] = first_window[i] + (length + window_length_ - 1) / window_length_;
  }
  // the quality bytes of every read an overlap uses: the forward string, or
  // the reversed one when transmute kept only that
  std::vector<uint64_t> quality_at(sequences_.size(), kNoQuality);
  std::vector<uint8_t> stored_reversed(sequences_.size(), 0);
  uint64_t quality_bytes = 0;
  for (const auto& o : overlaps) {
    const uint64_t q = o->q_id();
    if (quality_at[q] != kNoQuality) continue;
    const auto& sequence = sequences_[q];
    if (!sequence->quality().empty()) {
      quality_at[q] = quality_bytes;
      quality_bytes += sequence->quality().size();
    } else if (!sequence->reverse_quality().empty()) {
      quality_at[q] = quality_bytes;
      stored_reversed[q] = 1;
      quality_bytes += sequence->reverse_quality().size();
    }
  }
  // (run by the engine, possibly beside the first alignments)
  auto gather_qualities = [&](std::vector<char>* out) {
    std::vector<char>& qualities = *out;
    qualities.resize(quality_bytes);
    auto copy_qualities = [&](uint64_t begin, uint64_t end) {
      for (uint64_t q = begin; q < end; ++q) {
        if (quality_at[q] == kNoQuality) continue;
        const std::string& s =
            stored_reversed[q] ? sequences_[q]->reverse_quality() : sequences_[q]->quality();
        std::copy(s.begin(), s.end(), qualities.begin() + quality_at[q]);
      }
    };
    const uint64_t n = sequences_.size();
    const uint64_t workers = thread_pool_->size();
    if (workers <= 1 || quality_bytes < (1u << 24)) {
      copy_qualities(0, n);
    } else {
      const uint64_t step = (n + 4 * workers - 1) / (4 * workers);
      std::vector<std::future<void>> copies;
      for (uint64_t b = 0; b < n; b += step) {
        copies.emplace_back(thread_pool_->submit(copy_qualities, b, std::min(n, b + step)));
      }
      for (auto& f : copies) f.wait();
    }
  };

  std::vector<AlignTask> tasks;
  std::vector<uint64_t> task_overlap;
  collect_alignment_tasks(overlaps, &tasks, &task_overlap);
  std::vector<LayerSource> sources(tasks.size());
  for (uint64_t k = 0; k < tasks.size(); ++k) {
    const Overlap& o = *overlaps[task_overlap[k]];
    const auto& sequence = sequences_[o.q_id()];
    LayerSource& src = sources[k];
    src.overlap = static_cast<uint32_t>(task_overlap[k]);
    src.query_length = static_cast<uint32_t>(
        std::max(sequence->data().size(), sequence->reverse_complement().size()));
    src.first_window = first_window[o.t_id()];
    src.quality = quality_at[o.q_id()];
    src.mirrored = (o.strand() != 0) != (stored_reversed[o.q_id()] != 0) ? 1 : 0;
  }
  fprintf(stderr,
          "[racon_amd::layers] %lu layer sources (%.1f MB of read qualities) set up in %.1f ms\n",
          static_cast<unsigned long>(sources.size()), quality_bytes / 1e6,
          1e3 * seconds_since(t_inputs));
  engine_->align_layers(tasks, sources, window_length_, quality_threshold_,
                        first_window[targets_size], gather_qualities, layers);
  logger_->log("[racon::Polisher::initialize] aligned overlaps");
  return true;
}

void Polisher::collect_alignment_tasks(std::vector<std::unique_ptr<Overlap>>& overlaps,
                                       std::vector<AlignTask>* tasks_out,
                                       std::vector<uint64_t>* task_overlap_out) {
  std::vector<AlignTask>& tasks = *tasks_out;
  std::vector<uint64_t>& task_overlap = *task_overlap_out;
  for (uint64_t i = 0; i < overlaps.size(); ++i) {
    overlaps[i]->check_transmuted();
    if (overlaps[i]->needs_alignment()) {
      AlignTask t;
      overlaps[i]->alignment_input(sequences_, &t.query, &t.query_length,
                                   &t.target, &t.target_length);
      overlaps[i]->breaking_point_span(&t.t_begin, &t.t_end, &t.q_first);
      tasks.push_back(t);
      task_overlap.push_back(i);
    }
  }
  timings_.num_overlaps_aligned = tasks.size();
  if (!dump_overlaps_path_.empty()) {
    // "RHA1", u64 n, then per pair: u32 |q|, u32 |t|, q bytes, t bytes
    FILE* f = fopen(dump_overlaps_path_.c_str(), "wb");
    if (f == nullptr) {
      fprintf(stderr, "[racon::Polisher] error: unable to write %s\n",
              dump_overlaps_path_.c_str());
      exit(1);
    }
    const uint64_t n = tasks.size();
    fwrite("RHA1", 1, 4, f);
    fwrite(&n, 8, 1, f);
    for (const auto& t : tasks) {
      fwrite(&t.query_length, 4, 1, f);
      fwrite(&t.target_length, 4, 1, f);
      fwrite(t.query, 1, t.query_length, f);
      fwrite(t.target, 1, t.target_length, f);
    }
    fclose(f);
  }
}

void Polisher::find_overlap_breaking_points(
    std::vector<std::unique_ptr<Overlap>>& overlaps) {
  // polisher.cpp:462-484 with the per-overlap edlib calls batched through
  // the engine; SAM overlaps carry their CIGAR and skip alignment
  // (overlap.cpp:192).
  std::vector<AlignTask> tasks;
  std::vector<uint64_t> task_overlap;
  collect_alignment_tasks(overlaps, &tasks, &task_overlap);

  {
    std::vector<std::future<void>> futures;
    for (uint64_t i = 0; i < overlaps.size(); ++i) {
      if (!overlaps[i]->needs_alignment() && !overlaps[i]->cigar().empty()) {
        futures.emplace_back(thread_pool_->submit(
            [&](uint64_t j) -> void {
              overlaps[j]->find_breaking_points_from_cigar(window_length_);
            },
            i));
      }
    }
    for (const auto& it : futures) {
      it.wait();
    }
  }

  if (!tasks.empty()) {
    const bool done = engine_->align_breaking_points(
        tasks, window_length_, [&](uint64_t k, const uint32_t* points, uint64_t n_points) {
          overlaps[task_overlap[k]]->set_breaking_points(points, n_points);
        });
    if (!done) {
      engine_->align_overlaps(
          tasks, [&](uint64_t k, const uint8_t* ops, uint64_t n_ops) {
            overlaps[task_overlap[k]]->find_breaking_points_from_ops(ops, n_ops,
                                                                     window_length_);
          });
    }
  }
  logger_->log("[racon::Polisher::initialize] aligned overlaps");
}

void Polisher::polish(std::vector<std::unique_ptr<Sequence>>& dst,
                      bool drop_unpolished_sequences) {
  polish_streamed(drop_unpolished_sequences,
                  [&](std::unique_ptr<Sequence> s) { dst.emplace_back(std::move(s)); });
}

void Polisher::polish_streamed(bool drop_unpolished_sequences,
                               const std::function<void(std::unique_ptr<Sequence>)>& sink) {
  // polisher.cpp:486-548, with the stitching done target by target as soon
  // as the engine reports a target's windows final (same sequences, same
  // order, same tags)
  logger_->log();
  auto t_consensus = std::chrono::steady_clock::now();
  double stitch_s = 0;

  std::vector<uint8_t> status;
  std::string polished_data;
  uint32_t num_polished_windows = 0;
  uint64_t next = 0;  // first window not yet stitched
  auto stitch = [&](uint64_t ready) {
    auto t0 = std::chrono::steady_clock::now();
    for (; next < ready; ++next) {
      const uint64_t i = next;
      if (status[i] == kWindowPolishedChimeric) {
        windows_[i]->warn_chimeric();
      }
      num_polished_windows += status[i] != kWindowUnpolished ? 1 : 0;
      polished_data += windows_[i]->consensus();

      if (i == windows_.size() - 1 || windows_[i + 1]->rank() == 0) {
        double polished_ratio =
            num_polished_windows / static_cast<double>(windows_[i]->rank() + 1);

        if (!drop_unpolished_sequences || polished_ratio > 0) {
          std::string tags = type_ == PolisherType::kF ? "r" : "";
          tags += " LN:i:" + std::to_string(polished_data.size());
          tags += " RC:i:" + std::to_string(targets_coverages_[windows_[i]->id()]);
          tags += " XC:f:" + std::to_string(polished_ratio);
          sink(createSequence(sequences_[windows_[i]->id()]->name() + tags, polished_data));
        }
        num_polished_windows = 0;
        polished_data.clear();
      }
      windows_[i].reset();
    }
    stitch_s += seconds_since(t0);
  };
  engine_->generate_consensus(windows_, trim_, status, stitch);
  stitch(windows_.size());
  timings_.stitch = stitch_s;
  timings_.consensus = seconds_since(t_consensus) - stitch_s;

  logger_->log("[racon::Polisher::polish] generated consensus");

  std::vector<std::shared_ptr<Window>>().swap(windows_);
  std::vector<std::unique_ptr<Sequence>>().swap(sequences_);
}

}  // namespace racon
