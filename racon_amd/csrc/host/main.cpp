// racon command line (racon src/main.cpp:1-234): same options, defaults,
// validation and output format. The device switches -c/--cudapoa-batches,
// -b/--cuda-banded-alignment, --cudaaligner-batches and
// --cudaaligner-band-width are always compiled in.
//
// Extra long options for the benchmark driver (they do not change the
// output): --report <path> writes the phase timings as JSON;
// --dump-overlaps <path> writes the overlap alignment inputs;
// --dump-windows <path> writes the windows after initialization and exits.
//
// --serve <job list> polishes many jobs in one process on one device (each
// job's output file holds what racon prints for it alone): see serve.hpp;
// --serve-budget <GiB>, --serve-executors <n> and --serve-jobs <n> size the
// device pool, the kernel streams and the jobs prepared at once.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "polisher.hpp"
#include "sequence.hpp"
#include "serve.hpp"

#ifndef RACON_VERSION
#define RACON_VERSION "v1.4.17"
#endif

static const char* version = RACON_VERSION;
static const int32_t CUDAALIGNER_INPUT_CODE = 10000;
static const int32_t CUDAALIGNER_BAND_WIDTH_INPUT_CODE = 10001;
static const int32_t REPORT_INPUT_CODE = 10002;
static const int32_t DUMP_WINDOWS_CODE = 10003;
static const int32_t DUMP_OVERLAPS_CODE = 10004;
static const int32_t SERVE_CODE = 10005;
static const int32_t SERVE_BUDGET_CODE = 10006;
static const int32_t SERVE_EXECUTORS_CODE = 10007;
static const int32_t SERVE_JOBS_CODE = 10008;

static struct option options[] = {
    {"include-unpolished", no_argument, 0, 'u'},
    {"fragment-correction", no_argument, 0, 'f'},
    {"window-length", required_argument, 0, 'w'},
    {"quality-threshold", required_argument, 0, 'q'},
    {"error-threshold", required_argument, 0, 'e'},
    {"no-trimming", no_argument, 0, 'T'},
    {"match", required_argument, 0, 'm'},
    {"mismatch", required_argument, 0, 'x'},
    {"gap", required_argument, 0, 'g'},
    {"threads", required_argument, 0, 't'},
    {"version", no_argument, 0, 'v'},
    {"help", no_argument, 0, 'h'},
    {"cudapoa-batches", optional_argument, 0, 'c'},
    {"cuda-banded-alignment", no_argument, 0, 'b'},
    {"cudaaligner-batches", required_argument, 0, CUDAALIGNER_INPUT_CODE},
    {"cudaaligner-band-width", required_argument, 0, CUDAALIGNER_BAND_WIDTH_INPUT_CODE},
    {"report", required_argument, 0, REPORT_INPUT_CODE},
    {"dump-windows", required_argument, 0, DUMP_WINDOWS_CODE},
    {"dump-overlaps", required_argument, 0, DUMP_OVERLAPS_CODE},
    {"serve", required_argument, 0, SERVE_CODE},
    {"serve-budget", required_argument, 0, SERVE_BUDGET_CODE},
    {"serve-executors", required_argument, 0, SERVE_EXECUTORS_CODE},
    {"serve-jobs", required_argument, 0, SERVE_JOBS_CODE},
    {0, 0, 0, 0}};

static void help();

int main(int argc, char** argv) {
  std::vector<std::string> input_paths;

  uint32_t window_length = 500;
  double quality_threshold = 10.0;
  double error_threshold = 0.3;
  bool trim = true;

  int8_t match = 3;
  int8_t mismatch = -5;
  int8_t gap = -4;
  uint32_t type = 0;

  bool drop_unpolished_sequences = true;
  uint32_t num_threads = 1;

  uint32_t cudapoa_batches = 0;
  uint32_t cudaaligner_batches = 0;
  uint32_t cudaaligner_band_width = 0;
  bool cuda_banded_alignment = false;
  std::string report_path;
  std::string dump_windows_path, dump_overlaps_path;
  std::string serve_path;
  double serve_budget_gib = 8.0;
  uint32_t serve_executors = 4, serve_jobs = 4;

  std::string optstring = "ufw:q:e:m:x:g:t:hbc::";

  int32_t argument;
  while ((argument = getopt_long(argc, argv, optstring.c_str(), options, nullptr)) != -1) {
    switch (argument) {
      case 'u': drop_unpolished_sequences = false; break;
      case 'f': type = 1; break;
      case 'w': window_length = atoi(optarg); break;
      case 'q': quality_threshold = atof(optarg); break;
      case 'e': error_threshold = atof(optarg); break;
      case 'T': trim = false; break;
      case 'm': match = atoi(optarg); break;
      case 'x': mismatch = atoi(optarg); break;
      case 'g': gap = atoi(optarg); break;
      case 't': num_threads = atoi(optarg); break;
      case 'v': printf("%s\n", version); exit(0);
      case 'h': help(); exit(0);
      case 'c':
        // main.cpp:114-126: optional count, also accepted as the next word
        cudapoa_batches = 1;
        if (optarg == NULL && argv[optind] != NULL && argv[optind][0] != '-') {
          cudapoa_batches = atoi(argv[optind++]);
        }
        if (optarg != NULL) {
          cudapoa_batches = atoi(optarg);
        }
        break;
      case 'b': cuda_banded_alignment = true; break;
      case CUDAALIGNER_INPUT_CODE: cudaaligner_batches = atoi(optarg); break;
      case CUDAALIGNER_BAND_WIDTH_INPUT_CODE: cudaaligner_band_width = atoi(optarg); break;
      case REPORT_INPUT_CODE: report_path = optarg; break;
      case DUMP_WINDOWS_CODE: dump_windows_path = optarg; break;
      case DUMP_OVERLAPS_CODE: dump_overlaps_path = optarg; break;
      case SERVE_CODE: serve_path = optarg; break;
      case SERVE_BUDGET_CODE: serve_budget_gib = atof(optarg); break;
      case SERVE_EXECUTORS_CODE: serve_executors = atoi(optarg); break;
      case SERVE_JOBS_CODE: serve_jobs = atoi(optarg); break;
      default: exit(1);
    }
  }

  for (int32_t i = optind; i < argc; ++i) {
    input_paths.emplace_back(argv[i]);
  }

  if (!serve_path.empty()) {
    racon::ServeOptions so;
    so.type = type == 0 ? racon::PolisherType::kC : racon::PolisherType::kF;
    so.window_length = window_length;
    so.quality_threshold = quality_threshold;
    so.error_threshold = error_threshold;
    so.trim = trim;
    so.match = match;
    so.mismatch = mismatch;
    so.gap = gap;
    so.num_threads = num_threads;
    so.drop_unpolished = drop_unpolished_sequences;
    so.aligner_band_width = cudaaligner_band_width;
    so.budget_bytes = static_cast<uint64_t>(serve_budget_gib * (1ull << 30));
    so.executors = serve_executors;
    so.concurrent_jobs = serve_jobs;
    return racon::serve_jobs(racon::read_jobs(serve_path), so);
  }

  if (input_paths.size() < 3) {
    fprintf(stderr, "[racon::] error: missing input file(s)!\n");
    help();
    exit(1);
  }

  racon::PolisherTimings timings;
  {
    auto polisher = racon::createPolisher(
        input_paths[0], input_paths[1], input_paths[2],
        type == 0 ? racon::PolisherType::kC : racon::PolisherType::kF,
        window_length, quality_threshold, error_threshold, trim, match, mismatch,
        gap, num_threads, cudapoa_batches, cuda_banded_alignment,
        cudaaligner_batches, cudaaligner_band_width);

    polisher->set_dump_paths(dump_windows_path, dump_overlaps_path);
    polisher->initialize();
    if (!dump_windows_path.empty()) {
      // benchmark input only: windows written, no polishing
      polisher->dump_windows();
      return 0;
    }
    // each target is written as soon as its windows are polished (the bytes
    // of racon's output loop, main.cpp:164-166, in the same order)
    polisher->polish_streamed(drop_unpolished_sequences,
                              [](std::unique_ptr<racon::Sequence> s) {
                                fprintf(stdout, ">%s\n%s\n", s->name().c_str(),
                                        s->data().c_str());
                              });
    timings = polisher->timings();
  }
  fflush(stdout);

  if (!report_path.empty()) {
    FILE* f = fopen(report_path.c_str(), "w");
    if (f != nullptr) {
      fprintf(f,
              "{\"load_s\": %.6f, \"align_s\": %.6f, \"windows_s\": %.6f, "
              "\"consensus_s\": %.6f, \"stitch_s\": %.6f, \"num_windows\": %lu, "
              "\"num_overlaps_aligned\": %lu, \"threads\": %u}\n",
              timings.load, timings.align, timings.windows, timings.consensus,
              timings.stitch, static_cast<unsigned long>(timings.num_windows),
              static_cast<unsigned long>(timings.num_overlaps_aligned), num_threads);
      fclose(f);
    }
  }
  return 0;
}

static void help() {
  printf(
      "usage: racon [options ...] <sequences> <overlaps> <target sequences>\n"
      "\n"
      "    #default output is stdout\n"
      "    <sequences>\n"
      "        input file in FASTA/FASTQ format (can be compressed with gzip)\n"
      "        containing sequences used for correction\n"
      "    <overlaps>\n"
      "        input file in MHAP/PAF/SAM format (can be compressed with gzip)\n"
      "        containing overlaps between sequences and target sequences\n"
      "    <target sequences>\n"
      "        input file in FASTA/FASTQ format (can be compressed with gzip)\n"
      "        containing sequences which will be corrected\n"
      "\n"
      "    options:\n"
      "        -u, --include-unpolished\n"
      "            output unpolished target sequences\n"
      "        -f, --fragment-correction\n"
      "            perform fragment correction instead of contig polishing\n"
      "            (overlaps file should contain dual/self overlaps!)\n"
      "        -w, --window-length <int>\n"
      "            default: 500\n"
      "            size of window on which POA is performed\n"
      "        -q, --quality-threshold <float>\n"
      "            default: 10.0\n"
      "            threshold for average base quality of windows used in POA\n"
      "        -e, --error-threshold <float>\n"
      "            default: 0.3\n"
      "            maximum allowed error rate used for filtering overlaps\n"
      "        --no-trimming\n"
      "            disables consensus trimming at window ends\n"
      "        -m, --match <int>\n"
      "            default: 3\n"
      "            score for matching bases\n"
      "        -x, --mismatch <int>\n"
      "            default: -5\n"
      "            score for mismatching bases\n"
      "        -g, --gap <int>\n"
      "            default: -4\n"
      "            gap penalty (must be negative)\n"
      "        -t, --threads <int>\n"
      "            default: 1\n"
      "            number of threads\n"
      "        --version\n"
      "            prints the version number\n"
      "        -h, --help\n"
      "            prints the usage\n"
      "        -c, --cudapoa-batches <int>\n"
      "            default: 0\n"
      "            number of batches for GPU accelerated polishing\n"
      "        -b, --cuda-banded-alignment\n"
      "            use banding approximation for alignment on GPU\n"
      "        --cudaaligner-batches <int>\n"
      "            default: 0\n"
      "            number of batches for GPU accelerated alignment\n"
      "        --cudaaligner-band-width <int>\n"
      "            default: 0\n"
      "            Band width for GPU alignment. Must be >= 0. Non-zero allows user defined \n"
      "            band width, whereas 0 implies auto band width determination.\n");
}
