// Seeded synthetic polishing inputs: contigs, reads (FASTQ/FASTA) and PAF
// overlaps from true positions. Identical output on any machine (splitmix64,
// no library RNG).
//
//   simgen contig <seed> <out_prefix> <contig_bp,...> <coverage> <mean> <sd>
//          <error> <sub:ins:del> [--depth-gamma] [--min-len N] [--max-len N]
//   simgen frag   <seed> <out_prefix> <genome_bp> <reads> <mean> <sd> <error>
//
// contig mode writes <prefix>.contigs.fasta (uniform ACGT draft = truth with
// 1% errors), <prefix>.reads.fastq and <prefix>.paf (read -> contig). frag
// mode samples reads from one random genome and writes all-vs-all dual PAF
// overlaps between reads whose true spans overlap (<prefix>.reads.fastq,
// <prefix>.ava.paf).
//
// Read lengths are lognormal with mean m and sd s (mu = ln(m^2/sqrt(m^2+s^2)),
// sigma = sqrt(ln(1 + s^2/m^2))); errors are applied per base at the total
// rate with the given sub:ins:del split, indels x2 inside homopolymer runs of
// length >= 3; strand is 50/50; Phred qualities are uniform 8..30.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  uint32_t below(uint32_t n) { return static_cast<uint32_t>(uniform() * n); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  double gamma2() {  // Gamma(shape 2, scale 1) = sum of two exponentials
    double a = uniform(), b = uniform();
    if (a < 1e-300) a = 1e-300;
    if (b < 1e-300) b = 1e-300;
    return -std::log(a) - std::log(b);
  }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

std::string random_genome(Rng& r, uint64_t n) {
  std::string g(n, 'A');
  for (uint64_t i = 0; i < n; ++i) g[i] = kBases[r.next() & 3];
  return g;
}

char complement(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return c;
  }
}

std::string revcomp(const std::string& s) {
  std::string o(s.rbegin(), s.rend());
  for (auto& c : o) c = complement(c);
  return o;
}

struct ErrorModel {
  double rate = 0.12;
  double sub = 0.4, ins = 0.2, del = 0.4;
};

std::string mutate(Rng& r, const std::string& src, const ErrorModel& m) {
  std::string out;
  out.reserve(src.size() + src.size() / 8);
  size_t run = 0;
  for (size_t i = 0; i < src.size(); ++i) {
    run = (i > 0 && src[i] == src[i - 1]) ? run + 1 : 1;
    const double k = run >= 3 ? 2.0 : 1.0;  // homopolymer indel bias
    const double p_sub = m.rate * m.sub, p_ins = m.rate * m.ins * k,
                 p_del = m.rate * m.del * k;
    const double u = r.uniform();
    const char c = src[i];
    if (u < p_sub) {
      char b = kBases[r.below(3)];
      if (b == c) b = 'T';
      if (b == c) b = 'G';
      out += b;
    } else if (u < p_sub + p_ins) {
      out += c;
      out += kBases[r.next() & 3];
    } else if (u < p_sub + p_ins + p_del) {
      // deleted
    } else {
      out += c;
    }
  }
  if (out.empty()) out += 'A';
  return out;
}

uint32_t lognormal_length(Rng& r, double mean, double sd, uint32_t lo, uint32_t hi) {
  const double sigma = std::sqrt(std::log(1.0 + sd * sd / (mean * mean)));
  const double mu = std::log(mean * mean / std::sqrt(mean * mean + sd * sd));
  double x = std::exp(mu + sigma * r.normal());
  x = std::min<double>(std::max<double>(x, lo), hi);
  return static_cast<uint32_t>(x);
}

ErrorModel parse_split(double rate, const char* s) {
  ErrorModel m;
  m.rate = rate;
  double a = 0, b = 0, c = 0;
  if (sscanf(s, "%lf:%lf:%lf", &a, &b, &c) == 3 && a + b + c > 0) {
    const double t = a + b + c;
    m.sub = a / t;
    m.ins = b / t;
    m.del = c / t;
  }
  return m;
}

void write_fastq(FILE* f, const std::string& name, const std::string& seq, Rng& r) {
  std::string q(seq.size(), '!');
  for (auto& c : q) c = static_cast<char>(33 + 8 + r.below(23));
  fprintf(f, "@%s\n%s\n+\n%s\n", name.c_str(), seq.c_str(), q.c_str());
}

struct Read {
  std::string name;
  uint32_t contig;
  uint64_t begin, end;  // true span on the genome/contig
  bool reverse;
  uint32_t length;
};

int contig_mode(int argc, char** argv) {
  if (argc < 10) return 2;
  const uint64_t seed = strtoull(argv[2], nullptr, 10);
  const std::string prefix = argv[3];
  std::vector<uint64_t> sizes;
  for (char* tok = strtok(argv[4], ","); tok != nullptr; tok = strtok(nullptr, ",")) {
    sizes.push_back(strtoull(tok, nullptr, 10));
  }
  const double coverage = atof(argv[5]);
  const double mean = atof(argv[6]), sd = atof(argv[7]);
  const ErrorModel em = parse_split(atof(argv[8]), argv[9]);
  bool depth_gamma = false;
  uint32_t min_len = 500, max_len = 200000;
  ErrorModel em2 = em;
  double second_fraction = 0.0;
  for (int i = 10; i < argc; ++i) {
    if (!strcmp(argv[i], "--depth-gamma")) depth_gamma = true;
    if (!strcmp(argv[i], "--min-len") && i + 1 < argc) min_len = atoi(argv[++i]);
    if (!strcmp(argv[i], "--max-len") && i + 1 < argc) max_len = atoi(argv[++i]);
    if (!strcmp(argv[i], "--second") && i + 3 < argc) {
      // --second <fraction> <rate> <sub:ins:del>: a second read population
      second_fraction = atof(argv[++i]);
      const double rate = atof(argv[++i]);
      em2 = parse_split(rate, argv[++i]);
    }
  }
  Rng rng(seed);
  std::vector<std::string> truth, draft;
  FILE* fc = fopen((prefix + ".contigs.fasta").c_str(), "w");
  for (size_t k = 0; k < sizes.size(); ++k) {
    truth.push_back(random_genome(rng, sizes[k]));
    ErrorModel dm;
    dm.rate = 0.01;
    draft.push_back(mutate(rng, truth.back(), dm));
    fprintf(fc, ">contig%zu\n%s\n", k, draft.back().c_str());
  }
  fclose(fc);

  FILE* fq = fopen((prefix + ".reads.fastq").c_str(), "w");
  FILE* fp = fopen((prefix + ".paf").c_str(), "w");
  uint64_t n_reads = 0;
  for (size_t k = 0; k < sizes.size(); ++k) {
    const std::string& g = truth[k];
    const uint64_t G = g.size();
    // per-region depth multipliers (10 kbp regions): Gamma(2) scaled to the
    // target coverage and clipped to 5..200x
    std::vector<double> region_depth((G + 9999) / 10000 + 1, coverage);
    if (depth_gamma) {
      for (auto& d : region_depth) {
        d = std::min(200.0, std::max(5.0, coverage * rng.gamma2() / 2.0));
      }
    }
    double expected = 0;
    for (uint64_t r = 0; r * 10000 < G; ++r) {
      expected += region_depth[r] * std::min<uint64_t>(10000, G - r * 10000);
    }
    const uint64_t reads = static_cast<uint64_t>(expected / mean) + 1;
    for (uint64_t i = 0; i < reads; ++i) {
      // start drawn in proportion to the region depth
      uint64_t start;
      while (true) {
        start = static_cast<uint64_t>(rng.uniform() * G);
        const double d = region_depth[start / 10000];
        if (!depth_gamma || rng.uniform() * 200.0 <= d) break;
      }
      const uint32_t len = lognormal_length(rng, mean, sd, min_len, max_len);
      const uint64_t end = std::min<uint64_t>(G, start + len);
      if (end - start < 100) continue;
      const ErrorModel& m = rng.uniform() < second_fraction ? em2 : em;
      std::string seq = mutate(rng, g.substr(start, end - start), m);
      const bool rev = (rng.next() & 1) != 0;
      if (rev) seq = revcomp(seq);
      const std::string name = "r" + std::to_string(n_reads++);
      write_fastq(fq, name, seq, rng);
      // PAF against the draft: true span mapped through the 1% draft errors
      // approximately by scaling
      const double scale = static_cast<double>(draft[k].size()) / G;
      uint64_t tb = static_cast<uint64_t>(start * scale), te = static_cast<uint64_t>(end * scale);
      te = std::min<uint64_t>(te, draft[k].size());
      if (te <= tb) continue;
      const uint64_t span = std::max<uint64_t>(seq.size(), te - tb);
      fprintf(fp, "%s\t%zu\t0\t%zu\t%c\tcontig%zu\t%zu\t%lu\t%lu\t%lu\t%lu\t60\n", name.c_str(),
              seq.size(), seq.size(), rev ? '-' : '+', k, draft[k].size(),
              static_cast<unsigned long>(tb), static_cast<unsigned long>(te),
              static_cast<unsigned long>(std::min<uint64_t>(seq.size(), te - tb)),
              static_cast<unsigned long>(span));
    }
  }
  fclose(fq);
  fclose(fp);
  fprintf(stderr, "simgen: %lu reads\n", static_cast<unsigned long>(n_reads));
  return 0;
}

int frag_mode(int argc, char** argv) {
  if (argc < 9) return 2;
  const uint64_t seed = strtoull(argv[2], nullptr, 10);
  const std::string prefix = argv[3];
  const uint64_t G = strtoull(argv[4], nullptr, 10);
  const uint64_t n = strtoull(argv[5], nullptr, 10);
  const double mean = atof(argv[6]), sd = atof(argv[7]);
  ErrorModel em;
  em.rate = atof(argv[8]);
  uint32_t max_ovl = 40;
  for (int i = 9; i < argc; ++i) {
    if (!strcmp(argv[i], "--max-overlaps") && i + 1 < argc) max_ovl = atoi(argv[++i]);
  }
  Rng rng(seed);
  const std::string g = random_genome(rng, G);
  std::vector<Read> reads;
  FILE* fq = fopen((prefix + ".reads.fastq").c_str(), "w");
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t len = lognormal_length(rng, mean, sd, 1000, 100000);
    const uint64_t start = static_cast<uint64_t>(rng.uniform() * (G - 1000));
    const uint64_t end = std::min<uint64_t>(G, start + len);
    std::string seq = mutate(rng, g.substr(start, end - start), em);
    const bool rev = (rng.next() & 1) != 0;
    if (rev) seq = revcomp(seq);
    Read r{"f" + std::to_string(i), 0, start, end, rev, static_cast<uint32_t>(seq.size())};
    write_fastq(fq, r.name, seq, rng);
    reads.push_back(r);
  }
  fclose(fq);
  // all-vs-all dual overlaps from true spans, at most max_ovl per read
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(),
            [&](uint32_t a, uint32_t b) { return reads[a].begin < reads[b].begin; });
  FILE* fp = fopen((prefix + ".ava.paf").c_str(), "w");
  std::vector<uint32_t> count(n, 0);
  for (uint64_t ii = 0; ii < n; ++ii) {
    const Read& a = reads[order[ii]];
    for (uint64_t jj = ii + 1; jj < n; ++jj) {
      const Read& b = reads[order[jj]];
      if (b.begin >= a.end) break;
      const uint64_t ob = std::max(a.begin, b.begin), oe = std::min(a.end, b.end);
      if (oe <= ob + 500) continue;
      if (count[order[ii]] >= max_ovl || count[order[jj]] >= max_ovl) continue;
      ++count[order[ii]];
      ++count[order[jj]];
      // read coordinates of the shared span (scaled by the read/span ratio)
      auto span_of = [&](const Read& r, uint64_t& qb, uint64_t& qe) {
        const double s = static_cast<double>(r.length) / (r.end - r.begin);
        uint64_t fb = static_cast<uint64_t>((ob - r.begin) * s);
        uint64_t fe = std::min<uint64_t>(r.length, static_cast<uint64_t>((oe - r.begin) * s));
        if (r.reverse) {
          qb = r.length - fe;
          qe = r.length - fb;
        } else {
          qb = fb;
          qe = fe;
        }
      };
      uint64_t aqb, aqe, bqb, bqe;
      span_of(a, aqb, aqe);
      span_of(b, bqb, bqe);
      const char strand = a.reverse != b.reverse ? '-' : '+';
      const uint64_t alen = std::max(aqe - aqb, bqe - bqb);
      fprintf(fp, "%s\t%u\t%lu\t%lu\t%c\t%s\t%u\t%lu\t%lu\t%lu\t%lu\t255\n", a.name.c_str(),
              a.length, (unsigned long)aqb, (unsigned long)aqe, strand, b.name.c_str(), b.length,
              (unsigned long)bqb, (unsigned long)bqe, (unsigned long)std::min(aqe - aqb, bqe - bqb),
              (unsigned long)alen);
      fprintf(fp, "%s\t%u\t%lu\t%lu\t%c\t%s\t%u\t%lu\t%lu\t%lu\t%lu\t255\n", b.name.c_str(),
              b.length, (unsigned long)bqb, (unsigned long)bqe, strand, a.name.c_str(), a.length,
              (unsigned long)aqb, (unsigned long)aqe, (unsigned long)std::min(aqe - aqb, bqe - bqb),
              (unsigned long)alen);
    }
  }
  fclose(fp);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  int rc = 2;
  if (argc > 1 && !strcmp(argv[1], "contig")) rc = contig_mode(argc, argv);
  if (argc > 1 && !strcmp(argv[1], "frag")) rc = frag_mode(argc, argv);
  if (rc == 2) {
    fprintf(stderr,
            "usage: simgen contig <seed> <prefix> <bp,bp,...> <coverage> <mean> <sd> <error> "
            "<sub:ins:del> [--depth-gamma] [--min-len N] [--max-len N] "
            "[--second <fraction> <error> <sub:ins:del>]\n"
            "       simgen frag <seed> <prefix> <genome_bp> <reads> <mean> <sd> <error> "
            "[--max-overlaps N]\n");
  }
  return rc;
}
