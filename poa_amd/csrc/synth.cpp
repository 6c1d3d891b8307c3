// Seeded synthetic POA windows: a uniform-random ACGT template and reads derived from it by
// i.i.d. per-base substitutions, insertions and deletions. Deterministic across machines
// (splitmix64 stream; the only libm calls are log/exp/sqrt/cos for the lognormal length).
#include "mpoa_synth.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  int below(int n) { return (int)(next() % (uint64_t)n); }
  double normal() {  // Box-Muller, one value per call
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int geometric(double p) {  // P(L = 1) = p
    if (p <= 0.0 || p >= 1.0) return 1;
    int L = 1;
    while (uniform() >= p && L < 1000) L++;
    return L;
  }
};

const char kBase[4] = {'A', 'C', 'G', 'T'};

uint64_t mix(uint64_t seed, uint64_t idx) {
  Rng r(seed ^ (idx * 0xd1b54a32d192ed03ULL + 0x8cb92ba72f3d8dd7ULL));
  r.next();
  return r.next();
}

// One read: every template base meets exactly one event. Substitution draws one of the other
// three bases; an insertion emits L random bases and then the template base; a deletion skips
// L template bases (L = 1 when geo_p is 0).
void derive(Rng& r, const std::string& tpl, double sub, double ins, double del, double geo_p,
            std::string& out) {
  out.clear();
  size_t t = 0;
  while (t < tpl.size()) {
    double u = r.uniform();
    if (u < sub) {
      int b = r.below(3);
      int tb = tpl[t] == 'A' ? 0 : tpl[t] == 'C' ? 1 : tpl[t] == 'G' ? 2 : 3;
      out.push_back(kBase[b >= tb ? b + 1 : b]);
      t++;
    } else if (u < sub + ins) {
      int L = geo_p > 0 ? r.geometric(geo_p) : 1;
      for (int k = 0; k < L; k++) out.push_back(kBase[r.below(4)]);
      out.push_back(tpl[t]);
      t++;
    } else if (u < sub + ins + del) {
      int L = geo_p > 0 ? r.geometric(geo_p) : 1;
      t += (size_t)L;
    } else {
      out.push_back(tpl[t]);
      t++;
    }
  }
  if (out.empty()) out.push_back(tpl.empty() ? 'A' : tpl[0]);
}

void make_window(uint64_t seed, int tlen, int n_reads, double sub, double ins, double del,
                 double geo_p, std::vector<std::string>& reads) {
  Rng r(seed);
  std::string tpl(tlen, 'A');
  for (int i = 0; i < tlen; i++) tpl[i] = kBase[r.below(4)];
  reads.resize(n_reads);
  for (int i = 0; i < n_reads; i++) derive(r, tpl, sub, ins, del, geo_p, reads[i]);
}

struct BatchWin {
  int tlen, n_reads;
  double err;
};

BatchWin batch_window(uint64_t seed, int w, const mpoa_synth_batch_spec* s) {
  Rng r(mix(seed, (uint64_t)w));
  BatchWin b;
  b.n_reads = s->min_reads + r.below(s->max_reads - s->min_reads + 1);
  double len = std::exp(std::log(s->median_len) + s->sigma * r.normal());
  long L = std::lround(len);
  if (L < s->min_len) L = s->min_len;
  if (L > s->max_len) L = s->max_len;
  b.tlen = (int)L;
  b.err = s->err_lo + (s->err_hi - s->err_lo) * r.uniform();
  return b;
}

}  // namespace

extern "C" int mpoa_synth_window(uint64_t seed, int32_t tlen, int32_t n_reads, double sub,
                                 double ins, double del, double geo_p, char* seq, int64_t cap,
                                 int64_t* off) {
  if (tlen <= 0 || n_reads < 0 || !off) return -1;
  std::vector<std::string> reads;
  make_window(seed, tlen, n_reads, sub, ins, del, geo_p, reads);
  int64_t at = 0;
  off[0] = 0;
  for (int i = 0; i < n_reads; i++) {
    if (seq && at + (int64_t)reads[i].size() <= cap)
      std::memcpy(seq + at, reads[i].data(), reads[i].size());
    at += (int64_t)reads[i].size();
    off[i + 1] = at;
  }
  return seq && at > cap ? -4 : 0;
}

extern "C" int mpoa_synth_batch(uint64_t seed, const mpoa_synth_batch_spec* spec, int32_t n_windows,
                                int32_t* win_off, char* seq, int64_t cap, int64_t* seq_off,
                                int64_t* total_reads, int64_t* total_bases) {
  if (!spec || n_windows < 0 || spec->min_reads < 1 || spec->max_reads < spec->min_reads)
    return -1;
  int64_t reads = 0, bases = 0;
  if (win_off) win_off[0] = 0;
  if (seq_off) seq_off[0] = 0;
  std::vector<std::string> rv;
  for (int w = 0; w < n_windows; w++) {
    BatchWin b = batch_window(seed, w, spec);
    double frac = spec->split_sum > 0 ? 1.0 / spec->split_sum : 0.0;
    double sub = b.err * spec->split_sub * frac, ins = b.err * spec->split_ins * frac,
           del = b.err * spec->split_del * frac;
    make_window(mix(seed ^ 0x5eedULL, (uint64_t)w), b.tlen, b.n_reads, sub, ins, del,
                spec->geo_p, rv);
    for (const std::string& s : rv) {
      if (seq && seq_off) {
        if (bases + (int64_t)s.size() > cap) return -4;
        std::memcpy(seq + bases, s.data(), s.size());
      }
      bases += (int64_t)s.size();
      reads++;
      if (seq_off) seq_off[reads] = bases;
    }
    if (win_off) win_off[w + 1] = (int32_t)reads;
  }
  if (total_reads) *total_reads = reads;
  if (total_bases) *total_bases = bases;
  return 0;
}
