// Host check of the per-position seeding pieces the device runs (poa_amd/csrc/seed_core.h)
// against the oracle's sequential restatement of minipoa (oracle/oracle.cpp, included whole
// so that its file-local functions are reachable):
//   * sketch_step over every position, in order, emits exactly mm_sketch's minimizers
//     (src/minimizer.cpp:60-135) on random sequences with ambiguous bases, tandem repeats and
//     copied stretches (identical k-mers inside one window), for many (w, k);
//   * chain_log2 equals mg_log2 bit for bit, and chain_link equals comput_sc, on random
//     arguments and on every small gap;
//   * the radix sort's digit pass and insertion sort, composed as the device composes them,
//     order equal keys as the oracle's radix sort does (include/ksort.h:101-151).
// Prints "ok" or the first mismatch; exit status 1 on a mismatch.
#include "../oracle/oracle.cpp"

#include <cstdio>
#include <random>

#include "seed_core.h"

using mpoa::seedc::NONE;

static int check_sketch(std::mt19937_64& rng, int trial) {
  const int w = 1 + (int)(rng() % 24), k = 1 + (int)(rng() % 28);
  const int len = (int)(rng() % 700);
  std::string s;
  const char* acgt = "ACGT";
  const int alpha = 1 + (int)(rng() % 4);  // small alphabets give identical k-mers
  while ((int)s.size() < len) {
    const int mode = (int)(rng() % 10);
    if (mode == 0 && !s.empty()) {  // copy an earlier stretch
      const int from = (int)(rng() % s.size());
      const int n = 1 + (int)(rng() % 40);
      for (int t = 0; t < n && (int)s.size() < len; t++) s.push_back(s[from + t % (s.size() - from)]);
    } else if (mode == 1) {
      s.push_back(rng() % 2 ? 'N' : 'x');
    } else {
      s.push_back(acgt[rng() % alpha]);
    }
  }
  std::vector<mm128> want;
  const uint32_t rid = (uint32_t)(rng() % 1000);
  if (len > 0) mm_sketch(s.data(), len, w, k, rid, want);
  std::vector<uint64_t> xs(len);
  auto code = [&](int t) { return (int)code_of((unsigned char)s[t]); };
  for (int i = 0; i < len; i++) xs[i] = mpoa::seedc::kmer_x(code, i, k);
  auto xf = [&](int j) { return xs[j]; };
  std::vector<mm128> got;
  for (int i = 0; i < len; i++) {
    const int before = (int)got.size();
    const int n = mpoa::seedc::sketch_step(i, len, w, xf, [&](int j) {
      got.push_back({xs[j], (uint64_t)rid << 32 | (uint64_t)(uint32_t)j << 1});
    });
    if (n != (int)got.size() - before) {
      std::printf("FAIL trial %d: count %d at position %d\n", trial, n, i);
      return 1;
    }
  }
  if (got.size() != want.size()) {
    std::printf("FAIL trial %d (len %d w %d k %d): %zu minimizers, oracle %zu\n", trial, len, w,
                k, got.size(), want.size());
    return 1;
  }
  for (size_t t = 0; t < got.size(); t++) {
    if (got[t].x != want[t].x || got[t].y != want[t].y) {
      std::printf("FAIL trial %d (len %d w %d k %d): minimizer %zu differs\n", trial, len, w, k,
                  t);
      return 1;
    }
  }
  return 0;
}

static int check_chain(std::mt19937_64& rng) {
  for (int d = 1; d < 200000; d++) {
    const float v = (float)(d + 1);
    const float a = mpoa::seedc::chain_log2(v), b = mg_log2(v);
    if (__builtin_bit_cast(uint32_t, a) != __builtin_bit_cast(uint32_t, b)) {
      std::printf("FAIL log2(%d)\n", d + 1);
      return 1;
    }
  }
  for (int t = 0; t < 2000000; t++) {
    const int k = 3 + (int)(rng() % 26);
    const int max_dist = 50 + (int)(rng() % 2000);
    const float gap = (float)(0.01 * k), skip = (float)(0.0 * 0.01 * k);
    const int tpj = (int)(rng() % 100000), qpj = (int)(rng() % 100000);
    const int tpi = tpj + (int)(rng() % (rng() % 2 ? 40 : 3000));
    const int qpi = qpj + (int)(rng() % (rng() % 2 ? 40 : 3000)) - 20;
    mm128 ai = {(uint64_t)(uint32_t)tpi, (uint64_t)k << 32 | (uint32_t)qpi};
    mm128 aj = {(uint64_t)(uint32_t)tpj, (uint64_t)k << 32 | (uint32_t)qpj};
    const int want = comput_sc(&ai, &aj, max_dist, gap, skip);
    const int got = mpoa::seedc::chain_link(tpi, qpi, tpj, qpj, k, max_dist, gap, skip);
    if (want != got) {
      std::printf("FAIL link (%d,%d)-(%d,%d) k %d: %d, oracle %d\n", tpj, qpj, tpi, qpi, k, got,
                  want);
      return 1;
    }
  }
  return 0;
}

// The digit passes and insertion sorts of the device's radix-order emulation, composed as its
// stack composes them (pending buckets in any order), against the oracle's radix sort on keys
// with many ties: the order of equal keys (their payload) must be the same.
static int check_radix(std::mt19937_64& rng) {
  for (int t = 0; t < 3000; t++) {
    const int n = 65 + (int)(rng() % 3000);
    const int shape = (int)(rng() % 4);
    std::vector<mm128> a(n);
    const uint64_t distinct = 1 + rng() % (shape == 0 ? 8 : shape == 1 ? 300 : 100000);
    for (int i = 0; i < n; i++) {
      uint64_t key = rng() % distinct;
      if (shape == 2) key = key << 8 | 15;                       // minimizer-like x
      if (shape == 3) key = (uint64_t)7 << 32 | (key & 0xffff);  // anchor-like x
      a[i] = {key, (uint64_t)i};
    }
    std::vector<uint64_t> kx(n);
    std::vector<uint32_t> ki(n);
    for (int i = 0; i < n; i++) kx[i] = a[i].x, ki[i] = (uint32_t)i;
    radix_sort(a.data(), a.data() + n);
    std::vector<int> stk{0, n, 56};
    int cb[256], ce[256];
    while (!stk.empty()) {
      const int s = stk.back(), e = stk[stk.size() - 2], b = stk[stk.size() - 3];
      stk.resize(stk.size() - 3);
      for (int k = 0; k < 256; k++) ce[k] = 0;
      for (int i = b; i < e; i++) ce[(kx[i] >> s) & 255]++;
      mpoa::seedc::radix_walk(kx.data(), ki.data(), b, s, cb, ce);
      if (s == 0) continue;
      for (int k = (int)(rng() % 256), c = 0; c < 256; c++, k = (k + 1) & 255) {
        if (ce[k] - cb[k] > 64) stk.insert(stk.end(), {cb[k], ce[k], s > 8 ? s - 8 : 0});
        else if (ce[k] - cb[k] > 1) mpoa::seedc::insertion_run(kx.data(), ki.data(), cb[k], ce[k]);
      }
    }
    for (int i = 0; i < n; i++) {
      if (kx[i] != a[i].x || ki[i] != (uint32_t)a[i].y) {
        std::printf("FAIL radix order trial %d n %d shape %d at %d\n", t, n, shape, i);
        return 1;
      }
    }
  }
  return 0;
}

int main(int argc, char** argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 20000;
  std::mt19937_64 rng(12345);
  for (int t = 0; t < trials; t++)
    if (check_sketch(rng, t)) return 1;
  if (check_chain(rng)) return 1;
  if (check_radix(rng)) return 1;
  std::printf("ok %d sequences\n", trials);
  return 0;
}
