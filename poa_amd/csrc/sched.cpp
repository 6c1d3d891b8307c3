// Scheduling features: for every DP call of a batch (window w, read r >= 1 against the graph of
// reads 0..r-1), eight numbers known before anything runs. They feed the scheduling network
// (mlp.hip), whose predictions order the work queue and size the per-slot arenas.
//
//   0 read length                 4 anchor count: read minimizers shared with read 0
//   1 graph node count estimate   5 previous read's static band extent
//   2 graph edge count estimate   6 error-rate estimate of the window
//   3 chain coverage: shared      7 score width in bytes (int32: 4)
//     fraction of the read's minimizers
//
// Minimizers here are this file's own (forward strand, k = 11, w = 5, an invertible 64-bit
// mix as the order); they only estimate similarity and are not minipoa's seeding. A shared
// minimizer needs K intact bases in both reads, so coverage ~ (1 - err)^(2K).
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "mpoa.h"

namespace {

constexpr int K = 11, WIN = 5;

inline uint64_t mix(uint64_t x) {
  x = (x ^ (x >> 31)) * 0x7fb5d329728ea185ULL;
  x = (x ^ (x >> 27)) * 0x81dadef4bc2dd44dULL;
  return x ^ (x >> 33);
}

inline int code2(unsigned char c) {
  switch (c | 0x20) {
    case 'a': return 0;
    case 'c': return 1;
    case 'g': return 2;
    case 't': case 'u': return 3;
    default: return -1;
  }
}

// Window minimizers of a read (hash values, in order, consecutive duplicates dropped).
void minimizers(const char* s, int64_t len, std::vector<uint64_t>& out) {
  out.clear();
  const uint64_t mask = (1ULL << (2 * K)) - 1;
  uint64_t kmer = 0;
  int valid = 0;
  uint64_t ring[WIN];
  int filled = 0;
  uint64_t last = ~0ULL;
  for (int64_t i = 0; i < len; i++) {
    const int c = code2((unsigned char)s[i]);
    if (c < 0) {
      valid = 0;
      filled = 0;
      continue;
    }
    kmer = ((kmer << 2) | (uint64_t)c) & mask;
    if (++valid < K) continue;
    ring[filled % WIN] = mix(kmer);
    filled++;
    if (filled < WIN) continue;
    uint64_t mn = ring[0];
    for (int j = 1; j < WIN; j++) mn = std::min(mn, ring[j]);
    if (mn != last) out.push_back(mn);
    last = mn;
  }
}

void window_features(const mpoa_params& p, int32_t r0, int32_t r1, const char* seq,
                     const int64_t* seq_off, float* X) {
  const int n = r1 - r0;
  if (n < 2) return;
  std::vector<uint64_t> ref, cur;
  minimizers(seq + seq_off[r0], seq_off[r0 + 1] - seq_off[r0], ref);
  std::sort(ref.begin(), ref.end());
  ref.erase(std::unique(ref.begin(), ref.end()), ref.end());
  std::vector<float> cov(n, 0.f), shared(n, 0.f);
  double cov_sum = 0;
  for (int r = 1; r < n; r++) {
    minimizers(seq + seq_off[r0 + r], seq_off[r0 + r + 1] - seq_off[r0 + r], cur);
    int64_t c = 0;
    for (uint64_t h : cur) c += std::binary_search(ref.begin(), ref.end(), h) ? 1 : 0;
    shared[r] = (float)c;
    cov[r] = cur.empty() ? 0.f : (float)c / (float)cur.size();
    cov_sum += cov[r];
  }
  const double mean_cov = std::max(cov_sum / (n - 1), 1e-6);
  const double err = std::min(0.5, std::max(0.0, 1.0 - std::pow(mean_cov, 1.0 / (2 * K))));
  const double growth = 0.7 * err;
  const int64_t q0 = seq_off[r0 + 1] - seq_off[r0];
  double cum = 0;
  for (int r = 1; r < n; r++) {
    const int64_t q = seq_off[r0 + r + 1] - seq_off[r0 + r];
    const int64_t qp = seq_off[r0 + r] - seq_off[r0 + r - 1];
    const double nodes = 2.0 + (double)q0 + growth * cum;
    const double edges = nodes + growth * cum + r;
    const int mprev = (int)qp + 1;
    const double prev_w =
        p.band_f > 0 ? 2.0 * (p.band_b + mprev / p.band_f) : (double)(mprev + 1);
    float* x = X + (int64_t)(r - 1) * 8;
    x[0] = (float)q;
    x[1] = (float)nodes;
    x[2] = (float)edges;
    x[3] = cov[r];
    x[4] = shared[r];
    x[5] = (float)prev_w;
    x[6] = (float)err;
    x[7] = 4.f;
    cum += (double)q;
  }
}

}  // namespace

extern "C" int mpoa_sched_features(const mpoa_params* params, int32_t n_windows,
                                   const int32_t* win_off, const char* seq,
                                   const int64_t* seq_off, float* X, int64_t cap,
                                   int64_t* n_rows) {
  if (!params || n_windows < 0 || !win_off || (n_windows > 0 && (!seq_off || !seq)))
    return MPOA_E_ARG;
  std::vector<int64_t> row0(n_windows + 1, 0);
  for (int w = 0; w < n_windows; w++)
    row0[w + 1] = row0[w] + std::max(0, win_off[w + 1] - win_off[w] - 1);
  if (n_rows) *n_rows = row0[n_windows];
  if (!X) return MPOA_OK;
  if (cap < row0[n_windows]) return MPOA_E_OUT_CAP;
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++) {
    pool.emplace_back([&, t] {
      for (int w = t; w < n_windows; w += nt)
        window_features(*params, win_off[w], win_off[w + 1], seq, seq_off, X + row0[w] * 8);
    });
  }
  for (auto& th : pool) th.join();
  return MPOA_OK;
}
