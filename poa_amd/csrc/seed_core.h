// Pieces of -S seeding that the device runs per position or per anchor pair, written so that
// the host compiles the same code (tests/sketch_check.cpp runs them against the oracle):
//   the k-mer hash          src/minimizer.cpp:11-21
//   minimizer emission      src/minimizer.cpp:60-135 (mm_sketch, is_hpc 0, strand 0)
//   the chaining log2       include/mmpriv.h:12-20
//   the anchor-pair score   src/lchain.cpp:28-63 (comput_sc as minipoa calls it)
//   radix sort steps        include/ksort.h:101-151 (a digit pass, the insertion sort)
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace mpoa {
namespace seedc {

constexpr uint64_t NONE = ~0ull;        // an empty minimizer slot (x and y all ones)
constexpr int CHAIN_NEG = -1000000000;  // "no link" score (lchain.cpp:7)

// The invertible integer hash of a 2k-bit k-mer (minimizer.cpp:11-21); every step is masked
// to the k-mer's bits.
__host__ __device__ inline uint64_t kmer_hash(uint64_t v, uint64_t mask) {
  v = (~v + (v << 21)) & mask;
  v ^= v >> 24;
  v = (v + (v << 3) + (v << 8)) & mask;
  v ^= v >> 14;
  v = (v + (v << 2) + (v << 4)) & mask;
  v ^= v >> 28;
  v = (v + (v << 31)) & mask;
  return v;
}

// Latest position of the smallest x over positions [a, b] (a clipped at 0); -1 and NONE for an
// empty range.
template <class XF>
__host__ __device__ inline int latest_min(XF x, int a, int b, uint64_t& mx) {
  int at = -1;
  mx = NONE;
  for (int j = a < 0 ? 0 : a; j <= b; j++) {
    const uint64_t v = x(j);
    if (v <= mx) mx = v, at = j;
  }
  return at;
}

// The minimizers mm_sketch emits while it takes position i, in its order, as positions given
// to out(j); returns how many. x(j) is position j's x (hash << 8 | k where the k bases ending
// at j are all A/C/G/T, else NONE). Sequential mm_sketch keeps a w-slot ring; its state after
// position i is fixed by the last w x values alone:
//   * its minimum is the latest position of the smallest x among positions [i-w+1, i]
//     (a new x <= the minimum replaces it; a minimum leaving the ring is searched again
//     with >=, i.e. the latest of the smallest);
//   * its run counter l (valid bases since the last ambiguous one) enters only as l >= k+t,
//     which holds exactly when positions i-t .. i all have valid x;
// so with prev the minimum after position i-1 and v the valid x's ending at i:
//   v == w and prev valid:        the other positions in [i-w+1, i-1] with x == prev.x
//                                 (the first full window, minimizer.cpp:106-111);
//   x(i) <= prev.x:               prev, if v >= w+1 and prev valid (:112-114);
//   else prev at i-w (leaves):    prev if v >= w; then with cur the minimum of [i-w+1, i],
//                                 the other positions with x == cur.x if v >= w (:116-127);
// and after the last position, the minimum if valid (:131-132).
template <class XF, class OUT>
__host__ __device__ inline int sketch_step(int i, int len, int w, XF x, OUT out) {
  int n = 0;
  uint64_t px;
  const int pp = latest_min(x, i - w, i - 1, px);
  const uint64_t xi = x(i);
  int v = 0;  // valid x's ending at i, counted up to w + 1
  for (int j = i; j >= 0 && j >= i - w && x(j) != NONE; j--) v++;
  if (v == w && px != NONE) {
    for (int j = i - w + 1; j < i; j++)
      if (j >= 0 && j != pp && x(j) == px) out(j), n++;
  }
  if (xi <= px) {
    if (v >= w + 1 && px != NONE) out(pp), n++;
  } else if (pp == i - w) {
    if (v >= w) out(pp), n++;
    uint64_t cx;
    const int cp = latest_min(x, i - w + 1, i, cx);
    if (v >= w && cx != NONE) {
      for (int j = i - w + 1; j <= i; j++)
        if (j != cp && x(j) == cx) out(j), n++;
    }
  }
  if (i == len - 1) {
    uint64_t cx;
    const int cp = latest_min(x, len - w, len - 1, cx);
    if (cx != NONE) out(cp), n++;
  }
  return n;
}

// x of the k-mer ending at position i of codes[0..len) (0..3 bases, anything else
// ambiguous): hash << 8 | k, or NONE.
template <class CF>
__host__ __device__ inline uint64_t kmer_x(CF code, int i, int k) {
  if (i < k - 1) return NONE;
  const uint64_t mask = (k >= 32 ? ~0ull : (1ull << (2 * k)) - 1);
  uint64_t km = 0;
  for (int t = i - k + 1; t <= i; t++) {
    const int c = code(t);
    if (c > 3) return NONE;
    km = (km << 2 | (uint64_t)c) & mask;
  }
  return kmer_hash(km, mask) << 8 | (uint64_t)k;
}

// log2 approximation of the chaining penalty, exponent plus a quadratic in the mantissa
// (mmpriv.h:12-20; valid for v >= 2). Evaluated without fused multiply-adds, as upstream.
__host__ __device__ inline float chain_log2(float v) {
#pragma clang fp contract(off)
  uint32_t b = __builtin_bit_cast(uint32_t, v);
  float e = (float)((int)((b >> 23) & 255) - 128);
  b = (b & ~(255u << 23)) + (127u << 23);
  const float m = __builtin_bit_cast(float, b);
  e += (-0.34484843f * m + 2.02466578f) * m - 0.67487759f;
  return e;
}

// Score of chaining anchor j (earlier) to anchor i (lchain.cpp:28-63 with is_equal false,
// is_cdna 0, as mg_lchain_dp and chain_dp call it): CHAIN_NEG when the query step is not in
// (0, max_dist]. Positions are the anchors' last bases; span is j's k-mer span.
__host__ __device__ inline int chain_link(int tpi, int qpi, int tpj, int qpj, int span,
                                          int max_dist, float pen_gap, float pen_skip) {
#pragma clang fp contract(off)
  const int dq = qpi - qpj;
  if (dq <= 0 || dq > max_dist) return CHAIN_NEG;
  const int dr = tpi - tpj;
  const int dd = dr > dq ? dr - dq : dq - dr;
  const int dg = dr < dq ? dr : dq;
  int sc = span < dg ? span : dg;
  if (dd || dg > span) {
    const float lin = pen_gap * (float)dd + pen_skip * (float)dg;
    const float lg = dd >= 1 ? chain_log2((float)(dd + 1)) : 0.0f;
    if (dr == 0) ++sc;
    else sc -= (int)(lin < lg ? lin : lg);
  }
  return sc;
}

// One digit pass of minipoa's in-place radix sort (klib, ksort.h:112-129) on (kx, ki) from b:
// ce[0..256) holds the digit counts of the range on entry; on return cb/ce bound each digit's
// bucket. Entries move along cycles, each taken to the next free slot of its digit's bucket,
// which fixes the order equal keys leave in.
__host__ __device__ inline void radix_walk(uint64_t* kx, uint32_t* ki, int b, int s, int* cb,
                                           int* ce) {
  int at = b;
  for (int k = 0; k < 256; k++) {
    cb[k] = at;
    at += ce[k];
    ce[k] = at;
  }
  for (int k = 0; k < 256;) {
    if (cb[k] == ce[k]) {
      k++;
      continue;
    }
    int l = (int)((kx[cb[k]] >> s) & 255);
    if (l == k) {
      cb[k]++;
      continue;
    }
    uint64_t tx = kx[cb[k]];
    uint32_t ti = ki[cb[k]];
    do {  // carry the entry in hand to its bucket and pick up the one it displaces
      const uint64_t hx = tx;
      const uint32_t hi = ti;
      tx = kx[cb[l]];
      ti = ki[cb[l]];
      kx[cb[l]] = hx;
      ki[cb[l]] = hi;
      cb[l]++;
      l = (int)((tx >> s) & 255);
    } while (l != k);
    kx[cb[k]] = tx;
    ki[cb[k]] = ti;
    cb[k]++;
  }
  cb[0] = b;
  for (int k = 1; k < 256; k++) cb[k] = ce[k - 1];
}

// Stable insertion sort of (kx, ki)[b, e) by kx (ksort.h:102-111).
__host__ __device__ inline void insertion_run(uint64_t* kx, uint32_t* ki, int b, int e) {
  for (int i = b + 1; i < e; i++) {
    if (!(kx[i] < kx[i - 1])) continue;
    const uint64_t tx = kx[i];
    const uint32_t ti = ki[i];
    int j = i;
    for (; j > b && tx < kx[j - 1]; --j) {
      kx[j] = kx[j - 1];
      ki[j] = ki[j - 1];
    }
    kx[j] = tx;
    ki[j] = ti;
  }
}

}  // namespace seedc
}  // namespace mpoa
