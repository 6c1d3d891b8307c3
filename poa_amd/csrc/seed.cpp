// Host seeding for -S windows (see seed.h). Steps follow minipoa:
//   sketch          src/minimizer.cpp:60-135 (forward strand only, :96)
//   hash            src/minimizer.cpp:11-21
//   sort            include/ksort.h:101-151 (klib MSD radix sort on .x, not stable)
//   anchors         src/minimizer.cpp:404-442, 493-570 (collect_anchors_bycons)
//   local chains    src/lchain.cpp:9-238 (mg_lchain_dp, backtrack, compact)
//   global chain    src/lchain.cpp:242-386 (chain_dp)
//   log2            include/mmpriv.h:12-20
#pragma clang fp contract(off)

#include "seed.h"

#include <algorithm>
#include <cstring>

namespace mpoa {
namespace seed {
namespace {

constexpr int32_t NEG_BIG = 1000000000;  // lchain.cpp:7
constexpr uint64_t NONE = ~0ULL;

uint64_t mix_key(uint64_t key, uint64_t mask) {
  key = (~key + (key << 21)) & mask;
  key = key ^ key >> 24;
  key = ((key + (key << 3)) + (key << 8)) & mask;
  key = key ^ key >> 14;
  key = ((key + (key << 2)) + (key << 4)) & mask;
  key = key ^ key >> 28;
  key = (key + (key << 31)) & mask;
  return key;
}

// Window minimizers of codes[0..len): x = hash << 8 | span, y = rid << 32 | pos << 1.
// Ties inside a window are all kept, as minipoa does.
void sketch(const uint8_t* codes, int len, int w, int k, uint32_t rid, std::vector<Pair>& out) {
  const uint64_t mask = (1ULL << 2 * k) - 1;
  uint64_t fwd = 0;
  int run = 0, at = 0, min_at = 0;
  Pair ring[256];
  Pair best = {NONE, NONE};
  for (int j = 0; j < w; j++) ring[j] = {NONE, NONE};
  auto flush_ties = [&](int from, int to, bool incl) {
    for (int j = from; incl ? j <= to : j < to; ++j)
      if (best.x == ring[j].x && ring[j].y != best.y) out.push_back(ring[j]);
  };
  for (int i = 0; i < len; ++i) {
    const int c = codes[i];
    Pair cur = {NONE, NONE};
    if (c < 4) {
      const int span = run + 1 < k ? run + 1 : k;
      fwd = (fwd << 2 | (uint64_t)c) & mask;
      ++run;
      if (run >= k && span < 256) {
        cur.x = mix_key(fwd, mask) << 8 | (uint64_t)span;
        cur.y = (uint64_t)rid << 32 | (uint32_t)i << 1;
      }
    } else {
      run = 0;
    }
    ring[at] = cur;
    if (run == w + k - 1 && best.x != NONE) {
      flush_ties(at + 1, w, false);
      flush_ties(0, at, false);
    }
    if (cur.x <= best.x) {
      if (run >= w + k && best.x != NONE) out.push_back(best);
      best = cur;
      min_at = at;
    } else if (at == min_at) {
      if (run >= w + k - 1 && best.x != NONE) out.push_back(best);
      best.x = NONE;
      for (int j = at + 1; j < w; ++j)
        if (best.x >= ring[j].x) best = ring[j], min_at = j;
      for (int j = 0; j <= at; ++j)
        if (best.x >= ring[j].x) best = ring[j], min_at = j;
      if (run >= w + k - 1 && best.x != NONE) {
        flush_ties(at + 1, w, false);
        flush_ties(0, at, true);
      }
    }
    if (++at == w) at = 0;
  }
  if (best.x != NONE) out.push_back(best);
}

void insertion_sort(Pair* b, Pair* e) {
  for (Pair* i = b + 1; i < e; ++i) {
    if (i->x < (i - 1)->x) {
      const Pair v = *i;
      Pair* j = i;
      while (j > b && v.x < (j - 1)->x) {
        *j = *(j - 1);
        --j;
      }
      *j = v;
    }
  }
}

// One MSD pass on bits [shift, shift+8) with in-place cycle permutation, then recursion.
void msd_pass(Pair* b, Pair* e, int shift) {
  struct Range {
    Pair *lo, *hi;
  };
  Range r[256];
  for (int k = 0; k < 256; k++) r[k].lo = r[k].hi = b;
  for (Pair* p = b; p != e; ++p) ++r[p->x >> shift & 255].hi;
  for (int k = 1; k < 256; k++) {
    r[k].hi += r[k - 1].hi - b;
    r[k].lo = r[k - 1].hi;
  }
  for (int k = 0; k < 256;) {
    if (r[k].lo == r[k].hi) {
      ++k;
      continue;
    }
    int dst = (int)(r[k].lo->x >> shift & 255);
    if (dst == k) {
      ++r[k].lo;
      continue;
    }
    Pair carry = *r[k].lo;
    do {
      const Pair placed = carry;
      carry = *r[dst].lo;
      *r[dst].lo++ = placed;
      dst = (int)(carry.x >> shift & 255);
    } while (dst != k);
    *r[k].lo++ = carry;
  }
  r[0].lo = b;
  for (int k = 1; k < 256; k++) r[k].lo = r[k - 1].hi;
  if (shift == 0) return;
  const int next = shift > 8 ? shift - 8 : 0;
  for (int k = 0; k < 256; k++) {
    const int64_t cnt = r[k].hi - r[k].lo;
    if (cnt > 64) msd_pass(r[k].lo, r[k].hi, next);
    else if (cnt > 1) insertion_sort(r[k].lo, r[k].hi);
  }
}

void sort_x(Pair* b, Pair* e) {
  if (e - b <= 64) insertion_sort(b, e);
  else msd_pass(b, e, 56);
}

float log2_approx(float x) {
  union {
    float f;
    uint32_t u;
  } z = {x};
  float r = (float)(int)((z.u >> 23) & 255) - 128;
  z.u &= ~(255u << 23);
  z.u += 127u << 23;
  r += (-0.34484843f * z.f + 2.02466578f) * z.f - 0.67487759f;
  return r;
}

int32_t pair_score(const Pair& i, const Pair& j, int32_t max_dist, float pen_gap, float pen_skip) {
  const int32_t dq = (int32_t)i.y - (int32_t)j.y;
  if (dq <= 0 || dq > max_dist) return -NEG_BIG;
  const int32_t dr = (int32_t)(i.x - j.x);
  const int32_t dd = dr > dq ? dr - dq : dq - dr;
  const int32_t dg = dr < dq ? dr : dq;
  const int32_t span = j.y >> 32 & 0xff;
  int32_t sc = span < dg ? span : dg;
  if (dd || dg > span) {
    const float lin = pen_gap * (float)dd + pen_skip * (float)dg;
    const float lg = dd >= 1 ? log2_approx((float)(dd + 1)) : 0.0f;
    if (dr == 0) ++sc;
    else sc -= (int)(lin < lg ? lin : lg);
  }
  return sc;
}

int64_t walk_end(int32_t max_drop, const Pair* z, const int32_t* f, const int64_t* p, int32_t* t,
                 int64_t k) {
  int64_t i = z[k].y, end = -1, best_i = i;
  int32_t best = 0;
  if (i < 0 || t[i] != 0) return i;
  do {
    t[i] = 2;
    end = i = p[i];
    const int32_t s = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
    if (s > best) best = s, best_i = i;
    else if (best - s > max_drop) break;
  } while (i >= 0 && t[i] == 0);
  for (i = z[k].y; i >= 0 && i != end; i = p[i]) t[i] = 0;
  return best_i;
}

// Local chains of anchors a (sorted by x). Returns the chained anchors, chain after chain,
// each ascending; chains[c] = (strand<<63 | tpos<<32 | qpos of its end, first<<32 | end).
std::vector<Pair> local_chains(int max_dist, int bw, int max_skip, int max_iter, int min_cnt,
                               int min_sc, float pen_gap, float pen_skip,
                               const std::vector<Pair>& a, std::vector<Pair>& chains) {
  chains.clear();
  std::vector<Pair> out;
  const int64_t n = (int64_t)a.size();
  if (n == 0) return out;
  const int32_t max_drop = bw;
  if (max_dist < bw) max_dist = bw;
  std::vector<int64_t> p(n);
  std::vector<int32_t> f(n), v(n), tb(n + 1, 0);
  // t[-1] reads 0: minipoa indexes t[p[i]] with p[i] = -1, which lands on the zero high half
  // of its allocator's size header.
  int32_t* t = tb.data() + 1;
  int64_t st = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t best_j = -1;
    int32_t best = a[i].y >> 32 & 0xff, skipped = 0, iters = 0;
    while (st < i && (a[i].x >> 32 != a[st].x >> 32 || a[i].x > a[st].x + max_dist)) ++st;
    if (i - st > 5000) st = i - 5000;
    for (int64_t j = i - 1; j >= st; --j) {
      int32_t sc = pair_score(a[i], a[j], max_dist, pen_gap, pen_skip);
      if (sc == -NEG_BIG) continue;
      sc += f[j];
      if (sc > best) {
        best = sc, best_j = j;
        if (skipped > 0) --skipped;
      } else {
        if (sc != best && ++iters > max_iter) break;
        if (t[j] == (int32_t)i && ++skipped > max_skip) break;
      }
      if (p[j] >= 0) t[p[j]] = (int32_t)i;
    }
    f[i] = best, p[i] = best_j;
  }
  // backtrack (lchain.cpp:66-135)
  std::memset(t, 0, n * sizeof(int32_t));
  int64_t n_end = 0;
  for (int64_t i = n - 1; i >= 0; i--) {
    if (t[p[i]] == 0 && f[i] >= 0 && f[i] >= min_sc) {
      ++n_end;
      t[i] = 2;
    }
    if (p[i] >= 0) t[p[i]] = 1;
  }
  if (n_end == 0) return out;
  std::vector<Pair> z(n_end);
  for (int64_t i = 0, k = 0; i < n; ++i)
    if (t[i] == 2) z[k].x = (uint64_t)(int64_t)f[i], z[k++].y = (uint64_t)i;
  sort_x(z.data(), z.data() + n_end);
  std::memset(t, 0, n * sizeof(int32_t));
  int64_t n_v = 0;
  for (int64_t k = n_end - 1; k >= 0; --k) {
    if (t[z[k].y] != 0) continue;
    const int64_t v0 = n_v;
    const int64_t end = walk_end(max_drop, z.data(), f.data(), p.data(), t, k);
    int64_t i;
    for (i = (int64_t)z[k].y; i != end; i = p[i]) v[n_v++] = (int32_t)i, t[i] = 1;
    const int32_t sc = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
    if (sc >= min_sc && n_v > v0 && n_v - v0 >= min_cnt) {
      const uint64_t e = z[k].y;
      const uint64_t strand = a[e].x >> 63, tpos = (uint32_t)a[e].x, qpos = (uint32_t)a[e].y;
      chains.push_back({strand << 63 | tpos << 32 | qpos, (uint64_t)v0 << 32 | (uint64_t)n_v});
    } else {
      n_v = v0;
    }
  }
  out.reserve(n_v);
  for (const Pair& c : chains) {
    const int32_t s0 = (int32_t)(c.y >> 32), e0 = (int32_t)c.y;
    for (int32_t j = e0 - 1; j >= s0; j--) out.push_back(a[v[j]]);
  }
  return out;
}

int chain_tail_score(int end_t, int end_q, int lo, int hi, const Pair* a, const int* score) {
  int l = lo, r = hi;
  while (l + 1 < r) {
    const int mid = (l + r) / 2;
    const int tm = (uint32_t)a[mid].x, qm = (uint32_t)a[mid].y;
    if (tm <= end_t || qm <= end_q) l = mid;
    else r = mid;
  }
  if (r == hi) return -NEG_BIG;
  return score[hi - 1] - score[r];
}

// Best chain of local chains, then its anchors picked back to front at least min_w apart.
std::vector<Pair> global_chain(const std::vector<Pair>& lc, std::vector<Pair>& a, int min_w,
                               int tlen, int qlen, int max_dist, float pen_gap, float pen_skip) {
  const int n_l = (int)lc.size();
  std::vector<int32_t> f(a.size());
  std::vector<int> best_sc(n_l), prev(n_l);
  int top = -NEG_BIG, top_i = -1;
  for (int i = 0, st = 0; i < n_l; i++) {
    const uint64_t xi = lc[i].x, yi = lc[i].y;
    const int strand = (int)(xi >> 63), s_i = (int)(yi >> 32), e_i = (int32_t)yi;
    const int s_t = (int32_t)a[s_i].x, s_q = (int32_t)a[s_i].y;
    f[s_i] = a[s_i].y >> 32 & 0xff;
    for (int j = s_i + 1; j < e_i; j++)
      f[j] = f[j - 1] + pair_score(a[j], a[j - 1], max_dist, pen_gap, pen_skip);
    int best_j = -1, best = f[e_i - 1];
    while (st < i && (int)(lc[st].x >> 63) != strand) ++st;
    for (int j = i - 1; j >= st; --j) {
      const uint64_t xj = lc[j].x, yj = lc[j].y;
      const int e_t = (int)((xj >> 32) & 0x7fffffff), e_q = (int32_t)xj;
      const int s_j = (int)(yj >> 32);
      if (s_t < (int32_t)a[s_j].x) continue;
      const int sc = s_t > e_t && s_q > e_q
                         ? best_sc[j] + f[e_i - 1]
                         : best_sc[j] + chain_tail_score(e_t, e_q, s_i, e_i, a.data(), f.data());
      if (sc > best) best = sc, best_j = j;
    }
    best_sc[i] = best;
    prev[i] = best_j;
    if (best > top) top = best, top_i = i;
  }
  if (top_i < 0) return a;  // minipoa leaves the anchors as they are
  std::vector<Pair> out;
  int cur = top_i, pre = prev[top_i];
  uint64_t cur_y = lc[cur].y;
  int last_t = tlen, last_q = qlen, skipped = 0;
  int i = (int)(uint32_t)cur_y - 1;
  auto take = [&](int idx, int t, int q) {
    if (last_t - t >= min_w && last_q - q >= min_w) {
      if (skipped < min_w / 10) a[idx].y |= 1ULL << 40;
      skipped = 0;
      out.push_back(a[idx]);
      last_t = t, last_q = q;
    } else {
      skipped++;
    }
  };
  while (pre != -1) {
    const uint64_t px = lc[pre].x, py = lc[pre].y;
    const int pe_t = (int)((px >> 32) & 0x7fffffff), pe_q = (int32_t)px;
    while (i >= 0) {
      const int t = (int)(uint32_t)a[i].x, q = (int32_t)a[i].y;
      if (!(t > pe_t && q > pe_q)) break;
      take(i, t, q);
      i--;
      if ((uint64_t)(int64_t)i == (cur_y >> 32)) break;
    }
    cur = pre, pre = prev[pre], cur_y = py;
    i = (int)(uint32_t)cur_y - 1;
  }
  while (i >= 0) {
    take(i, (int)(uint32_t)a[i].x, (int32_t)a[i].y);
    if ((uint64_t)(int64_t)i == (cur_y >> 32)) break;
    i--;
  }
  std::reverse(out.begin(), out.end());
  return out;
}

}  // namespace

std::vector<int> guide_order(const Config& c, int n_reads, const uint8_t* codes,
                             const int64_t* off) {
  std::vector<int> ord(n_reads);
  for (int i = 0; i < n_reads; i++) ord[i] = i;
  if (n_reads < 2) return ord;
  std::vector<Pair> mins;
  std::vector<int> len(n_reads);
  int maxl = 0;
  for (int i = 0; i < n_reads; i++) {
    len[i] = (int)(off[i + 1] - off[i]);
    maxl = std::max(maxl, len[i]);
    sketch(codes + off[i], len[i], c.w, c.k, (uint32_t)i, mins);
  }
  if (mins.empty()) return ord;
  // minimizers grouped by hash: each read's total, and for every pair of reads the smaller of
  // their counts in each group, summed (minimizer.cpp:211-259)
  sort_x(mins.data(), mins.data() + mins.size());
  const size_t n = (size_t)n_reads;
  std::vector<int64_t> total(n, 0), shared(n * n, 0);
  std::vector<int> cnt(n, 0), present;
  for (size_t g0 = 0; g0 < mins.size();) {
    size_t g1 = g0;
    while (g1 < mins.size() && mins[g1].x == mins[g0].x) g1++;
    present.clear();
    for (size_t t = g0; t < g1; t++) {
      const int r = (int)(mins[t].y >> 32);
      if (cnt[r]++ == 0) present.push_back(r);
      total[r]++;
    }
    for (size_t a = 0; a < present.size(); a++)
      for (size_t b = 0; b < present.size(); b++)
        if (present[a] < present[b])
          shared[(size_t)present[b] * n + present[a]] += std::min(cnt[present[a]], cnt[present[b]]);
    for (int r : present) cnt[r] = 0;
    g0 = g1;
  }
  // length-weighted Jaccard similarity of every ordered pair (minimizer.cpp:264-290): the
  // weights make it asymmetric; the first pair of the largest value opens the order
  std::vector<double> sim(n * n, 0.0);
  double best = -1.0;
  int bi = -1, bj = -1;
  for (size_t i = 0; i < n; i++) {
    for (size_t j = 0; j < n; j++) {
      if (i == j) continue;
      const int64_t sh = i > j ? shared[i * n + j] : shared[j * n + i];
      const int64_t tot = total[i] + total[j] - sh;
      double jac = 0.0;
      if (tot > 0) {
        jac = (0.0 + (double)sh) / (double)tot;
        jac = jac * ((0.51 * len[i] / maxl) + (0.49 * len[j] / maxl));
      }
      sim[i * n + j] = jac;
      if (jac > best) {
        best = jac;
        bi = (int)i;
        bj = (int)j;
      }
    }
  }
  // then, one read at a time, the one whose similarities from the reads already ordered sum
  // highest, summed in their order (minimizer.cpp:295-330)
  ord[0] = bi;
  ord[1] = bj;
  std::vector<char> placed(n, 0);
  placed[bi] = placed[bj] = 1;
  for (int at = 2; at < n_reads; at++) {
    double top = -1.0;
    int pick = n_reads;
    for (int r = 0; r < n_reads; r++) {
      if (placed[r]) continue;
      double sum = 0.0;
      for (int t = 0; t < at; t++) sum += sim[(size_t)ord[t] * n + r];
      if (sum > top) {
        top = sum;
        pick = r;
      }
    }
    ord[at] = pick;
    placed[pick] = 1;
  }
  return ord;
}

void Seeder::init(const Config& c, int n_reads, const uint8_t* codes, const int64_t* off) {
  cfg_ = c;
  n_reads_ = n_reads;
  mins_.clear();
  start_.assign(n_reads + 2, 0);
  for (int i = 0; i < n_reads; i++) {
    sketch(codes + off[i], (int)(off[i + 1] - off[i]), c.w, c.k, (uint32_t)i, mins_);
    start_[i + 1] = (int64_t)mins_.size();
  }
  for (int i = 0; i < n_reads; i++) sort_x(mins_.data() + start_[i], mins_.data() + start_[i + 1]);
}

std::vector<Pair> Seeder::anchors(int rid, int qlen, const uint8_t* cons, int cons_len) {
  const int n = n_reads_;
  mins_.resize(start_[n]);
  sketch(cons, cons_len, cfg_.w, cfg_.k, (uint32_t)n, mins_);
  start_[n + 1] = (int64_t)mins_.size();
  sort_x(mins_.data() + start_[n], mins_.data() + start_[n + 1]);
  // Join the consensus's and the read's sorted minimizers on the hash.
  std::vector<Pair> a;
  int64_t i = start_[n], j = start_[rid];
  const int64_t ie = start_[n + 1], je = start_[rid + 1];
  while (i < ie && j < je) {
    const uint64_t xi = mins_[i].x, xj = mins_[j].x;
    if (xi < xj) {
      ++i;
    } else if (xi > xj) {
      ++j;
    } else {
      int64_t ii = i, jj = j;
      for (; ii < ie && mins_[ii].x == xi; ++ii) {
        const uint64_t yi = mins_[ii].y;
        for (jj = j; jj < je && mins_[jj].x == xj; ++jj) {
          const uint64_t yj = mins_[jj].y;
          const uint32_t span = (uint32_t)(mins_[jj].x & 0xff);
          a.push_back({(yi & 0x7fffffff00000000ULL) | ((uint32_t)yi >> 1),
                       (uint64_t)span << 32 | ((uint32_t)yj >> 1)});
        }
      }
      i = ii, j = jj;
    }
  }
  sort_x(a.data(), a.data() + a.size());
  int min_w = std::min(20000, (cons_len + 24 - 1) / 24);
  if (cfg_.poa_w) min_w = std::min(min_w, cfg_.poa_w + cfg_.k);
  const int max_dist = cfg_.bw;
  const float pen_gap = (float)(0.01 * cfg_.k), pen_skip = (float)(0.0 * 0.01 * cfg_.k);
  std::vector<Pair> lc;
  std::vector<Pair> chained =
      local_chains(max_dist, cfg_.bw, 25, 50, 3, 100, pen_gap, pen_skip, a, lc);
  if (lc.empty()) return {};
  sort_x(lc.data(), lc.data() + lc.size());
  return global_chain(lc, chained, min_w, cons_len, qlen, max_dist, pen_gap, pen_skip);
}

}  // namespace seed
}  // namespace mpoa
