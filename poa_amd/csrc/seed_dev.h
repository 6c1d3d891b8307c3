// Device seeding of -S reads (included by pipeline.hip inside namespace mpoa). minipoa seeds
// each read on the host against the consensus of the graph so far (align.cpp:1752-1766,
// minimizer.cpp:537-570): minimizers of the consensus, anchors shared with the read, local
// chains (lchain.cpp:158-238), the best global chain (lchain.cpp:261-386), and sub-windows
// between its anchors. Here the workgroup that committed a read seeds the next one:
//   * every position's k-mer x in parallel, then every position's emitted minimizers
//     (seedc::sketch_step, a per-position form of mm_sketch) written at prefix-sum offsets;
//   * sorts by x as bitonic sorts of (x, index) in LDS: the order of a stable sort, which is
//     the klib radix sort's order whenever keys are distinct or at most 64 (insertion sort);
//     with equal keys among more than 64, the radix sort's own steps (klib_order);
//   * the join with the read's sorted minimizers by binary search per consensus minimizer;
//   * the local-chain DP with the anchors in LDS: one wavefront scores 64 predecessors at once,
//     then takes minipoa's early-exit walk over them in order with scalar reads;
//   * chain ends picked in parallel, chain walks, global chain and sub-windows on one lane.

// Block-wide exclusive prefix sum over NT threads; sh holds NT/64 ints of LDS.
template <int NT>
__device__ __forceinline__ int seed_excl_sum(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const int u = __shfl_up(incl, s, 64);
    if (lane >= s) incl += u;
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w2 = 0; w2 < NT / 64; w2++) {
    const int t = sh[w2];
    if (w2 < wave) before += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// x of every position of a sequence of len bases (code(i): 0..3, else ambiguous) into hx.
template <int NT, class CF>
__device__ void seed_positions(CF code, int len, int k, uint64_t* hx) {
  for (int i = threadIdx.x; i < len; i += NT) hx[i] = seedc::kmer_x(code, i, k);
  __syncthreads();
}

// mm_sketch's minimizers of positions [0, len) with x in hx, in its order, into (ox, oy):
// y = rid << 32 | position << 1. Returns the count, or -1 beyond cap (nothing past cap is
// written).
template <int NT>
__device__ int seed_emit(const uint64_t* hx, int len, int w, uint32_t rid, uint64_t* ox,
                         uint64_t* oy, int cap, int* sh) {
  auto xf = [&](int j) { return hx[j]; };
  int base = 0;
  for (int c0 = 0; c0 < len; c0 += NT) {
    const int i = c0 + threadIdx.x;
    const int cnt = i < len ? seedc::sketch_step(i, len, w, xf, [](int) {}) : 0;
    int total;
    const int ex = seed_excl_sum<NT>(cnt, sh, &total);
    if (cnt > 0 && base + ex + cnt <= cap) {
      int o = base + ex;
      seedc::sketch_step(i, len, w, xf, [&](int j) {
        ox[o] = hx[j];
        oy[o] = (uint64_t)rid << 32 | (uint64_t)(uint32_t)j << 1;
        o++;
      });
    }
    base += total;
  }
  __syncthreads();
  return base <= cap ? base : -1;
}

// LDS the radix-order emulation needs besides the keys: bucket bounds, the range stack.
constexpr int SEED_SORT_AUX = 4096;

// minipoa's in-place MSD radix sort on x (klib, ksort.h:101-151: 8-bit digits from the top
// byte, buckets above 64 entries sorted again on the next digit, smaller ones by insertion
// sort) applied to (kx, ki)[0..n), n > 64, step for step where the order of equal keys depends
// on it: the in-place cycle walk that moves every entry to its digit's bucket, on one lane.
// Digit counts and the insertion sorts of the small buckets (one bucket per thread) run in
// parallel; pending buckets wait on a stack (disjoint ranges, so their order does not matter).
template <int NT>
__device__ void klib_order(uint64_t* kx, uint32_t* ki, int n, int* aux) {
  const int tid = threadIdx.x;
  int* cb = aux;          // 256 bucket begins
  int* ce = aux + 256;    // 256 bucket ends
  int* stk = aux + 512;   // (begin, end, shift) triples
  int* top = aux + 1020;  // stack height; aux[1021..1023]: the range taken
  constexpr int STK = 160;
  if (tid == 0) {
    stk[0] = 0; stk[1] = n; stk[2] = 56;
    *top = 1;
  }
  __syncthreads();
  while (true) {
    if (tid == 0 && *top > 0) {
      const int t = --*top;
      aux[1021] = stk[3 * t]; aux[1022] = stk[3 * t + 1]; aux[1023] = stk[3 * t + 2];
    } else if (tid == 0) {
      aux[1021] = -1;
    }
    __syncthreads();
    const int b = aux[1021], e = aux[1022], s = aux[1023];
    if (b < 0) break;
    for (int t = tid; t < 256; t += NT) ce[t] = 0;
    __syncthreads();
    for (int i = b + tid; i < e; i += NT) atomicAdd(&ce[(int)((kx[i] >> s) & 255)], 1);
    __syncthreads();
    if (tid == 0) seedc::radix_walk(kx, ki, b, s, cb, ce);
    __syncthreads();
    if (s > 0) {
      const int s2 = s > 8 ? s - 8 : 0;
      for (int k = tid; k < 256; k += NT) {
        const int kb = cb[k], ke = ce[k];
        if (ke - kb > 64) {
          const int t = atomicAdd(top, 1);
          if (t < STK) { stk[3 * t] = kb; stk[3 * t + 1] = ke; stk[3 * t + 2] = s2; }
        } else if (ke - kb > 1) {
          seedc::insertion_run(kx, ki, kb, ke);
        }
      }
    }
    __syncthreads();
  }
}

// (sx, sy)[0..n) sorted by x into (dx, dy) in minipoa's order (ksort.h:101-151): a bitonic
// sort of (x, index) pairs in lds gives the order of a stable sort, which is the radix sort's
// whenever the keys are distinct or at most 64 (insertion sort); with equal keys among more
// than 64 entries, klib_order repeats the radix sort's own steps. False when n does not fit.
template <int NT>
__device__ bool seed_sort(const uint64_t* sx, const uint64_t* sy, uint64_t* dx, uint64_t* dy,
                          int n, uint8_t* lds, int lds_bytes) {
  const int tid = threadIdx.x;
  if (n <= 0) return true;
  int P = 1;
  while (P < n) P <<= 1;
  if ((int64_t)P * 12 + SEED_SORT_AUX > lds_bytes) return false;
  uint64_t* kx = reinterpret_cast<uint64_t*>(lds);
  uint32_t* ki = reinterpret_cast<uint32_t*>(kx + P);
  int* aux = reinterpret_cast<int*>(ki + P);
  for (int t = tid; t < P; t += NT) {
    kx[t] = t < n ? sx[t] : seedc::NONE;
    ki[t] = (uint32_t)t;
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = tid; t < P / 2; t += NT) {
        const int lo = 2 * stride * (t / stride) + (t % stride), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t xl = kx[lo], xh = kx[hi];
        const uint32_t il = ki[lo], ih = ki[hi];
        const bool gt = xl > xh || (xl == xh && il > ih);
        if (gt == up) {
          kx[lo] = xh; kx[hi] = xl;
          ki[lo] = ih; ki[hi] = il;
        }
      }
      __syncthreads();
    }
  }
  bool tie = false;
  for (int t = tid; t + 1 < n; t += NT) tie |= kx[t] == kx[t + 1];
  tie = __syncthreads_or(tie);
  if (tie && n > 64) {
    for (int t = tid; t < n; t += NT) {
      kx[t] = sx[t];
      ki[t] = (uint32_t)t;
    }
    __syncthreads();
    klib_order<NT>(kx, ki, n, aux);
  }
  for (int t = tid; t < n; t += NT) {
    dx[t] = kx[t];
    dy[t] = sy[ki[t]];
  }
  __syncthreads();
  return true;
}

// First index in x[0..n) (sorted) with x >= v, or > v when upper.
__device__ __forceinline__ int seed_bound(const uint64_t* x, int n, uint64_t v, bool upper) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const uint64_t m = x[mid];
    if (m < v || (upper && m == v)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Where read r's minimizers start in the window's minimizer arrays (capacity qlen + 2w + 8 per
// read): the host's offsets when it gave them, else the read's offset in the window plus
// 2w + 8 per read before it.
__device__ __forceinline__ int64_t seed_rm0(const SeedDev& d, const KernelArgs& a,
                                            const WinDesc& wd, int r) {
  if (d.rm_off) return d.rm_off[r];
  return (a.read_off[wd.read0 + r] - a.read_off[wd.read0]) + (int64_t)r * (2 * d.w + 8);
}

// Minimizers of read r of window wd, sorted by x (the whole workgroup). False when they
// outgrow their capacity or the LDS.
template <int NT>
__device__ bool seed_read_minimizers(const KernelArgs& a, const WinDesc& wd, const SeedDev& d,
                                     int r, uint8_t* lds, int lds_bytes, int* sh) {
  const int64_t q0 = a.read_off[wd.read0 + r], base0 = a.read_off[wd.read0];
  const int len = (int)(a.read_off[wd.read0 + r + 1] - q0);
  const uint8_t* seq = a.read_seq + q0;
  uint64_t* hx = d.hx + (q0 - base0);
  seed_positions<NT>([&](int i) { return code_of(seq[i]); }, len, d.k, hx);
  const int64_t m0 = seed_rm0(d, a, wd, r);
  const int cap = len + 2 * d.w + 8;
  const int n = seed_emit<NT>(hx, len, d.w, (uint32_t)r, d.rtx + m0, d.rty + m0, cap, sh);
  bool ok = n >= 0;
  if (ok) ok = seed_sort<NT>(d.rtx + m0, d.rty + m0, d.rmx + m0, d.rmy + m0, n, lds, lds_bytes);
  if (threadIdx.x == 0) {
    d.rm_cnt[r] = ok ? n : 0;
    if (!ok) atomicOr(d.flag, SEED_WHY_READS);
  }
  __syncthreads();
  return ok;
}

// Minimizers of read r (one workgroup per read), sorted by x.
template <int NT>
__global__ __launch_bounds__(NT) void k_seed_reads(SeedArgs s) {
  __shared__ uint64_t lds8[(SEED_SORT * 12 + SEED_SORT_AUX) / 8];
  __shared__ int sh[NT / 64];
  seed_read_minimizers<NT>(s.graph, s.graph.win[0], s.dev, blockIdx.x,
                           reinterpret_cast<uint8_t*>(lds8), (int)sizeof(lds8), sh);
}

// The pieces of the local chains kept in LDS (one int per anchor each).
struct ChainLds {
  int *tp, *qp, *f, *p, *t, *v;
};

// Seeds read r1 of window wd against the consensus the commit of read r1 - 1 just built
// (A.cons_*, S.cons_len) and writes its sub-windows (alignment()'s -S branch,
// align.cpp:1752-1797) at index ri of the per-read outputs (n_seg, n_splice, segs, splice).
// lds: lds_bytes of free LDS. Returns false when a capacity is exceeded (*flag set, S.status
// MPOA_E_OVERFLOW) or the read ends the window (MPOA_E_EMPTY_WINDOW).
template <int NT>
__device__ bool seed_next(const KernelArgs& a, const WinDesc& wd, const SeedDev& d, int r1,
                          int ri, Arena& A, WinShared& S, uint8_t* lds, int lds_bytes) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t rq0 = a.read_off[wd.read0 + r1];
  const int qlen = (int)(a.read_off[wd.read0 + r1 + 1] - rq0);
  int* sh = reinterpret_cast<int*>(lds + lds_bytes - 64);
  int* ctl = sh + 8;  // one-lane results shared with the block
  lds_bytes -= 64;
  const int k = d.k, max_dist = d.bw;
  auto fail = [&](int why) {
    if (tid == 0) {
      atomicOr(d.flag, why);
      S.status = MPOA_E_OVERFLOW;
      if (d.n_seg) {  // absent when only the anchors are asked for
        d.n_seg[ri] = 0;
        d.n_splice[ri] = 0;
      }
    }
    __syncthreads();
    return false;
  };
  if (S.n_nodes <= 2) {  // alignment(): a graph without bases takes the read as a new path
    if (tid == 0) { d.n_seg[ri] = 0; d.n_splice[ri] = 0; }
    __syncthreads();
    return true;
  }
  const int L = S.cons_len;
  if (L > d.cons_cap) return fail(SEED_WHY_CONS);
  // Consensus minimizers (collect_anchors_bycons, minimizer.cpp:541-551), sorted by x.
  uint64_t* hx = d.hx + d.hx_cons;
  seed_positions<NT>([&](int i) { return code_of(A.cons_ch[i]); }, L, k, hx);
  const int n_c = seed_emit<NT>(hx, L, d.w, (uint32_t)wd.n_reads, d.ctx, d.cty, d.cons_cap, sh);
  if (n_c < 0) return fail(SEED_WHY_CONS_MM);
  if (!seed_sort<NT>(d.ctx, d.cty, d.cmx, d.cmy, n_c, lds, lds_bytes))
    return fail(SEED_WHY_CONS_SORT);
  // Anchors (collect_anchors, minimizer.cpp:404-442): for each consensus minimizer in order,
  // every read minimizer with the same x in order; x = rid << 32 | tpos, y = span << 32 | qpos.
  const uint64_t* qx = d.rmx + seed_rm0(d, a, wd, r1);
  const uint64_t* qy = d.rmy + seed_rm0(d, a, wd, r1);
  const int n_q = d.rm_cnt[r1];
  int n_a = 0;
  for (int c0 = 0; c0 < n_c; c0 += NT) {
    const int i = c0 + tid;
    int lo = 0, hi = 0;
    if (i < n_c) {
      const uint64_t x = d.cmx[i];
      lo = seed_bound(qx, n_q, x, false);
      hi = lo < n_q && qx[lo] == x ? seed_bound(qx, n_q, x, true) : lo;
    }
    int total;
    const int ex = seed_excl_sum<NT>(hi - lo, sh, &total);
    if (hi > lo && n_a + ex + (hi - lo) <= SEED_ANCHORS) {
      const uint64_t cy = d.cmy[i];
      const uint64_t ax = (cy & 0x7fffffff00000000ull) | (uint64_t)((uint32_t)cy >> 1);
      for (int j = lo; j < hi; j++) {
        d.ax[n_a + ex + j - lo] = ax;
        d.ay[n_a + ex + j - lo] = (qx[j] & 0xff) << 32 | (uint64_t)((uint32_t)qy[j] >> 1);
      }
    }
    n_a += total;
  }
  __syncthreads();
  if (n_a > SEED_ANCHORS) return fail(SEED_WHY_ANCHORS);
  if (!seed_sort<NT>(d.ax, d.ay, d.bx, d.by, n_a, lds, lds_bytes))
    return fail(SEED_WHY_ANCHOR_SORT);

  // Local chains (mg_lchain_dp, lchain.cpp:158-238, with dp_chaining's arguments,
  // minimizer.cpp:493-500: max_dist = bw, max_skip 25, max_iter 50, min_cnt 3, min_sc 100).
  ChainLds C;
  // six arrays of na ints (the picked anchors reuse v: 2 * SEED_SEGS ints), then the sorts
  const int na = max((n_a + 63) / 64 * 64, 2 * SEED_SEGS);
  if (6 * na * 4 > lds_bytes) return fail(SEED_WHY_ANCHORS);
  C.tp = reinterpret_cast<int*>(lds);
  C.qp = C.tp + na; C.f = C.qp + na; C.p = C.f + na;
  C.t = C.p + na; C.v = C.t + na;
  uint8_t* slds = reinterpret_cast<uint8_t*>(C.v + na);
  const int slds_bytes = lds_bytes - 6 * na * 4;
  for (int i = tid; i < n_a; i += NT) {
    C.tp[i] = (int)(uint32_t)d.bx[i];
    C.qp[i] = (int)(uint32_t)d.by[i];
    C.t[i] = 0;
  }
  __syncthreads();
  if (wave == 0) {
    int st = 0;
    for (int i = 0; i < n_a; i++) {
      const int tpi = C.tp[i], qpi = C.qp[i];
      while (st < i && tpi > C.tp[st] + max_dist) st++;
      if (i - st > 5000) st = i - 5000;
      int max_f = k, max_j = -1, n_skip = 0, n_iter = 0;
      bool stop = false;
      for (int jb = i - 1; jb >= st && !stop; jb -= 64) {
        const int j = jb - lane;
        int sc = 0;
        bool valid = false, hit = false;
        if (j >= st) {
          sc = seedc::chain_link(tpi, qpi, C.tp[j], C.qp[j], k, max_dist, d.pen_gap,
                                 d.pen_skip);
          valid = sc != seedc::CHAIN_NEG;
          if (valid) {
            sc += C.f[j];
            const int pj = C.p[j];
            // minipoa marks t[p[j]] = i for each scored j as it walks and counts a skip at a
            // marked j. Every lane marks before any lane reads: marks come only from larger
            // j (p[j] < j), and one from a j past the walk's stop reaches only anchors below
            // that j, which this i never checks (later i look for their own index).
            if (pj >= 0) C.t[pj] = i;
            hit = C.t[j] == i;
          }
        }
        uint64_t vm = __ballot(valid);
        const uint64_t hm = __ballot(hit);
        while (vm) {
          const int l = __builtin_ctzll(vm);
          vm &= vm - 1;
          const int sl = __builtin_amdgcn_readlane(sc, l);
          if (sl > max_f) {
            max_f = sl;
            max_j = jb - l;
            if (n_skip > 0) --n_skip;
          } else {
            if (sl != max_f && ++n_iter > 50) { stop = true; break; }
            if (((hm >> l) & 1) && ++n_skip > 25) { stop = true; break; }
          }
        }
      }
      if (lane == 0) {
        C.f[i] = max_f;
        C.p[i] = max_j;
      }
    }
  }
  __syncthreads();
  // Chain ends (mg_chain_backtrack, lchain.cpp:66-81): anchors with f >= 100 whose predecessor
  // no later anchor also has (t[-1] reads 0 upstream, so a missing predecessor qualifies).
  for (int i = tid; i < n_a; i += NT) C.v[i] = -1;
  __syncthreads();
  for (int i = tid; i < n_a; i += NT)
    if (C.p[i] >= 0) atomicMax(&C.v[C.p[i]], i);
  __syncthreads();
  int n_z = 0;
  for (int c0 = 0; c0 < n_a; c0 += NT) {
    const int i = c0 + tid;
    const bool z = i < n_a && C.f[i] >= 100 && (C.p[i] < 0 || C.v[C.p[i]] == i);
    int total;
    const int ex = seed_excl_sum<NT>(z ? 1 : 0, sh, &total);
    if (z) {
      d.zx[n_z + ex] = (uint64_t)(int64_t)C.f[i];
      d.zy[n_z + ex] = (uint64_t)i;
    }
    n_z += total;
  }
  __syncthreads();
  if (!seed_sort<NT>(d.zx, d.zy, d.zsx, d.zsy, n_z, slds, slds_bytes))
    return fail(SEED_WHY_END_SORT);
  // Chains from the highest end score down (lchain.cpp:83-133): walk back to the best cut
  // (mg_chain_bk_end), keep chains of >= 3 anchors scoring >= 100.
  const int max_drop = d.bw;
  if (tid == 0) {
    for (int i = 0; i < n_a; i++) C.t[i] = 0;
    int n_v = 0, n_u = 0;
    for (int kk = n_z - 1; kk >= 0; --kk) {
      const int zi = (int)d.zsy[kk], zf = (int)d.zsx[kk];
      if (C.t[zi] != 0) continue;
      const int n_v0 = n_v;
      int end_i = -1, max_i = zi, max_s = 0, i = zi;
      do {
        C.t[i] = 2;
        end_i = i = C.p[i];
        const int s2 = i < 0 ? zf : zf - C.f[i];
        if (s2 > max_s) max_s = s2, max_i = i;
        else if (max_s - s2 > max_drop) break;
      } while (i >= 0 && C.t[i] == 0);
      for (i = zi; i >= 0 && i != end_i; i = C.p[i]) C.t[i] = 0;
      for (i = zi; i != max_i; i = C.p[i]) C.v[n_v++] = i, C.t[i] = 1;
      const int sc = i < 0 ? zf : zf - C.f[i];
      if (sc >= 100 && n_v > n_v0 && n_v - n_v0 >= 3) {
        d.zx[n_u] = (uint64_t)(uint32_t)C.tp[zi] << 32 | (uint32_t)C.qp[zi];
        d.zy[n_u] = (uint64_t)n_v0 << 32 | (uint32_t)n_v;
        n_u++;
      } else {
        n_v = n_v0;
      }
    }
    ctl[0] = n_v;
    ctl[1] = n_u;
  }
  __syncthreads();
  const int n_v = ctl[0], n_u = ctl[1];
  // Chained anchors chain after chain, each ascending (compact_a, lchain.cpp:136-156), into
  // t (tpos) and p (qpos); chains sorted by end (dp_chaining, minimizer.cpp:505).
  for (int c = tid; c < n_u; c += NT) {
    const int s0 = (int)(d.zy[c] >> 32), e0 = (int)(uint32_t)d.zy[c];
    for (int j = e0 - 1, o = s0; j >= s0; j--, o++) {
      C.t[o] = C.tp[C.v[j]];
      C.p[o] = C.qp[C.v[j]];
    }
  }
  __syncthreads();
  if (!seed_sort<NT>(d.zx, d.zy, d.zsx, d.zsy, n_u, slds, slds_bytes))
    return fail(SEED_WHY_CHAIN_SORT);
  // The global chain (chain_dp, lchain.cpp:261-386) and its anchors at least min_w apart,
  // then the sub-windows between them (alignment(), align.cpp:1771-1797). One lane.
  int* ctp = C.t;
  int* cqp = C.p;
  int* f2 = C.f;
  int* cs = C.tp;
  int* pc = C.qp;
  int* pick = C.v;  // picked anchors, back to front: (tpos, qpos) pairs, SEED_SEGS kept
  SegDesc* segs = d.segs + (int64_t)ri * SEED_SEGS;
  int* splice = d.splice + (int64_t)ri * 2 * SEED_SPLICE;
  if (tid == 0) {
    int n_pick = 0, filter_num = 0;
    // a picked anchor; `mark`: fewer than min_w / 10 anchors were passed over since the one
    // picked before it (chain_dp's band-mode flag, bit 40 of y; the sub-windows ignore it)
    auto take = [&](int ct, int cq, bool mark) {
      if (d.pick_out) {
        if (n_pick < SEED_ANCHORS) {
          d.pick_out[2 * n_pick] = (uint64_t)wd.n_reads << 32 | (uint32_t)ct;
          d.pick_out[2 * n_pick + 1] =
              (uint64_t)k << 32 | (uint32_t)cq | (mark ? 1ull << 40 : 0ull);
        }
      } else if (n_pick < SEED_SEGS) {
        pick[2 * n_pick] = ct;
        pick[2 * n_pick + 1] = cq;
      }
      n_pick++;
    };
    if (n_u > 0) {
      int gmax = seedc::CHAIN_NEG, gi = -1;
      for (int i = 0; i < n_u; i++) {
        const uint64_t yi = d.zsy[i];
        const int s_i = (int)(yi >> 32), e_i = (int)(uint32_t)yi;
        const int stp = ctp[s_i], sqp = cqp[s_i];
        f2[s_i] = k;
        for (int j = s_i + 1; j < e_i; j++)
          f2[j] = f2[j - 1] + seedc::chain_link(ctp[j], cqp[j], ctp[j - 1], cqp[j - 1], k,
                                                max_dist, d.pen_gap, d.pen_skip);
        int max_j = -1, max_s = f2[e_i - 1];
        for (int j = i - 1; j >= 0; --j) {
          const uint64_t xj = d.zsx[j], yj = d.zsy[j];
          const int etp = (int)((xj >> 32) & 0x7fffffff), eqp = (int)(uint32_t)xj;
          if (stp < ctp[(int)(yj >> 32)]) continue;
          int s1;
          if (stp > etp && sqp > eqp) {
            s1 = cs[j] + f2[e_i - 1];
          } else {
            // get_local_chain_score (lchain.cpp:249-259)
            int lo = s_i, hi = e_i;
            while (lo + 1 < hi) {
              const int mid = (lo + hi) / 2;
              if (ctp[mid] <= etp || cqp[mid] <= eqp) lo = mid;
              else hi = mid;
            }
            s1 = cs[j] + (hi == e_i ? seedc::CHAIN_NEG : f2[e_i - 1] - f2[hi]);
          }
          if (s1 > max_s) max_s = s1, max_j = j;
        }
        cs[i] = max_s;
        pc[i] = max_j;
        if (max_s > gmax) gmax = max_s, gi = i;
      }
      int min_w = min(20000, (L + 24 - 1) / 24);
      if (d.poa_w) min_w = min(min_w, d.poa_w + k);
      int cur = gi, pre = pc[gi];
      uint64_t cy = d.zsy[cur];
      int last_t = L, last_q = qlen, i = (int)(uint32_t)cy - 1;
      while (pre != -1) {
        const uint64_t px = d.zsx[pre], py = d.zsy[pre];
        const int pet = (int)((px >> 32) & 0x7fffffff), peq = (int)(uint32_t)px;
        while (i >= 0) {
          const int ct = ctp[i], cq = cqp[i];
          if (!(ct > pet && cq > peq)) break;
          if (last_t - ct >= min_w && last_q - cq >= min_w) {
            take(ct, cq, filter_num < min_w / 10);
            filter_num = 0;
            last_t = ct, last_q = cq;
          } else {
            filter_num++;
          }
          i--;
          if (i == (int)(cy >> 32)) break;
        }
        cur = pre, pre = pc[pre], cy = py;
        i = (int)(uint32_t)cy - 1;
      }
      while (i >= 0) {
        const int ct = ctp[i], cq = cqp[i];
        if (last_t - ct >= min_w && last_q - cq >= min_w) {
          take(ct, cq, filter_num < min_w / 10);
          filter_num = 0;
          last_t = ct, last_q = cq;
        } else {
          filter_num++;
        }
        if (i == (int)(cy >> 32)) break;
        i--;
      }
    }
    ctl[3] = n_pick;
    if (d.pick_out) {  // anchors only
      *d.n_pick_out = n_pick;
      ctl[2] = 1;
    } else {
      ctl[2] = 0;
    }
  }
  __syncthreads();
  if (ctl[2]) return true;
  if (tid == 0) {
    const int n_pick = min(ctl[3], SEED_SEGS + 1);
    // Sub-windows: up to each picked anchor's first base, then its k-1 bases spliced in as
    // the consensus nodes they matched, the last one opening the next sub-window.
    int status = MPOA_OK, why = 0, n_seg = 0, n_sp = 0;
    int beg_id = 0, beg_rank = 0, beg_q = 0, j = 0, res_pos = 0;
    int64_t dp_used = 0;
    if (n_pick >= SEED_SEGS) status = MPOA_E_OVERFLOW, why = SEED_WHY_SEGS;  // n_pick + 1
    // A sub-window of length -1 (anchors overlapping by one base) is a DP call without res
    // entries (null_call_score); shorter ones end the window where minipoa reads past its
    // query (MPOA_E_EMPTY_WINDOW), as does one that ends before it begins in ranks.
    auto add_seg = [&](int end_id, int end_rank, int sub_len) {
      if (sub_len < -1 || end_rank - beg_rank + 1 <= 0) {
        status = MPOA_E_EMPTY_WINDOW;
        return;
      }
      if (n_seg >= SEED_SEGS) {
        status = MPOA_E_OVERFLOW, why = SEED_WHY_SEGS;
        return;
      }
      const int rows = end_rank - beg_rank + 1;
      const int64_t need = ((int64_t)rows * ((sub_len + 1) / 8 + 2) * 8 + 63) / 64 * 64;
      SegDesc g{};
      g.beg_id = beg_id;
      g.end_id = end_id;
      g.q0 = j;
      g.qlen = sub_len;
      g.res0 = res_pos;
      g.dp0 = dp_used;
      g.dp_len = need;
      dp_used += 3 * need;
      if (dp_used > d.dp_cap) {
        status = MPOA_E_OVERFLOW, why = SEED_WHY_POOL;
        return;
      }
      segs[n_seg++] = g;
    };
    for (int pi = n_pick - 1; pi >= 0 && status == MPOA_OK; pi--) {
      const int tpos = pick[2 * pi], qpos = pick[2 * pi + 1];
      const int end_t = tpos - k + 1, end_q = qpos - k + 1;
      if (end_t < 0 || end_t + k > L) {
        status = MPOA_E_OVERFLOW, why = SEED_WHY_SPLICE;
        break;
      }
      const int sub_len = end_q - beg_q;
      add_seg(A.cons_id[end_t], A.cons_rank[end_t], sub_len);
      if (status != MPOA_OK) break;
      res_pos += sub_len + 1;
      j += sub_len;
      int last_id = 0, last_rank = 0;
      for (int kk = 0; kk < k; kk++, j++) {
        last_id = A.cons_id[end_t + kk];
        last_rank = A.cons_rank[end_t + kk];
        if (kk != k - 1) {
          if (n_sp >= SEED_SPLICE) {
            status = MPOA_E_OVERFLOW, why = SEED_WHY_SPLICE;
            break;
          }
          splice[2 * n_sp] = res_pos++;
          splice[2 * n_sp + 1] = last_id;
          n_sp++;
        }
      }
      beg_id = last_id;
      beg_rank = last_rank;
      beg_q = j;
    }
    if (status == MPOA_OK) add_seg(1, S.n_nodes - 1, qlen - beg_q);
    if (status != MPOA_OK) {
      n_seg = 0;
      n_sp = 0;
      S.status = status;
      if (status == MPOA_E_OVERFLOW) atomicOr(d.flag, why);
    }
    d.n_seg[ri] = n_seg;
    d.n_splice[ri] = n_sp;
  }
  __syncthreads();
  return S.status == MPOA_OK;
}
