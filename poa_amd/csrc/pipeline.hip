// Device pipeline: every step of minipoa's per-read loop (src/main.cpp:110-121) for a set of
// independent windows, one workgroup per window, windows pulled from a device work queue by
// a persistent grid.
//
//   per read:  DP plan (band, row offsets)  -> banded DP rows (all lanes)
//              -> backtrack (one lane)       -> add_path (one lane)
//              -> topological re-sort + hlen/tlen + next read's DP plan (one lane)
//   per window: heaviest-bundle consensus or row-column MSA (one lane + all lanes)
//
// Arithmetic is minipoa's int32 (NEG_INF = 0xc0c0c0c0, no saturation): every score, path,
// node, edge, MSA column and consensus base equals minipoa's on the same input.
//
// Layout in HBM (per slot = one resident workgroup; reused window after window):
//   node arrays (SoA, node id), edge records (in- and out-lists, singly linked in insertion
//   order), rank-ordered DP plan (pred CSR, band, row offsets), DP matrices H/D/I with ragged
//   rows of (Be-Bs)*8 int32 at row_off[row].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "pipeline.h"
#include "seed_core.h"

namespace mpoa {

constexpr int Q = 8;                      // band quantum (src/align.cpp:13,19)
constexpr int NEG_INF = (int)0xc0c0c0c0;  // src/align.cpp:22
constexpr int CODE_N = 4;
constexpr int M_OP = 1, D_OP = 2, I_OP = 4, ALL_OP = 7;
constexpr int SCAN_ID = INT32_MIN;        // identity of the prefix max

__device__ __forceinline__ int code_of(uint8_t c) {
  switch (c) {
    case 0: case 1: case 2: case 3: return c;
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return CODE_N;
  }
}
__device__ __forceinline__ uint8_t char_of(int code) {
  return (uint8_t)(code == 0 ? 'A' : code == 1 ? 'C' : code == 2 ? 'G' : code == 3 ? 'T' : 'N');
}

// Window-level scalar state shared by the workgroup.
struct WinShared {
  int n_nodes, n_in_e, n_out_e;
  int status;
  int read;       // current read index (seq_id)
  int qlen, m;
  int n_rows;     // rows of the current DP (whole graph: n_nodes)
  int beg_i, end_i;
  int score;
  int cons_len;
  int msa_cols;
  int64_t cells, edge_cells;
  int64_t call_cells, call_ecells;  // of the last DP call
  int64_t call_logical;             // its logical cells, sum over rows 1..n-1 of Me-Ms+1
  int64_t logical;                  // the window's
  int call_width;                   // widest row of the last DP call
  int bt_rc;                        // backtrack result of the last DP call
  bool bt_serial;                   // it ran on one lane
  int kahn_over;                    // the LDS sort's stack would have overflowed
  int64_t total;  // DP storage of the current call
  int job;
  int scan[16];
  int row_bs, row_be;
  int row_wmax;   // widest row of the current static-band plan (8*block_num)
  uint32_t fmin;  // int16 form: smallest finite H of the call, less FINITE_CUT
  bool narrow_bad;       // the call's int16 form did not hold its values: re-run in int32
  int64_t narrow_cells;  // band cells of the window's calls in the int16 form
  bool q_lds;
  int q_cap;
  int64_t ticks[PH_N];
};

__device__ __forceinline__ int64_t now() { return (int64_t)wall_clock64(); }

// Waits for this wave's outstanding HBM loads (and stores) only: s_waitcnt vmcnt(0) with the
// export and LDS counters left at their maxima. Ends every branch that reads HBM inside a DP
// row, so that at the branch's join no load is pending and the compiler's wait insertion has
// no reason to wait there, on the LDS-only path, for the previous row's stores.
__device__ __forceinline__ void drain_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---------------------------------------------------------------------------------------
// Graph edits (one lane). Follow graph::add_adj / node_t::add_in_adj / add_out_adj
// (src/graph.cpp:13-20, include/node.h:23-61) and graph::add_path (graph.cpp:170-235).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int new_node(Arena& A, WinShared& S, const SlotCaps& c, int base) {
  int id = S.n_nodes;
  if (id >= c.nodes) { S.status = MPOA_E_OVERFLOW; return -1; }
  S.n_nodes = id + 1;
  A.base[id] = (uint8_t)base;
  A.par[id] = id;
  for (int k = 0; k < NB; k++) A.aligned[id * NB + k] = -1;
  A.aligned[id * NB + base] = id;
  A.ind[id] = 0;
  A.in_head[id] = A.in_tail[id] = A.out_head[id] = A.out_tail[id] = -1;
  A.in_cnt[id] = 0;
  A.out_cnt[id] = 0;
  A.last_read[id] = -1;
  return id;
}

__device__ __forceinline__ void add_adj(Arena& A, WinShared& S, const SlotCaps& c, bool keep_path, int seq_id,
                        int from, int to, int pos, int path0) {
  if (from == to) return;
  // out-list of `from`
  int e;
  for (e = A.out_head[from]; e != -1; e = A.oe_next[e])
    if (A.oe_dst[e] == to) break;
  if (e != -1) {
    A.oe_w[e]++;
  } else {
    e = S.n_out_e;
    if (e >= c.edges) { S.status = MPOA_E_OVERFLOW; return; }
    S.n_out_e = e + 1;
    A.oe_dst[e] = to; A.oe_w[e] = 1; A.oe_next[e] = -1;
    if (A.out_tail[from] == -1) A.out_head[from] = e; else A.oe_next[A.out_tail[from]] = e;
    A.out_tail[from] = e;
    A.out_cnt[from]++;
  }
  // in-list of `to`
  A.ind[to]++;
  for (e = A.in_head[to]; e != -1; e = A.ie_next[e])
    if (A.ie_src[e] == from) break;
  if (e != -1) {
    A.ie_w[e]++;
  } else {
    e = S.n_in_e;
    if (e >= c.edges) { S.status = MPOA_E_OVERFLOW; return; }
    S.n_in_e = e + 1;
    A.ie_src[e] = from; A.ie_w[e] = 1; A.ie_next[e] = -1;
    if (A.in_tail[to] == -1) A.in_head[to] = e; else A.ie_next[A.in_tail[to]] = e;
    A.in_tail[to] = e;
    A.in_cnt[to]++;
  }
  // ids/idp: first visit of this read records the node at read position pos.
  if (A.last_read[to] != seq_id) {
    A.last_read[to] = seq_id;
    if (keep_path && to != 1 && pos >= 0) A.path[path0 + pos] = to;
  }
}

__device__ __forceinline__ void add_path(Arena& A, WinShared& S, const SlotCaps& c, bool keep_path, int seq_id,
                         int m, int path0) {
  int anchor = -1, pos = 0;
  for (int r = 0; r < m && S.status == MPOA_OK; r++) {
    int cur = A.res_from[r];
    const int b = A.res_base[r];
    if (cur < 0) {
      const int al = A.res_aligned[r];
      if (al >= 0) {
        int ex = A.aligned[al * NB + b];
        if (ex == -1) {
          cur = new_node(A, S, c, b);
          if (cur < 0) return;
          A.aligned[al * NB + b] = cur;
          A.par[cur] = al;
        } else {
          cur = ex;
        }
      } else {
        cur = new_node(A, S, c, b);
        if (cur < 0) return;
      }
    }
    if (anchor != -1) add_adj(A, S, c, keep_path, seq_id, anchor, cur, pos++, path0);
    anchor = cur;
  }
  add_adj(A, S, c, keep_path, seq_id, anchor, 1, pos, path0);
}

// add_path over all lanes. When a read's path visits each node at most once (a run of the
// same node, X -> X, is one visit: add_adj skips self-loops), every in-list and out-list
// touched by one read belongs to a distinct node and the edits are independent; new nodes take
// ids in res order (a prefix sum), exactly as the serial walk assigns them. A read whose res
// reuses an existing node at two positions apart, or would create the same aligned node twice,
// falls back to the serial walk. Returns false when that fallback is needed (nothing has been
// modified then). stamp: N ints of scratch.
template <int NT>
__device__ __forceinline__ int block_excl_sum(int v, WinShared& S, int* total);

template <int NT>
__device__ __forceinline__ bool add_path_parallel(Arena& A, WinShared& S, const SlotCaps& c, bool keep_path,
                                  int seq_id, int m, int path0, int* cur, int* stamp) {
  bool dup = false;
  // 0. existing node of each res entry (its own id, or the aligned node it reuses), stamped
  //    with the position of the run it opens; a second run of the same node finds a stamp
  {
    auto reused = [&](int r) {
      const int f = A.res_from[r];
      if (f >= 0) return f;
      const int al = A.res_aligned[r];
      return al >= 0 ? A.aligned[al * NB + A.res_base[r]] : -1;
    };
    for (int r = threadIdx.x; r < m; r += NT) {
      const int v = reused(r);
      if (v >= 0 && (r == 0 || reused(r - 1) != v)) stamp[v] = -1;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < m; r += NT) {
      const int v = reused(r);
      if (v >= 0 && (r == 0 || reused(r - 1) != v) && atomicExch(&stamp[v], r) != -1) dup = true;
    }
    if (__syncthreads_or(dup)) return false;
  }
  // 1. claim empty aligned slots (key al*NB+b) for the earliest res entry: atomicMin of
  //    INT32_MIN + r (every claim is below -1, the empty mark; node ids are >= 0)
  for (int r = threadIdx.x; r < m; r += NT) {
    if (A.res_from[r] < 0 && A.res_aligned[r] >= 0) {
      const int key = A.res_aligned[r] * NB + A.res_base[r];
      if (A.aligned[key] < 0) atomicMin(&A.aligned[key], INT32_MIN + r);
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < m; r += NT) {
    if (A.res_from[r] < 0 && A.res_aligned[r] >= 0) {
      const int key = A.res_aligned[r] * NB + A.res_base[r];
      const int v = A.aligned[key];
      if (v < -1 && v != INT32_MIN + r) dup = true;
    }
  }
  if (__syncthreads_or(dup)) {
    for (int r = threadIdx.x; r < m; r += NT) {  // undo the claims
      if (A.res_from[r] < 0 && A.res_aligned[r] >= 0) {
        const int key = A.res_aligned[r] * NB + A.res_base[r];
        if (A.aligned[key] < -1) A.aligned[key] = -1;
      }
    }
    __syncthreads();
    return false;
  }
  // 2. node of every res entry; new ids by prefix sum in res order
  const int n0 = S.n_nodes;
  int base = 0;
  for (int r0 = 0; r0 < m; r0 += NT) {
    const int r = r0 + threadIdx.x;
    int is_new = 0, v = -1;
    if (r < m) {
      v = A.res_from[r];
      if (v < 0) {
        const int al = A.res_aligned[r];
        if (al >= 0) {
          const int ex = A.aligned[al * NB + A.res_base[r]];
          if (ex >= 0) v = ex;
          else is_new = 1;
        } else {
          is_new = 1;
        }
      }
    }
    int tot;
    const int ex = block_excl_sum<NT>(is_new, S, &tot);
    if (r < m) {
      if (is_new) {
        v = n0 + base + ex;
        if (v < c.nodes) {
          const int b = A.res_base[r];
          A.base[v] = (uint8_t)b;
          for (int k = 0; k < NB; k++) A.aligned[v * NB + k] = -1;
          A.aligned[v * NB + b] = v;
          A.ind[v] = 0;
          A.in_head[v] = A.in_tail[v] = A.out_head[v] = A.out_tail[v] = -1;
          A.in_cnt[v] = 0;
          A.out_cnt[v] = 0;
          A.last_read[v] = -1;
          const int al = A.res_aligned[r];
          if (A.res_from[r] < 0 && al >= 0) {
            A.aligned[al * NB + b] = v;
            A.par[v] = al;
          } else {
            A.par[v] = v;
          }
        }
      }
      cur[r] = v;
    }
    base += tot;
  }
  if (n0 + base > c.nodes) {
    if (threadIdx.x == 0) S.status = MPOA_E_OVERFLOW;
    __syncthreads();
    return true;
  }
  __syncthreads();
  if (threadIdx.x == 0) S.n_nodes = n0 + base;
  // 3. edges cur[r-1] -> cur[r] (r = 1..m-1) and cur[m-1] -> sink, at read position r-1
  for (int r = 1 + threadIdx.x; r <= m; r += NT) {
    const int from = cur[r - 1], to = r < m ? cur[r] : 1, pos = r - 1;
    if (from == to) continue;
    int e;
    for (e = A.out_head[from]; e != -1; e = A.oe_next[e])
      if (A.oe_dst[e] == to) break;
    if (e != -1) {
      A.oe_w[e]++;
    } else {
      e = atomicAdd(&S.n_out_e, 1);
      if (e >= c.edges) { S.status = MPOA_E_OVERFLOW; continue; }
      A.oe_dst[e] = to; A.oe_w[e] = 1; A.oe_next[e] = -1;
      if (A.out_tail[from] == -1) A.out_head[from] = e; else A.oe_next[A.out_tail[from]] = e;
      A.out_tail[from] = e;
      A.out_cnt[from]++;
    }
    A.ind[to]++;
    for (e = A.in_head[to]; e != -1; e = A.ie_next[e])
      if (A.ie_src[e] == from) break;
    if (e != -1) {
      A.ie_w[e]++;
    } else {
      e = atomicAdd(&S.n_in_e, 1);
      if (e >= c.edges) { S.status = MPOA_E_OVERFLOW; continue; }
      A.ie_src[e] = from; A.ie_w[e] = 1; A.ie_next[e] = -1;
      if (A.in_tail[to] == -1) A.in_head[to] = e; else A.ie_next[A.in_tail[to]] = e;
      A.in_tail[to] = e;
      A.in_cnt[to]++;
    }
    if (A.last_read[to] != seq_id) {
      A.last_read[to] = seq_id;
      if (keep_path && to != 1) A.path[path0 + pos] = to;
    }
  }
  __syncthreads();
  return true;
}

// graph::topsort op=0 (src/graph.cpp:59-123), split into:
//   1. the LIFO Kahn order (stack seeded with the zero in-degree nodes in ascending id
//      order) by one wavefront over an out-edge CSR built in parallel: in LDS when it fits,
//      else in the slot's HBM scratch (kahn_walk);
//   2. in parallel over ranks: rank arrays, the DP plan (predecessor CSR in node.in order with
//      rank, weight and base; per-row records), and the parents of the hlen/tlen trees (the
//      max-in-weight predecessor / max-out-weight successor, first wins);
//   3. hlen = 1 + hlen(parent) (source 0) and tlen = 1 + tlen(child) (sink 1) as depths in
//      those trees, by blocked list ranking (tree_depth_blocked).

// Block-wide exclusive prefix sum of v (NT threads); *total gets the sum over all threads.
template <int NT>
__device__ __forceinline__ int block_excl_sum(int v, WinShared& S, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const int u = __shfl_up(incl, s, 64);
    if (lane >= s) incl += u;
  }
  if (lane == 63) S.scan[wave] = incl;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w2 = 0; w2 < NT / 64; w2++) {
    const int t = S.scan[w2];
    if (w2 < wave) before += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// Stack capacity of the wave walks: the stack holds the released nodes not yet popped, fewer
// than ten in these alignment graphs; a walk that would exceed it starts over with one lane.
constexpr int KAHN_STK = 512;

// The LIFO Kahn walk of graph::topsort op=0 (src/graph.cpp:59-123) into A.rank_id / A.rank,
// over an out-edge CSR (ooff: N + 1 offsets, odst: targets in out-list order), in-degrees deg
// and a stack of KAHN_STK entries. Three storage forms share it: the LDS form (int16 CSR, byte
// degrees); for graphs whose CSR does not fit the tile's LDS, the form with byte degrees in LDS
// and the CSR in HBM as records (REC: each out-edge's target with the target's own out-edge
// range, one load per edge); and the HBM form (int32 CSR in the slot's scratch arrays, degrees
// decremented with atomics so that every access to a counter is ordered at L2; the stack stays
// in LDS). Returns false when a degree exceeds deg_max or the stack would overflow (the caller
// then walks with one lane).
template <int NT, bool HBM, bool REC, class Off, class Deg, class Stk>
__device__ __forceinline__ bool kahn_walk(Arena& A, WinShared& S, Off* ooff, Off* odst, int4* orec,
                                          Deg* deg, Stk* stk, int deg_max) {
  const int N = S.n_nodes;
  // out-degree prefix sums (node id order)
  int base = 0, wide = 0;
  for (int v0 = 0; v0 < N; v0 += NT) {
    const int v = v0 + threadIdx.x;
    const int oc = v < N ? A.out_cnt[v] : 0;
    int tot;
    const int ex = block_excl_sum<NT>(oc, S, &tot);
    if (v < N) {
      const int ic = A.in_cnt[v];
      wide |= ic > deg_max;
      ooff[v] = (Off)(base + ex);
      deg[v] = (Deg)ic;
    }
    base += tot;
  }
  if (threadIdx.x == 0) ooff[N] = (Off)base;
  if (__syncthreads_or(wide)) return false;
  for (int v = threadIdx.x; v < N; v += NT) {
    int k = ooff[v];
    for (int e = A.out_head[v]; e != -1; e = A.oe_next[e]) {
      const int d = A.oe_dst[e];
      if constexpr (REC) orec[k++] = make_int4(d, ooff[d], ooff[d + 1], 0);
      else odst[k++] = (Off)d;
    }
  }
  __syncthreads();
  // The walk, by wave 0. Zero in-degree nodes go on the stack in ascending id order; each
  // popped node's out-edges are processed by one lane each, in out-list order: the degree
  // decrements (distinct targets: no two lanes touch one counter), then the targets that
  // reached zero are pushed in edge order. The last one pushed is the next node popped, so it
  // is kept in a register, with its out-edge range loaded alongside the decrements.
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) {
    const int lane = threadIdx.x & 63;
    int sp = 0;
    bool over = false;
    for (int v0 = 0; v0 < N && !over; v0 += 64) {
      const int v = v0 + lane;
      const bool z = v < N && deg[v] == 0;
      const uint64_t mk = __ballot(z);
      over = sp + __popcll(mk) > KAHN_STK;
      if (z && !over) stk[sp + __popcll(mk & ((1ULL << lane) - 1))] = (Stk)v;
      sp += __popcll(mk);
    }
    // Ranks leave in blocks of 64 (lane k holds the node of rank r - (r & 63) + k): two vector
    // stores per 64 pops instead of two single-lane stores per pop.
    int r = 0, buf_u = 0;
    int next = -1, next_o0 = 0, next_o1 = 0;
    while (!over) {
      int u, o0, o1;
      if (next >= 0) {
        u = next;
        o0 = next_o0;
        o1 = next_o1;
      } else {
        if (sp == 0) break;
        u = stk[--sp];
        o0 = ooff[u];
        o1 = ooff[u + 1];
      }
      if (lane == (r & 63)) buf_u = u;
      if ((r & 63) == 63) {
        A.rank_id[r - 63 + lane] = buf_u;
        A.rank[buf_u] = r - 63 + lane;
      }
      r++;
      const int cnt = o1 - o0;
      next = -1;
      for (int b0 = 0; b0 < cnt; b0 += 64) {
        const int k = b0 + lane;
        int v = 0, vo0 = 0, vo1 = 0;
        bool z = false;
        if (k < cnt) {
          if constexpr (REC) {
            const int4 e = orec[o0 + k];
            v = e.x;
            vo0 = e.y;
            vo1 = e.z;
          } else {
            // the target's edge range first: read before the degree store in program order,
            // its loads go out with the degree's instead of after the store
            v = odst[o0 + k];
            vo0 = ooff[v];
            vo1 = ooff[v + 1];
          }
          int d;
          if constexpr (HBM) {
            d = atomicSub(&deg[v], 1) - 1;
          } else {
            d = deg[v] - 1;
            deg[v] = (Deg)d;
          }
          z = d == 0;
        }
        const uint64_t mk = __ballot(z);
        if (sp + __popcll(mk) > KAHN_STK) {
          over = true;
          break;
        }
        if (mk) {
          // an earlier pending next goes under the new pushes
          if (next >= 0 && lane == 0) stk[sp] = (Stk)next;
          if (next >= 0) sp++;
          const int last = 63 - __clzll(mk);
          const uint64_t rest = mk & ~(1ULL << last);
          if (z && lane != last) stk[sp + __popcll(rest & ((1ULL << lane) - 1))] = (Stk)v;
          sp += __popcll(rest);
          next = __builtin_amdgcn_readlane(v, last);
          next_o0 = __builtin_amdgcn_readlane(vo0, last);
          next_o1 = __builtin_amdgcn_readlane(vo1, last);
        }
      }
    }
    if (lane < (r & 63)) {
      A.rank_id[r - (r & 63) + lane] = buf_u;
      A.rank[buf_u] = r - (r & 63) + lane;
    }
    if (lane == 0) {
      S.n_rows = r;
      S.kahn_over = over;
    }
  }
  __syncthreads();
  return !S.kahn_over;
}

// The walk with the graph in LDS (lds_bytes available at lds): int16 CSR, byte in-degrees, a
// KAHN_STK-entry int16 stack. False when it does not fit, a degree exceeds 255 or the stack
// would overflow.
template <int NT>
__device__ __forceinline__ bool kahn_lds(Arena& A, WinShared& S, uint8_t* lds, int lds_bytes) {
  const int N = S.n_nodes, E = S.n_out_e;
  if (N > 32767 || E > 32767 || 2LL * (N + 1 + E + KAHN_STK) + N > lds_bytes) return false;
  int16_t* ooff = reinterpret_cast<int16_t*>(lds);  // N + 1
  int16_t* odst = ooff + N + 1;
  int16_t* stk = odst + E;
  uint8_t* deg = reinterpret_cast<uint8_t*>(stk + KAHN_STK);
  return kahn_walk<NT, false, false>(A, S, ooff, odst, nullptr, deg, stk, 255);
}

// The walk with byte in-degrees and the stack in LDS, the CSR in HBM: offsets in pred_off, and
// for each out-edge a record of its target and the target's own out-edge range in the row
// records' space (free until the DP plan is filled), so that one load per edge gives the next
// node and where its edges are. False when it does not fit, a degree exceeds 255 or the stack
// would overflow.
template <int NT>
__device__ __forceinline__ bool kahn_mid(Arena& A, WinShared& S, const SlotCaps& c, uint8_t* lds,
                                         int lds_bytes) {
  const int N = S.n_nodes, E = S.n_out_e;
  if ((int64_t)sizeof(int) * KAHN_STK + N > lds_bytes || 4LL * E > 12LL * c.nodes) return false;
  int* stk = reinterpret_cast<int*>(lds);
  uint8_t* deg = lds + sizeof(int) * KAHN_STK;
  return kahn_walk<NT, false, true>(A, S, A.pred_off, A.pred_rank,
                                    reinterpret_cast<int4*>(A.rowrec), deg, stk, 255);
}

// The walk with the graph in HBM: the CSR in pred_off (N + 1) and pred_rank (E), the degrees in
// t0, all three free until the DP plan is filled after the sort; the stack in LDS.
template <int NT>
__device__ __forceinline__ bool kahn_hbm(Arena& A, WinShared& S, uint8_t* lds, int lds_bytes) {
  if ((int)sizeof(int) * KAHN_STK > lds_bytes) return false;
  return kahn_walk<NT, true, false>(A, S, A.pred_off, A.pred_rank, nullptr, A.t0,
                                    reinterpret_cast<int*>(lds), 0x7fffffff);
}

// Kahn order on the HBM lists (one lane), for a graph whose wave walk would pass KAHN_STK
// stack entries.
__device__ __forceinline__ void kahn_global(Arena& A, WinShared& S) {
  const int N = S.n_nodes;
  int* deg = A.t0;
  int* stk = A.t1;
  int sp = 0;
  for (int v = 0; v < N; v++) {
    deg[v] = A.in_cnt[v];
    if (deg[v] == 0) stk[sp++] = v;
  }
  int r = 0;
  while (sp > 0) {
    const int u = stk[--sp];
    A.rank[u] = r;
    A.rank_id[r] = u;
    r++;
    for (int e = A.out_head[u]; e != -1; e = A.oe_next[e]) {
      const int v = A.oe_dst[e];
      if (--deg[v] == 0) stk[sp++] = v;
    }
  }
  S.n_rows = r;
}

// Depth of every rank in a parent forest given as par[r] (-1 at roots; UP: parents have smaller
// ranks), out[r] = base + the number of links from r to its root, by blocked list ranking.
// Blocks of 64 ranks in parent-first order: a lane whose parent lies in an earlier
// block reads that parent's depth; links to parents inside the block are resolved by pointer
// jumping across the lanes (Wyllie's list ranking: at most six steps for 63 links). dep holds
// the depths the later blocks read; out gets base + depth. Wave 0 only. In LDS (int16 arrays)
// operations of one wavefront complete in issue order, so the next block sees every store; in
// HBM (int32 arrays) each block's stores are waited for and the depths are read with
// device-coherent loads. Not inlined: inside the window kernel it would push the register
// allocation into scratch.
template <bool UP, bool HBM, class T>
__device__ __noinline__ void tree_depth_blocked(int n, const T* par, T* dep, int* out, int base) {
  const int lane = threadIdx.x & 63;
  for (int b = 0; b < n; b += 64) {
    const int hi = n - 1 - b;  // highest rank of the block when parents have larger ranks
    const int r = UP ? b + lane : hi - lane;
    const bool act = b + lane < n;
    const int p = act ? par[r] : -1;
    int d = 0, link = lane, done = 1;
    if (p >= 0) {
      if (UP ? p >= b : p <= hi) {
        d = 1;
        link = UP ? p - b : hi - p;
        done = 0;
      } else if constexpr (HBM) {
        d = __hip_atomic_load(dep + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
      } else {
        d = dep[p] + 1;
      }
    }
    for (int step = 0; step < 7 && !__all(done); step++) {
      const int dl = __shfl(d, link, 64);
      const int donel = __shfl(done, link, 64);
      const int ll = __shfl(link, link, 64);
      if (!done) {
        d += dl;
        if (donel) done = 1; else link = ll;
      }
    }
    if (act) {
      dep[r] = (T)d;
      out[r] = base + d;
    }
    if constexpr (HBM) drain_vmem();
  }
}

template <int NT>
__device__ __forceinline__ void topsort0(Arena& A, WinShared& S, const SlotCaps& c, bool with_len, uint8_t* lds,
                         int lds_bytes, int forms) {
  const int64_t tk = now();
  const bool lds_forms = (forms & (MPOA_FORM_SORT_HBM | MPOA_FORM_SORT_LANE)) == 0;
  if ((forms & MPOA_FORM_SORT_LANE) ||
      (!(lds_forms && (kahn_lds<NT>(A, S, lds, lds_bytes) || kahn_mid<NT>(A, S, c, lds, lds_bytes))) &&
       !kahn_hbm<NT>(A, S, lds, lds_bytes))) {
    if (threadIdx.x == 0) kahn_global(A, S);
    __syncthreads();
  }
  if (threadIdx.x == 0) S.ticks[PH_KAHN] += now() - tk;
  const int n = S.n_rows;
  // DP plan in rank order: predecessor CSR offsets (block scan of in-degrees), then each lane
  // walks its rank's in-list.
  int base = 0;
  for (int r0 = 0; r0 < n; r0 += NT) {
    const int r = r0 + threadIdx.x;
    const int ic = r < n ? A.in_cnt[A.rank_id[r]] : 0;
    int tot;
    const int ex = block_excl_sum<NT>(ic, S, &tot);
    if (r < n) A.pred_off[r] = base + ex;
    base += tot;
  }
  if (threadIdx.x == 0) A.pred_off[n] = base;
  __syncthreads();
  int* hpar = A.t0;  // max-in-weight predecessor rank (hlen tree)
  int* tchd = A.t1;  // max-out-weight successor rank (tlen tree)
  // the two trees (and the depths) in LDS when they fit; the sort's LDS arrays are dead here
  const bool lds_len = with_len && n <= 32767 && 6LL * n <= lds_bytes;
  int16_t* lpar = reinterpret_cast<int16_t*>(lds);
  int16_t* lchd = lpar + n;
  int16_t* ldep = lchd + n;
  for (int r = threadIdx.x; r < n; r += NT) {
    const int u = A.rank_id[r];
    const int k0 = A.pred_off[r];
    const int ind10 = A.ind[u] / 10;
    int k = k0, wmax = -1, max_pre = -1, pr3[3] = {0, 0, 0}, w3[3] = {0, 0, 0}, pk = 0;
    int omax = -1, max_suc = -1, last_suc = -1;
    // the in-list and the out-list are walked in one loop, so that their loads overlap
    int ei = A.in_head[u], eo = A.out_head[u];
    while (ei != -1 || eo != -1) {
      if (ei != -1) {
        const int src = A.ie_src[ei], w = A.ie_w[ei], pr = A.rank[src];
        const int b = A.base[src];
        ei = A.ie_next[ei];
        A.pred_rank[k] = pr;
        A.pred_w[k] = w;
        A.pred_base[k] = (uint8_t)(b | (w >= ind10 ? 0x80 : 0));
        if (k - k0 < 3) {
          pr3[k - k0] = pr;
          w3[k - k0] = w;
          pk |= b << (3 * (k - k0));
          if (w >= ind10) pk |= 1 << (9 + (k - k0));
        }
        k++;
        if (w > wmax) { wmax = w; max_pre = pr; }
      }
      if (eo != -1) {
        const int sr = A.rank[A.oe_dst[eo]], ow = A.oe_w[eo];
        eo = A.oe_next[eo];
        if (ow > omax) { omax = ow; max_suc = sr; }
        last_suc = max(last_suc, sr);
      }
    }
    int* rec = A.rowrec + 12LL * r;
    rec[2] = k0;
    rec[3] = k - k0;
    reinterpret_cast<int4*>(rec)[1] = make_int4(pr3[0], pr3[1], pr3[2], pk);
    rec[8] = w3[0];
    rec[9] = w3[1];
    rec[10] = w3[2];
    rec[11] = last_suc;
    if (lds_len) {
      lpar[r] = (int16_t)max_pre;
      lchd[r] = (int16_t)max_suc;
    } else {
      hpar[r] = max_pre;
      tchd[r] = max_suc;
    }
  }
  __syncthreads();
  if (with_len) {
    const int64_t td = now();
    if (lds_len) {
      if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) {
        tree_depth_blocked<true, false>(n, lpar, ldep, A.hlen, 0);
        tree_depth_blocked<false, false>(n, lchd, ldep, A.tlen, 1);
      }
      __syncthreads();
    } else {
      if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) {
        tree_depth_blocked<true, true>(n, hpar, A.t2, A.hlen, 0);
        tree_depth_blocked<false, true>(n, tchd, A.t2, A.tlen, 1);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) S.ticks[PH_DEPTH] += now() - td;
  }
}

// graph::topsort op=1 (src/graph.cpp:24-57): LIFO order over aligned-node groups; members
// inherit their column head's rank. Returns the number of groups.
__device__ __forceinline__ int topsort1(Arena& A, WinShared& S) {
  const int N = S.n_nodes;
  int* deg = A.t0;
  int* stk = A.t1;
  int sp = 0;
  for (int v = 0; v < N; v++) {
    A.rank[v] = -1;
    if (A.par[v] != v) continue;
    int sum = 0;
    for (int k = 0; k < NB; k++) {
      int s = A.aligned[v * NB + k];
      if (s >= 0) sum += A.in_cnt[s];
    }
    deg[v] = sum;
    if (sum == 0) stk[sp++] = v;
  }
  int r = 0;
  while (sp > 0) {
    int u = stk[--sp];
    A.rank[u] = r;
    for (int k = 0; k < NB; k++) {
      int s = A.aligned[u * NB + k];
      if (s < 0) continue;
      A.rank[s] = r;
      for (int e = A.out_head[s]; e != -1; e = A.oe_next[e]) {
        int v = A.par[A.oe_dst[e]];
        if (--deg[v] == 0) stk[sp++] = v;
      }
    }
    r++;
  }
  return r;
}

// ---------------------------------------------------------------------------------------
// Banded DP (src/align.cpp:1357-1514).
// ---------------------------------------------------------------------------------------

// Per-row DP record (12 int32, one per rank), written by the topological sort (predecessors)
// and the band plan (band, storage offsets); read once per row.
//   [0] row_off: offset of the row in the DP matrices
//   [1] Bs | Be << 16      [2] k0: first predecessor in the predecessor CSR   [3] n_pred
//   [4..6] ranks of the first three predecessors
//   [7] bits 0-8: their bases (3 bits each); bits 9-11: their in-weight >= ind/10 (the
//       backtrack's match rule, src/align.cpp:1609)
//   [8..10] in-weights of the first three predecessors
//   [11] the largest rank among the successors (-1: none): the last row that reads this one
struct RowRec {
  int row_off, bsbe, k0, np, p0, p1, p2, pk, w0, w1, w2, last;
};
__device__ __forceinline__ RowRec rowrec_of(const int4& a, const int4& b, const int4& c) {
  return RowRec{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
}
__device__ __forceinline__ RowRec load_rowrec(const int* rr, int r) {
  const int4* v = reinterpret_cast<const int4*>(rr + 12LL * r);
  return rowrec_of(v[0], v[1], v[2]);
}

// LDS ring of the H and D bands of the last RING rows (row r in slot r % RING, column c of the
// band at offset c - 8*Bs[r]). Rows wider than WR are not cached.
template <int RING, int WR, int QCAP>
struct Ring {
  int H[RING * WR];
  int D[RING * WR];
  int row[RING];
  int bsbe[RING];
  uint8_t q[QCAP];         // shifted query codes when the read fits
  // Row records of the current 32-row chunk, with the fourth to sixth predecessors as
  // (local rank << 3) | base in place of the in-weights (only the backtrack reads those, from
  // the records in HBM).
  int4 rec[3 * 32];
  static constexpr int kRing = RING, kWR = WR, kQ = QCAP;
};

// Orders this workgroup's LDS accesses across a row boundary (no wait on global stores).
template <int NT>
__device__ __forceinline__ void lds_barrier() {
  if (NT == 64) {
    __builtin_amdgcn_wave_barrier();
  } else {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }
}

// Inclusive prefix max across a wavefront with DPP row shifts and row broadcasts.
__device__ __forceinline__ int wave_incl_max(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(SCAN_ID, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = max(v, __builtin_amdgcn_update_dpp(SCAN_ID, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = max(v, __builtin_amdgcn_update_dpp(SCAN_ID, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = max(v, __builtin_amdgcn_update_dpp(SCAN_ID, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = max(v, __builtin_amdgcn_update_dpp(SCAN_ID, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, __builtin_amdgcn_update_dpp(SCAN_ID, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

// Inclusive prefix min across a wavefront (the same DPP steps; identity INT32_MAX).
__device__ __forceinline__ int wave_incl_min(int v) {
  constexpr int ID = 0x7fffffff;
  v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

// Exclusive prefix max of v over the workgroup's threads, and the maximum over all of them.
// The wave partials alternate between two halves of S.scan (ph, flipped by the caller on
// every call): a call's reads of one half are then ordered before the next write to it by the
// barrier of the call in between, so one barrier per call suffices.
template <int NT>
__device__ __forceinline__ int block_excl_max(int v, WinShared& S, int* total, int ph) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_incl_max(v);
  int excl = __builtin_amdgcn_update_dpp(SCAN_ID, incl, 0x138, 0xf, 0xf, false);  // wave_shr:1
  if (lane == 0) excl = SCAN_ID;
  const int wtot = __builtin_amdgcn_readlane(incl, 63);
  if (NT == 64) {
    *total = wtot;
    return excl;
  }
  int* part = S.scan + 8 * ph;
  if (lane == 0) part[wave] = wtot;
  lds_barrier<NT>();
  int before = SCAN_ID, all = SCAN_ID;
#pragma unroll
  for (int w2 = 0; w2 < NT / 64; w2++) {
    const int t = part[w2];
    if (w2 < wave) before = max(before, t);
    all = max(all, t);
  }
  *total = all;
  return max(excl, before);
}

// Score storage. A DP call keeps H, D and I either as int32 (minipoa's own width) or, when
// every finite value of the call fits, as int16 ("narrow"): finite values are stored as they
// are and every value derived from NEG_INF is stored as -32768, which reads back as NEG_INF.
// In registers the arithmetic stays int32, so a finite cell gets exactly minipoa's value and a
// cell derived from NEG_INF stays far below every finite one (minipoa's NEG_INF offsets differ,
// which no comparison between a finite and a NEG_INF-derived value can see).
constexpr int16_t N16 = -32768;        // stored form of every NEG_INF-derived value
constexpr int FINITE_CUT = -(1 << 29);  // values above this are finite (NEG_INF offsets stay below)
__device__ __forceinline__ int widen(int v) { return v; }
__device__ __forceinline__ int widen(int16_t v) { return v == N16 ? NEG_INF : (int)v; }
__device__ __forceinline__ int widen_lo(uint32_t w) { return widen((int16_t)(w & 0xffff)); }
__device__ __forceinline__ int widen_hi(uint32_t w) { return widen((int16_t)(w >> 16)); }
// A value stored as int16: NEG_INF-derived values clamp to N16; finite values are in range
// (the call's bound check).
__device__ __forceinline__ uint32_t pack16(int a, int b) {
  return ((uint32_t)max(a, -32768) & 0xffffu) | ((uint32_t)max(b, -32768) << 16);
}
template <class T> __device__ __forceinline__ T store_of(int v);
template <> __device__ __forceinline__ int store_of<int>(int v) { return v; }
template <> __device__ __forceinline__ int16_t store_of<int16_t>(int v) {
  return (int16_t)max(v, -32768);
}

// CPL consecutive stored scores at base[idx] (idx a multiple of CPL), widened to int32: one
// aligned vector load.
template <int CPL>
__device__ __forceinline__ void load_cols(const int16_t* base, int idx, int* out) {
  if constexpr (CPL == 1) {
    out[0] = widen(base[idx]);
  } else if constexpr (CPL == 2) {
    const uint32_t v = *reinterpret_cast<const uint32_t*>(base + idx);
    out[0] = widen_lo(v); out[1] = widen_hi(v);
  } else if constexpr (CPL == 4) {
    const uint2 v = *reinterpret_cast<const uint2*>(base + idx);
    out[0] = widen_lo(v.x); out[1] = widen_hi(v.x); out[2] = widen_lo(v.y); out[3] = widen_hi(v.y);
  } else {
    static_assert(CPL == 8, "CPL");
    const uint4 v = *reinterpret_cast<const uint4*>(base + idx);
    out[0] = widen_lo(v.x); out[1] = widen_hi(v.x); out[2] = widen_lo(v.y); out[3] = widen_hi(v.y);
    out[4] = widen_lo(v.z); out[5] = widen_hi(v.z); out[6] = widen_lo(v.w); out[7] = widen_hi(v.w);
  }
}
template <int CPL>
__device__ __forceinline__ void load_cols(int16_t* base, int idx, int* out) {
  load_cols<CPL>(const_cast<const int16_t*>(base), idx, out);
}
template <int CPL>
__device__ __forceinline__ void store_cols(int16_t* base, int idx, const int* v) {
  if constexpr (CPL == 1) {
    base[idx] = store_of<int16_t>(v[0]);
  } else if constexpr (CPL == 2) {
    *reinterpret_cast<uint32_t*>(base + idx) = pack16(v[0], v[1]);
  } else if constexpr (CPL == 4) {
    *reinterpret_cast<uint2*>(base + idx) = make_uint2(pack16(v[0], v[1]), pack16(v[2], v[3]));
  } else {
    *reinterpret_cast<uint4*>(base + idx) = make_uint4(pack16(v[0], v[1]), pack16(v[2], v[3]),
                                                       pack16(v[4], v[5]), pack16(v[6], v[7]));
  }
}

// CPL consecutive int32 at base[idx] (idx a multiple of CPL): one aligned vector load.
template <int CPL, class P>
__device__ __forceinline__ void load_cols(P base, int idx, int* out) {
  if constexpr (CPL == 1) {
    out[0] = base[idx];
  } else if constexpr (CPL == 2) {
    const int2 v = *reinterpret_cast<const int2*>(base + idx);
    out[0] = v.x; out[1] = v.y;
  } else if constexpr (CPL == 4) {
    const int4 v = *reinterpret_cast<const int4*>(base + idx);
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
  } else {
    static_assert(CPL == 8, "CPL");
    const int4 a = *reinterpret_cast<const int4*>(base + idx);
    const int4 b = *reinterpret_cast<const int4*>(base + idx + 4);
    out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
    out[4] = b.x; out[5] = b.y; out[6] = b.z; out[7] = b.w;
  }
}
template <int CPL, class P>
__device__ __forceinline__ void store_cols(P base, int idx, const int* v) {
  if constexpr (CPL == 1) {
    base[idx] = v[0];
  } else if constexpr (CPL == 2) {
    *reinterpret_cast<int2*>(base + idx) = make_int2(v[0], v[1]);
  } else if constexpr (CPL == 4) {
    *reinterpret_cast<int4*>(base + idx) = make_int4(v[0], v[1], v[2], v[3]);
  } else {
    *reinterpret_cast<int4*>(base + idx) = make_int4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<int4*>(base + idx + 4) = make_int4(v[4], v[5], v[6], v[7]);
  }
}

// M/D-source accumulation of one predecessor row into a thread's CPL columns [cb, cb+CPL).
// Hrow/Drow point at the predecessor band's first stored column (8*Bs[p]); len is the number
// of int32 readable there. Because every band starts and ends on a multiple of 8 columns, a
// thread's columns are either all inside the predecessor's band or all outside it: inside,
// one aligned vector load per matrix; outside, a clamped (in-bounds) load whose values the
// range masks discard. Column cb's diagonal is H[cb-1], one scalar load.
template <int CPL, class P>
__device__ __forceinline__ void accumulate(P Hrow, P Drow, int len, int pofs, int cb, int c0,
                                           int c1, int mlo, int mhi, int dlo, int dhi, int pbase,
                                           const int* qc, int match, int mismatch, int o, int e,
                                           int* h, int* d) {
  const int idx = cb - pofs;
  const int safe = min(max(idx, 0), len - CPL);
  int hv[CPL], dv[CPL];
  load_cols<CPL>(Hrow, safe, hv);
  load_cols<CPL>(Drow, safe, dv);
  const int hprev = widen(Hrow[min(max(idx - 1, 0), len - 1)]);
  // A predecessor whose bands cover the row's whole pass (the common case, a wave-uniform
  // test) needs no per-column masks.
  const bool whole = mlo <= c0 && mhi >= c1 && dlo <= c0 && dhi >= c1;
  if (whole) {
#pragma unroll
    for (int l = 0; l < CPL; l++) {
      const int hp = l == 0 ? (cb == c0 ? NEG_INF : hprev) : hv[l - 1];
      h[l] = max(h[l], hp + (pbase == qc[l] ? match : mismatch));
      d[l] = max(d[l], max(dv[l] + e, hv[l] + o));
    }
    return;
  }
  const int mlo_l = mlo - cb, mhi_l = mhi - cb, dlo_l = dlo - cb, dhi_l = dhi - cb;
#pragma unroll
  for (int l = 0; l < CPL; l++) {
    const int hp = l == 0 ? (cb == c0 ? NEG_INF : hprev) : hv[l - 1];
    const int cand = hp + (pbase == qc[l] ? match : mismatch);
    const int mv = max(h[l], cand);
    h[l] = (l >= mlo_l && l < mhi_l) ? mv : h[l];
    const int dd = max(d[l], max(dv[l] + e, hv[l] + o));
    d[l] = (l >= dlo_l && l < dhi_l) ? dd : d[l];
  }
}

// CPL query codes for columns [cb, cb+CPL) from the shifted query qs (qs[c] = code of the
// query character column c consumes; qs[0] = N). cb is a multiple of CPL.
template <int CPL, class P>
__device__ __forceinline__ void load_q(P qs, int cb, int cap, int* qc) {
  const int at = min(cb, cap - CPL);
  if constexpr (CPL == 1) {
    qc[0] = qs[at];
  } else if constexpr (CPL == 2) {
    const uint32_t v = *reinterpret_cast<const uint16_t*>(qs + at);
    qc[0] = v & 0xff; qc[1] = v >> 8;
  } else if constexpr (CPL == 4) {
    const uint32_t v = *reinterpret_cast<const uint32_t*>(qs + at);
#pragma unroll
    for (int l = 0; l < 4; l++) qc[l] = (v >> (8 * l)) & 0xff;
  } else {
    const uint2 v = *reinterpret_cast<const uint2*>(qs + at);
#pragma unroll
    for (int l = 0; l < 4; l++) {
      qc[l] = (v.x >> (8 * l)) & 0xff;
      qc[4 + l] = (v.y >> (8 * l)) & 0xff;
    }
  }
}


// One predecessor's M/D-source contribution to this thread's columns (the band conditions of
// src/align.cpp:1467 and :1484 as contiguous column ranges), from the LDS row cache or from
// HBM. The cache directory lives in registers, lane k describing slot k (every wave keeps the
// same copy): meta_row, the local row the slot holds (-1: none); meta_bsbe, its band; a
// lookup is one compare and a ballot, with no LDS access. Slot k holds its row's band at
// R.H / R.D + k * sw.
__device__ __forceinline__ int cache_slot(int meta_row, int p) {
  const uint64_t hit = __ballot(meta_row == p);
  return hit ? (int)__builtin_ctzll(hit) : -1;
}

// The DP matrices of one call in their storage type T (int or int16_t): HBM rows (H, D, I at
// row_off) and the LDS row cache (slot k at RH / RD + k * sw).
template <class T>
struct DpMats {
  T *H, *D, *I;
  T *RH, *RD;
};

template <int CPL, class T>
__device__ __forceinline__ void pred_contrib(const Arena& A, const DpMats<T>& X, const DpScores& sc,
                                             int p, int pbase, int cb, int c0, int c1,
                                             const int* qc, int* h, int* d, int meta_row,
                                             int meta_bsbe, int sw) {
  if (p == 0) pbase = CODE_N;
  const int slot = cache_slot(meta_row, p);
  if (slot >= 0) {
    const int bsbe = __builtin_amdgcn_readlane(meta_bsbe, slot);
    const int bs = bsbe & 0xffff, be = bsbe >> 16;
    const int mlo = max(c0, bs == 0 ? 0 : (bs + 1) * Q), mhi = min(c1, (be - 1) * Q);
    const int dlo = max(c0, bs * Q), dhi = min(c1, (be - 1) * Q);
    accumulate<CPL>(X.RH + slot * sw, X.RD + slot * sw, sw, bs * Q, cb, c0, c1, mlo, mhi, dlo,
                    dhi, pbase, qc, sc.match, sc.mismatch, sc.o1, sc.e1, h, d);
  } else {
    const int bs = A.Bs[p], be = A.Be[p], off = A.row_off[p];
    const int mlo = max(c0, bs == 0 ? 0 : (bs + 1) * Q), mhi = min(c1, (be - 1) * Q);
    const int dlo = max(c0, bs * Q), dhi = min(c1, (be - 1) * Q);
    accumulate<CPL>(X.H + off, X.D + off, (be - bs) * Q, bs * Q, cb, c0, c1, mlo, mhi, dlo, dhi,
                    pbase, qc, sc.match, sc.mismatch, sc.o1, sc.e1, h, d);
    drain_vmem();
  }
}

// One DP row i (local) for rows in rank window [beg_i, end_i]: M/D sources from every
// in-window predecessor, then the I/H scan as an exact prefix max:
//   I[j] = j*e + max(NEG_INF, max_{k<j} (G[k] + o - (k+1)*e)),  H = max(G, I),
// G = max(M-source, D-source). This equals the serial recurrence of align.cpp:1507-1514,
// I[j] = max(I[j-1]+e, H[j-1]+o), because o <= e (gap_open <= 0) and int32 never wraps here.
// Writes H, D, I (and the NEG_INF sentinel block) to HBM and, when the row has a cache slot
// (slot_i >= 0), H and D to LDS. The common path reads LDS only: on gfx9 a wait for any HBM
// load also waits for every store issued before it, so HBM reads here (predecessors outside
// the cache, predecessors after the third, a query too long for LDS) sit in branches of their
// own.
// Returns the first argmax column (local) of the row's H, for the adaptive band; *np_in gets
// the number of in-window predecessors. In the int16 form, fmin keeps the thread's smallest
// finite H as an unsigned offset from FINITE_CUT (NEG_INF-derived values wrap to above 2^31
// and never lower it).
template <int NT, int CPL, bool QLDS, class T, class RingT>
__device__ __forceinline__ int dp_row(const Arena& A, WinShared& S, RingT& R, const DpMats<T>& X,
                                      const DpScores& sc, int i, const RowRec& rr, int bs, int be,
                                      bool want_argmax, int* np_in, int meta_row, int meta_bsbe,
                                      int sw, int slot_i, int& scan_ph, uint32_t& fmin) {
  const int tid = threadIdx.x;
  const int c0 = bs * Q, W = (be - bs - 1) * Q, c1 = c0 + W;
  T* Hi = X.H + rr.row_off;  // column c at Hi[c - c0]
  T* Di = X.D + rr.row_off;
  T* Ii = X.I + rr.row_off;
  const bool cache_i = slot_i >= 0;
  T* RHi = X.RH + (cache_i ? slot_i : 0) * sw;  // column c at RHi[c - c0]
  T* RDi = X.RD + (cache_i ? slot_i : 0) * sw;
  const int o = sc.o1, e = sc.e1;
  const int beg = S.beg_i;
  const int n3 = min(rr.np, 3);
  int carry = SCAN_ID;
  int best = SCAN_ID, best_j = 0x7fffffff;
  int npi = 0;
  for (int pb = c0; pb < c1; pb += NT * CPL) {
    const int cb = pb + tid * CPL;
    int h[CPL], d[CPL];
#pragma unroll
    for (int l = 0; l < CPL; l++) h[l] = d[l] = NEG_INF;
    int qc[CPL];
    if constexpr (QLDS) load_q<CPL>(R.q, cb, RingT::kQ, qc);
    else load_q<CPL>(A.qsh, cb, S.q_cap, qc);
    npi = 0;
    for (int k = 0; k < n3; k++) {
      const int prank = k == 0 ? rr.p0 : k == 1 ? rr.p1 : rr.p2;
      const int p = prank - beg;
      if (p < 0) continue;
      npi++;
      pred_contrib<CPL>(A, X, sc, p, (rr.pk >> (3 * k)) & 7, cb, c0, c1, qc, h, d, meta_row,
                        meta_bsbe, sw);
    }
    if (rr.np > 3) {
      // the fourth to sixth from the record (LDS), later ones from HBM
      const int n6 = min(rr.np, 6);
      for (int k = 3; k < n6; k++) {
        const int x = k == 3 ? rr.w0 : k == 4 ? rr.w1 : rr.w2;
        const int p = x >> 3;
        if (p < 0) continue;
        npi++;
        pred_contrib<CPL>(A, X, sc, p, x & 7, cb, c0, c1, qc, h, d, meta_row, meta_bsbe, sw);
      }
      if (rr.np > 6) {
        for (int k = 6; k < rr.np; k++) {
          const int p = A.pred_rank[rr.k0 + k] - beg;
          const int pbase = A.pred_base[rr.k0 + k] & 7;
          drain_vmem();
          if (p < 0) continue;
          npi++;
          pred_contrib<CPL>(A, X, sc, p, pbase, cb, c0, c1, qc, h, d, meta_row, meta_bsbe, sw);
        }
        drain_vmem();
      }
    }
    int g[CPL], t[CPL];
    int lane_max = SCAN_ID;
#pragma unroll
    for (int l = 0; l < CPL; l++) {
      const int c = cb + l;
      g[l] = max(h[l], d[l]);
      t[l] = c < c1 ? g[l] + o - (c - c0 + 1) * e : SCAN_ID;
      lane_max = max(lane_max, t[l]);
    }
    int total;
    int run = max(block_excl_max<NT>(lane_max, S, &total, scan_ph), carry);
    scan_ph ^= 1;
    if (cb < c1) {  // W and cb - c0 are multiples of CPL: a thread's columns are all in or out
      int hv[CPL], iv[CPL];
#pragma unroll
      for (int l = 0; l < CPL; l++) {
        const int j = cb + l - c0;
        iv[l] = j * e + max(NEG_INF, run);
        hv[l] = max(g[l], iv[l]);
        run = max(run, t[l]);
        if (want_argmax && hv[l] > best) { best = hv[l]; best_j = j; }
        if constexpr (sizeof(T) == 2) fmin = min(fmin, (uint32_t)(hv[l] - FINITE_CUT));
      }
      const int j0 = cb - c0;
      store_cols<CPL>(Hi, j0, hv);
      store_cols<CPL>(Di, j0, d);
      store_cols<CPL>(Ii, j0, iv);
      if (cache_i) {
        store_cols<CPL>(RHi, j0, hv);
        store_cols<CPL>(RDi, j0, d);
      }
    }
    carry = max(carry, total);
  }
  *np_in = npi;
  // NEG_INF sentinel block after the band (align.cpp:1505-1506,1515).
  if (tid < Q) Hi[W + tid] = Di[W + tid] = Ii[W + tid] = store_of<T>(NEG_INF);
  if (!want_argmax) return 0;
  // First argmax over the row (value desc, column asc): the wave's maximum, then the smallest
  // column holding it (two DPP reductions).
  {
    const int own = best;
    best = __builtin_amdgcn_readlane(wave_incl_max(own), 63);
    best_j = __builtin_amdgcn_readlane(wave_incl_min(own == best ? best_j : 0x7fffffff), 63);
  }
  if (NT > 64) {
    const int wave = tid >> 6;
    lds_barrier<NT>();
    if ((tid & 63) == 0) { S.scan[2 * wave] = best; S.scan[2 * wave + 1] = best_j; }
    lds_barrier<NT>();
    best = S.scan[0]; best_j = S.scan[1];
    for (int w2 = 1; w2 < NT / 64; w2++) {
      const int ov = S.scan[2 * w2], oj = S.scan[2 * w2 + 1];
      if (ov > best || (ov == best && oj < best_j)) { best = ov; best_j = oj; }
    }
    lds_barrier<NT>();
  }
  return best_j == 0x7fffffff ? 0 : best_j;
}

// Static band per row (src/align.cpp:1378-1394) and ragged row offsets; all lanes.
// Returns false on arena overflow.
template <int NT>
__device__ __forceinline__ bool plan_static(const Arena& A, WinShared& S, const DpScores& sc, const SlotCaps& cap,
                            int w) {
  const int n = S.n_rows, m = S.m;
  const int beg = S.beg_i, end = S.end_i;
  // Block-wide exclusive scan of row sizes in chunks of NT.
  int64_t base = 0;
  for (int r0 = 0; r0 < n; r0 += NT) {
    const int r = r0 + threadIdx.x;
    int len = 0;
    if (r < n) {
      int bs, be;
      if (sc.f > 0) {
        const int aci = beg + r;
        const int lo = A.hlen[aci] - A.hlen[beg], hi = m - A.tlen[aci] + A.tlen[end];
        const int Ms = max(0, min(lo, hi) - w), Me = min(m, max(lo, hi) + w);
        bs = Ms / Q;
        be = Me / Q + 2;
        if (r > 0) atomicAdd(reinterpret_cast<unsigned long long*>(&S.call_logical),
                             (unsigned long long)(Me - Ms + 1));
      } else {
        bs = 0;
        be = m / Q + 2;
        if (r > 0) atomicAdd(reinterpret_cast<unsigned long long*>(&S.call_logical),
                             (unsigned long long)(m + 1));
      }
      A.Bs[r] = bs;
      A.Be[r] = be;
      len = (be - bs) * Q;
      atomicMax(&S.row_wmax, len - Q);
    }
    // prefix sum via shuffles + shared
    int v = len;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      int u = __shfl_up(v, s, 64);
      if (lane >= s) v += u;
    }
    if (lane == 63) S.scan[wave] = v;
    __syncthreads();
    int before = 0, all = 0;
    for (int w2 = 0; w2 < NT / 64; w2++) {
      if (w2 < wave) before += S.scan[w2];
      all += S.scan[w2];
    }
    __syncthreads();
    if (r < n) {
      const int off = (int)(base + before + v - len);
      A.row_off[r] = off;
      int* rec = A.rowrec + 12LL * (beg + r);
      rec[0] = off;
      rec[1] = A.Bs[r] | (A.Be[r] << 16);
    }
    base += all;
  }
  if (threadIdx.x == 0) {
    A.row_off[n] = (int)base;
    S.total = base;
  }
  __syncthreads();
  return base <= cap.dp;
}

// Backtrack (src/align.cpp:1572-1736), one lane. Emits res[acj-1] for every consumed query
// column, so the result is already in forward order.
template <class T>
__device__ __forceinline__ int backtrack(const Arena& A, WinShared& S, const DpScores& sc,
                                         const DpMats<T>& X) {
  const int n = S.n_rows, m = S.m, beg = S.beg_i;
  const int match = sc.match, mismatch = sc.mismatch, o1 = sc.o1, e1 = sc.e1;
  auto Hat = [&](int r, int j) { return widen(X.H[A.row_off[r] + j]); };
  auto Dat = [&](int r, int j) { return widen(X.D[A.row_off[r] + j]); };
  auto Iat = [&](int r, int j) { return widen(X.I[A.row_off[r] + j]); };
  int i = n - 1, acj = m, op = ALL_OP;
  auto emit_sub = [&](int pid, int p) {
    const int pbase = p == 0 ? CODE_N : A.base[pid];
    const int qc = A.qseq[acj - 1];
    if (pbase == qc) {
      A.res_from[acj - 1] = pid;
      A.res_base[acj - 1] = A.base[pid];
      A.res_aligned[acj - 1] = -1;
    } else {
      const int par = A.par[pid];
      const int al = A.aligned[par * NB + qc];
      A.res_from[acj - 1] = al;
      A.res_base[acj - 1] = (uint8_t)qc;
      A.res_aligned[acj - 1] = al != -1 ? -1 : par;
    }
  };
  while (i > 0 || acj > 0) {
    const int Bsi = A.Bs[i], Bei = A.Be[i];
    const int j = acj - Bsi * Q;
    if (acj > (Bei - 1) * Q || acj < Bsi * Q) return MPOA_E_BAND;
    const int k0 = A.pred_off[beg + i], k1 = A.pred_off[beg + i + 1];
    const int u = A.rank_id[beg + i];
    const int h = Hat(i, j);
    if ((op & M_OP) && acj > 0) {
      int bk = -1, bw = 0, bp = 0;
      const int qc = A.qseq[acj - 1];
      for (int k = k0; k < k1; k++) {
        const int p = A.pred_rank[k] - beg;
        if (p < 0) continue;
        const int pbase = p == 0 ? CODE_N : (A.pred_base[k] & 7);
        if (pbase != qc) continue;
        const int pj = acj - A.Bs[p] * Q;
        const int bn = A.Be[p] - A.Bs[p] - 1;
        if (pj - 1 >= 0 && pj - 1 < bn * Q && h == Hat(p, pj - 1) + match &&
            (bk == -1 || A.pred_w[k] > bw)) {
          bk = k; bw = A.pred_w[k]; bp = p;
        }
      }
      if (bk != -1 && bw >= A.ind[u] / 10) {
        emit_sub(A.rank_id[beg + bp], bp);
        op = ALL_OP;
        i = bp;
        acj--;
        continue;
      }
    }
    if (op & D_OP) {
      if (op == D_OP || h == Dat(i, j)) {
        const int dv = Dat(i, j);
        int bk = -1, bw = 0, bp = 0, bop = 0;
        for (int pass = 0; pass < 2 && bk == -1; pass++) {
          for (int k = k0; k < k1; k++) {
            const int p = A.pred_rank[k] - beg;
            if (p < 0) continue;
            if (A.Bs[p] <= acj / Q && acj / Q < A.Be[p] - 1) {
              const int pj = acj - A.Bs[p] * Q;
              const int src = pass == 0 ? Hat(p, pj) + o1 : Dat(p, pj) + e1;
              if (dv == src && (bk == -1 || A.pred_w[k] > bw)) {
                bk = k; bw = A.pred_w[k]; bp = p;
                bop = pass == 0 ? (M_OP | I_OP) : D_OP;
              }
            }
          }
        }
        if (bk != -1) {
          i = bp;
          op = bop;
          continue;
        }
      }
    }
    if ((op & I_OP) && acj > 0 && j - 1 >= 0) {
      const int iv = Iat(i, j);
      if (op == I_OP || h == iv) {
        const int qc = A.qseq[acj - 1];
        if (iv == Hat(i, j - 1) + o1) {
          A.res_from[acj - 1] = -1; A.res_base[acj - 1] = (uint8_t)qc; A.res_aligned[acj - 1] = -1;
          op = M_OP | D_OP;
          acj--;
          continue;
        }
        if (iv == Iat(i, j - 1) + e1) {
          A.res_from[acj - 1] = -1; A.res_base[acj - 1] = (uint8_t)qc; A.res_aligned[acj - 1] = -1;
          op = I_OP;
          acj--;
          continue;
        }
      }
    }
    if ((op & M_OP) && acj > 0) {
      int bk = -1, bw = 0, bp = 0;
      const int qc = A.qseq[acj - 1];
      for (int k = k0; k < k1; k++) {
        const int p = A.pred_rank[k] - beg;
        if (p < 0) continue;
        const int pbase = p == 0 ? CODE_N : (A.pred_base[k] & 7);
        const int pj = acj - A.Bs[p] * Q;
        const int bn = A.Be[p] - A.Bs[p] - 1;
        if (pj - 1 >= 0 && pj - 1 < bn * Q &&
            h == Hat(p, pj - 1) + (pbase == qc ? match : mismatch) &&
            (bk == -1 || A.pred_w[k] > bw)) {
          bk = k; bw = A.pred_w[k]; bp = p;
        }
      }
      if (bk != -1) {
        op = ALL_OP;
        emit_sub(A.rank_id[beg + bp], bp);
        i = bp;
        acj--;
        continue;
      }
    }
    return MPOA_E_BACKTRACK;
  }
  return MPOA_OK;
}

// The backtrack (align.cpp:1572-1736, the same moves and tie-breaks as backtrack() above) by
// one wavefront. Lane k < np holds the current row's k-th predecessor (in-list order) with
// its band, row offset and row record; lane 63 reads the current cell. Each move costs one
// round trip to HBM for the cells, plus one for the predecessors' records when the row
// changes; the "largest in-weight, first wins" choices are wave maxima of
// (weight << 6 | 63 - k). Moves are logged as local predecessor ranks (-1: insertion) in
// `log` (LDS when the query fits) and turned into res_t entries afterwards by every thread
// (resolve_res).
// Returns MPOA_OK, an error, or 1 when a row has more than 63 predecessors (the caller then
// runs backtrack()). Called by wave 0 only.
// Best candidate lane for the backtrack: largest weight, lowest lane; -1 if none.
__device__ __forceinline__ int bt_pick(bool cand, int w, int lane) {
  const int key = cand ? (w << 6) | (63 - lane) : -1;
  const int best = __builtin_amdgcn_readlane(wave_incl_max(key), 63);
  return best < 0 ? -1 : 63 - (best & 63);
}

// A run of match moves along consecutive ranks, checked by the whole wave at once: lane t
// takes the move from (row i - t, column acj - t) to (i - t - 1, acj - t - 1) and checks that
// minipoa's first rule (the match-only step, align.cpp:1591-1626) takes exactly it: the row's
// heaviest in-window predecessor (first of the largest in-weight) is row i - t - 1, its base is
// the query's, its in-weight passes the ind/10 rule, the column is inside the row's band and
// the predecessor's, and H equals the predecessor's diagonal H plus the match score. The
// diagonal cell is the one lane t + 1 loads for its own move. Lane 0 also needs op to hold M.
// Lanes 0..K (K <= 63) take part. Moves of the leading lanes that pass are taken (at most K
// per call, two round trips), and logged as their predecessor's local rank. Returns their number; *nbs, *nbe, *noff, *nrr get
// the geometry and record of the row reached.
template <class T>
__device__ __forceinline__ int bt_run(const Arena& A, const WinShared& S, const T* H, int match,
                                      int i, int acj, int op, int K, int* log, int* nbs, int* nbe,
                                      int* noff, RowRec* nrr) {
  const int lane = threadIdx.x & 63;
  const int beg = S.beg_i;
  const int r = i - lane, a = acj - lane;
  const bool live = lane <= K && r >= 0 && a >= 0;
  int bs = 0, be = 0, off = 0, qc = -1;
  RowRec rec{};
  if (live) {
    bs = A.Bs[r];
    be = A.Be[r];
    off = (int)A.row_off[r];
    rec = load_rowrec(A.rowrec, beg + r);
    if (a > 0) qc = A.qseq[a - 1];
  }
  const bool in_band = live && a >= bs * Q && a <= (be - 1) * Q;
  T rh = 0;  // widened after the branch: no wait inside it
  if (in_band) rh = H[off + a - bs * Q];
  const int h = widen(rh);
  const int up_bs = __shfl_down(bs, 1, 64), up_be = __shfl_down(be, 1, 64);
  const int up_h = __shfl_down(h, 1, 64);
  bool ok = in_band && r >= 1 && a >= 1 && lane < K && rec.np >= 1 && rec.np <= 3;
  if (ok) {
    int kh = -1, wmax = -1, ph = -1;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int prank = k == 0 ? rec.p0 : k == 1 ? rec.p1 : rec.p2;
      const int w = k == 0 ? rec.w0 : k == 1 ? rec.w1 : rec.w2;
      if (k < rec.np && prank >= beg && w > wmax) {
        wmax = w;
        kh = k;
        ph = prank - beg;
      }
    }
    ok = kh >= 0 && ph == r - 1;
    if (ok) {
      const int pbase = ph == 0 ? CODE_N : (rec.pk >> (3 * kh)) & 7;
      const int rule = (rec.pk >> (9 + kh)) & 1;
      const int pj = a - up_bs * Q;
      ok = pbase == qc && rule && pj - 1 >= 0 && pj - 1 < (up_be - up_bs - 1) * Q &&
           h == up_h + match;
    }
  }
  if (lane == 0 && !(op & M_OP)) ok = false;
  const uint64_t good = __ballot(ok);
  const int L = (int)__builtin_ctzll(~good);
  if (lane < L) log[a - 1] = r - 1;
  *nbs = __builtin_amdgcn_readlane(bs, L);
  *nbe = __builtin_amdgcn_readlane(be, L);
  *noff = __builtin_amdgcn_readlane(off, L);
  nrr->row_off = __builtin_amdgcn_readlane(rec.row_off, L);
  nrr->bsbe = __builtin_amdgcn_readlane(rec.bsbe, L);
  nrr->k0 = __builtin_amdgcn_readlane(rec.k0, L);
  nrr->np = __builtin_amdgcn_readlane(rec.np, L);
  nrr->p0 = __builtin_amdgcn_readlane(rec.p0, L);
  nrr->p1 = __builtin_amdgcn_readlane(rec.p1, L);
  nrr->p2 = __builtin_amdgcn_readlane(rec.p2, L);
  nrr->pk = __builtin_amdgcn_readlane(rec.pk, L);
  nrr->w0 = __builtin_amdgcn_readlane(rec.w0, L);
  nrr->w1 = __builtin_amdgcn_readlane(rec.w1, L);
  nrr->w2 = __builtin_amdgcn_readlane(rec.w2, L);
  nrr->last = __builtin_amdgcn_readlane(rec.last, L);
  return L;
}

// The heaviest in-window predecessor of local row r (the first of the largest in-weight, the
// one the match-only move takes when it is a candidate, align.cpp:1591-1626), packed as its
// local rank | base << 24 | (in-weight >= ind/10) << 27; -1 for row 0, a row without one, or a
// row with more than 63 predecessors (the one-lane backtrack's).
__device__ __forceinline__ int heaviest_pred(const Arena& A, int beg, int r) {
  if (r == 0) return -1;
  const RowRec rec = load_rowrec(A.rowrec, beg + r);
  if (rec.np < 1 || rec.np > 63) return -1;
  int best = -1, bw = -1, b8 = 0;
  for (int k = 0; k < rec.np; k++) {
    int prank, w, b;
    if (k < 3) {
      prank = k == 0 ? rec.p0 : k == 1 ? rec.p1 : rec.p2;
      w = k == 0 ? rec.w0 : k == 1 ? rec.w1 : rec.w2;
      b = ((rec.pk >> (3 * k)) & 7) | (((rec.pk >> (9 + k)) & 1) << 7);
    } else {
      prank = A.pred_rank[rec.k0 + k];
      w = A.pred_w[rec.k0 + k];
      b = A.pred_base[rec.k0 + k];
    }
    if (prank < beg || w <= bw) continue;
    bw = w;
    best = prank - beg;
    b8 = b;
  }
  return best < 0 ? -1 : best | ((b8 & 7) << 24) | (((b8 >> 7) & 1) << 27);
}

// A run of match moves along the heaviest-predecessor chain, checked by the whole wave at
// once: lane t takes row r_t (r_0 = i, r_{t+1} = the heaviest in-window predecessor of r_t,
// from hp) at column acj - t and checks that minipoa's first rule takes the move to
// (r_{t+1}, acj - t - 1): r_{t+1} is a candidate (its base is the query's, the column is
// inside both bands, H equals its diagonal H plus the match score), so as the heaviest it is
// the one chosen, and its in-weight passes the ind/10 rule. Otherwise as bt_run.
template <class T>
__device__ __forceinline__ int bt_chain(const Arena& A, const WinShared& S, const T* H,
                                        const int* hp, int match, int i, int acj, int op, int K,
                                        int* log, int* ni, int* nbs, int* nbe, int* noff,
                                        RowRec* nrr) {
  const int lane = threadIdx.x & 63;
  const int beg = S.beg_i;
  int r = -1, hk = -1;
  for (int t = 0, c = i; t <= K; t++) {  // the chain, one LDS read a step
    const int v = c >= 0 ? hp[c] : -1;
    if (lane == t) {
      r = c;
      hk = v;
    }
    c = v < 0 ? -1 : (v & 0xffffff);
  }
  const int a = acj - lane;
  const bool live = lane <= K && r >= 0 && a >= 0;
  int bs = 0, be = 0, off = 0, qc = -1;
  RowRec rec{};
  if (live) {
    bs = A.Bs[r];
    be = A.Be[r];
    off = (int)A.row_off[r];
    rec = load_rowrec(A.rowrec, beg + r);
    if (a > 0) qc = A.qseq[a - 1];
  }
  const bool in_band = live && a >= bs * Q && a <= (be - 1) * Q;
  T rh = 0;  // widened after the branch: no wait inside it
  if (in_band) rh = H[off + a - bs * Q];
  const int h = widen(rh);
  const int up_bs = __shfl_down(bs, 1, 64), up_be = __shfl_down(be, 1, 64);
  const int up_h = __shfl_down(h, 1, 64);
  const int p = hk & 0xffffff;
  bool ok = in_band && r >= 1 && a >= 1 && lane < K && hk >= 0;
  if (ok) {
    const int pbase = p == 0 ? CODE_N : (hk >> 24) & 7;
    const int pj = a - up_bs * Q;  // lane t + 1 holds row p: its band and diagonal cell
    ok = pbase == qc && ((hk >> 27) & 1) && pj - 1 >= 0 &&
         pj - 1 < (up_be - up_bs - 1) * Q && h == up_h + match;
  }
  if (lane == 0 && !(op & M_OP)) ok = false;
  const uint64_t good = __ballot(ok);
  const int L = (int)__builtin_ctzll(~good);
  if (lane < L) log[a - 1] = p;
  *ni = __builtin_amdgcn_readlane(r, L);
  *nbs = __builtin_amdgcn_readlane(bs, L);
  *nbe = __builtin_amdgcn_readlane(be, L);
  *noff = __builtin_amdgcn_readlane(off, L);
  nrr->row_off = __builtin_amdgcn_readlane(rec.row_off, L);
  nrr->bsbe = __builtin_amdgcn_readlane(rec.bsbe, L);
  nrr->k0 = __builtin_amdgcn_readlane(rec.k0, L);
  nrr->np = __builtin_amdgcn_readlane(rec.np, L);
  nrr->p0 = __builtin_amdgcn_readlane(rec.p0, L);
  nrr->p1 = __builtin_amdgcn_readlane(rec.p1, L);
  nrr->p2 = __builtin_amdgcn_readlane(rec.p2, L);
  nrr->pk = __builtin_amdgcn_readlane(rec.pk, L);
  nrr->w0 = __builtin_amdgcn_readlane(rec.w0, L);
  nrr->w1 = __builtin_amdgcn_readlane(rec.w1, L);
  nrr->w2 = __builtin_amdgcn_readlane(rec.w2, L);
  nrr->last = __builtin_amdgcn_readlane(rec.last, L);
  return L;
}

// hp: the heaviest in-window predecessor of every row (heaviest_pred) in LDS, or null; with
// it, runs follow the heaviest-predecessor chain (bt_chain) and are tried after every match
// or substitution move, else they follow consecutive ranks (bt_run).
template <class T>
__device__ __forceinline__ int backtrack_wave(const Arena& A, WinShared& S, const DpScores& sc,
                                              const DpMats<T>& X, int* log, const int* hp) {
  const int lane = threadIdx.x & 63;
  const int n = S.n_rows, m = S.m, beg = S.beg_i;
  const int match = sc.match, mismatch = sc.mismatch, o1 = sc.o1, e1 = sc.e1;
  int i = n - 1, acj = m, op = ALL_OP;
  // current row (uniform)
  int bs_i = A.Bs[i], be_i = A.Be[i], off_i = (int)A.row_off[i];
  RowRec rr = load_rowrec(A.rowrec, beg + i);
  // this lane's predecessor of the current row
  int p = -1, pw = 0, pbase = 0, prule = 0, pbs = 0, pbe = 0, poff = 0;
  RowRec prr{};
  bool setup = true;
  // Runs of match moves are tried along the heaviest-predecessor chain after every match or
  // substitution move when hp is given (about 88% of the moves of these graphs lie in such runs, 9 moves long on
  // average), else along consecutive ranks after two match moves in a row to the previous rank
  // (where a run is likely to go on); and again after a run that used its whole width. The
  // width starts at 8 moves and doubles after such a run (up to 63), since a run's loads cost
  // more the more lanes take part.
  bool run = false;
  int width = 8, streak = 0;
  for (;;) {
    if (run && (op & M_OP) && acj > 0 && i > 0) {
      int ni = i, nbs, nbe, noff;
      RowRec nrr;
      const int L = hp ? bt_chain(A, S, X.H, hp, match, i, acj, op, width, log, &ni, &nbs,
                                  &nbe, &noff, &nrr)
                       : bt_run(A, S, X.H, match, i, acj, op, width, log, &nbs, &nbe, &noff,
                                &nrr);
      if (!hp) ni = i - L;  // consecutive ranks
      run = L == width;  // a full run may go on; otherwise the next move is not such a match
      width = run ? min(2 * width, 63) : 8;
      if (!run) streak = 0;
      if (L > 0) {
        i = ni;
        acj -= L;
        op = ALL_OP;
        bs_i = nbs;
        be_i = nbe;
        off_i = noff;
        rr = nrr;
        setup = true;
        continue;
      }
    }
    if (setup) {
      if (rr.np > 63) return 1;
      p = -1;
      if (lane < rr.np) {
        int prank, b8;
        if (lane < 3) {
          prank = lane == 0 ? rr.p0 : lane == 1 ? rr.p1 : rr.p2;
          pw = lane == 0 ? rr.w0 : lane == 1 ? rr.w1 : rr.w2;
          b8 = ((rr.pk >> (3 * lane)) & 7) | (((rr.pk >> (9 + lane)) & 1) << 7);
        } else {
          prank = A.pred_rank[rr.k0 + lane];
          pw = A.pred_w[rr.k0 + lane];
          b8 = A.pred_base[rr.k0 + lane];
        }
        p = prank - beg;
        pbase = p == 0 ? CODE_N : (b8 & 7);
        prule = b8 >> 7;
        if (p >= 0) {
          pbs = A.Bs[p];
          pbe = A.Be[p];
          poff = (int)A.row_off[p];
          prr = load_rowrec(A.rowrec, beg + p);
        }
      }
      setup = false;
    }
    if (i <= 0 && acj <= 0) break;
    const int j = acj - bs_i * Q;
    if (acj > (be_i - 1) * Q || acj < bs_i * Q) return MPOA_E_BAND;
    // One round trip: the predecessors' cells (lanes < np) and the current cell (lane 63).
    const bool valid = p >= 0;
    const int pj = acj - pbs * Q;
    const bool m_in = valid && acj > 0 && pj - 1 >= 0 && pj - 1 < (pbe - pbs - 1) * Q;
    const bool d_in = valid && pbs <= acj / Q && acj / Q < pbe - 1;
    // The stored values are loaded in the branches and widened after them, so that every
    // load of the move is issued before the one wait for them.
    T r_hdiag = 0, r_hup = 0, r_dup = 0;
    if (m_in) r_hdiag = X.H[poff + pj - 1];
    if (d_in) {
      r_hup = X.H[poff + pj];
      r_dup = X.D[poff + pj];
    }
    T r_h = 0, r_d = 0, r_i = 0, r_hm = 0, r_im = 0;
    int c_q = 0;
    if (lane == 63) {
      r_h = X.H[off_i + j];
      r_d = X.D[off_i + j];
      r_i = X.I[off_i + j];
      if (j >= 1) {
        r_hm = X.H[off_i + j - 1];
        r_im = X.I[off_i + j - 1];
      }
      if (acj > 0) c_q = A.qseq[acj - 1];
    }
    const int hdiag = widen(r_hdiag), hup = widen(r_hup), dup = widen(r_dup);
    const int h = __builtin_amdgcn_readlane(widen(r_h), 63);
    const int dv = __builtin_amdgcn_readlane(widen(r_d), 63);
    const int iv = __builtin_amdgcn_readlane(widen(r_i), 63);
    const int hm = __builtin_amdgcn_readlane(widen(r_hm), 63);
    const int im = __builtin_amdgcn_readlane(widen(r_im), 63);
    const int qc = __builtin_amdgcn_readlane(c_q, 63);
    // kind: 1 substitution move (logged), 2 deletion move, 3 insertion (logged)
    int k = -1, kind = 0, nop = op;
    // 1. match-only move with the in-weight rule (align.cpp:1591-1626)
    bool matched = false;
    if ((op & M_OP) && acj > 0) {
      k = bt_pick(m_in && pbase == qc && h == hdiag + match, pw, lane);
      if (k >= 0 && __builtin_amdgcn_readlane(prule, k)) {
        kind = 1;
        nop = ALL_OP;
        matched = true;
      }
    }
    // 2. deletion: open first, then extend (align.cpp:1627-1667)
    if (kind == 0 && (op & D_OP) && (op == D_OP || h == dv)) {
      k = bt_pick(d_in && dv == hup + o1, pw, lane);
      nop = M_OP | I_OP;
      if (k < 0) {
        k = bt_pick(d_in && dv == dup + e1, pw, lane);
        nop = D_OP;
      }
      if (k >= 0) kind = 2;
    }
    // 3. insertion: open first, then extend (align.cpp:1668-1688)
    if (kind == 0 && (op & I_OP) && acj > 0 && j - 1 >= 0 && (op == I_OP || h == iv)) {
      if (iv == hm + o1) {
        kind = 3;
        nop = M_OP | D_OP;
      } else if (iv == im + e1) {
        kind = 3;
        nop = I_OP;
      }
    }
    // 4. any substitution, no in-weight rule (align.cpp:1689-1728)
    if (kind == 0 && (op & M_OP) && acj > 0) {
      k = bt_pick(m_in && h == hdiag + (pbase == qc ? match : mismatch), pw, lane);
      if (k >= 0) {
        kind = 1;
        nop = ALL_OP;
      }
    }
    if (kind == 0) return MPOA_E_BACKTRACK;
    if (hp) {
      run = kind == 1;  // after a match or a substitution a chain run most likely follows
    } else {
      streak = matched && __builtin_amdgcn_readlane(p, k) == i - 1 ? streak + 1 : 0;
      run = streak >= 2;
    }
    if (kind != 2) {
      if (lane == 0) log[acj - 1] = kind == 1 ? __builtin_amdgcn_readlane(p, k) : -1;
      acj--;
    }
    op = nop;
    if (kind != 3) {
      i = __builtin_amdgcn_readlane(p, k);
      bs_i = __builtin_amdgcn_readlane(pbs, k);
      be_i = __builtin_amdgcn_readlane(pbe, k);
      off_i = __builtin_amdgcn_readlane(poff, k);
      rr.row_off = __builtin_amdgcn_readlane(prr.row_off, k);
      rr.bsbe = __builtin_amdgcn_readlane(prr.bsbe, k);
      rr.k0 = __builtin_amdgcn_readlane(prr.k0, k);
      rr.np = __builtin_amdgcn_readlane(prr.np, k);
      rr.p0 = __builtin_amdgcn_readlane(prr.p0, k);
      rr.p1 = __builtin_amdgcn_readlane(prr.p1, k);
      rr.p2 = __builtin_amdgcn_readlane(prr.p2, k);
      rr.pk = __builtin_amdgcn_readlane(prr.pk, k);
      rr.w0 = __builtin_amdgcn_readlane(prr.w0, k);
      rr.w1 = __builtin_amdgcn_readlane(prr.w1, k);
      rr.w2 = __builtin_amdgcn_readlane(prr.w2, k);
      setup = true;
    }
  }
  return MPOA_OK;
}

// res_t entries from the logged moves of backtrack_wave (emit_sub of align.cpp:1603-1624; the
// log holds a local predecessor rank per query column, -1 for an insertion): a predecessor
// whose base is the query's is that node; otherwise the node aligned to its column head with
// the query base, or a new node aligned to the head. All threads.
template <int NT>
__device__ __forceinline__ void resolve_res(const Arena& A, const WinShared& S, const int* log) {
  const int beg = S.beg_i;
  for (int x = threadIdx.x; x < S.m; x += NT) {
    const int p = log[x];
    const int qc = A.qseq[x];
    if (p < 0) {
      A.res_from[x] = -1;
      A.res_base[x] = (uint8_t)qc;
      A.res_aligned[x] = -1;
      continue;
    }
    const int pid = A.rank_id[beg + p];
    const int pb = p == 0 ? CODE_N : A.base[pid];
    if (pb == qc) {
      A.res_from[x] = pid;
      A.res_base[x] = A.base[pid];
      A.res_aligned[x] = -1;
    } else {
      const int par = A.par[pid];
      const int al = A.aligned[par * NB + qc];
      A.res_from[x] = al;
      A.res_base[x] = (uint8_t)qc;
      A.res_aligned[x] = al != -1 ? -1 : par;
    }
  }
}


// Adaptive band of row i before its DP (src/align.cpp:1433-1444); one lane.
__device__ __forceinline__ void adaptive_band_row(const Arena& A, WinShared& S, int i, int w) {
  const int beg = S.beg_i, end = S.end_i, m = S.m;
  const int aci = beg + i;
  const int lo = A.hlen[aci] - A.hlen[beg], hi = m - A.tlen[aci] + A.tlen[end];
  const int ow = i == 0 ? 0 : A.Ow[i];
  const int Ms = max(0, min(min(A.Pl[i], lo + A.Ol[i]), hi) - w - ow);
  const int Me = min(m, max(max(A.Pr[i], lo + A.Or[i]), hi) + w + ow);
  A.Bs[i] = Ms / Q;
  A.Be[i] = Me / Q + 2;
  if (i > 0) S.call_logical += Me - Ms + 1;
}

// Adaptive-band update after row i (src/align.cpp:1516-1568); one lane, argmax given.
__device__ __forceinline__ void adaptive_update(const Arena& A, WinShared& S, int i, int max_j, int w) {
  const int beg = S.beg_i, n = S.n_rows;
  const int max_acj = A.Bs[i] * Q + max_j;
  const int u = A.rank_id[beg + i];
  int max_slp = 0;
  bool is_match = false;
  A.Mp[i] = max_acj;
  const int k0 = A.pred_off[beg + i], k1 = A.pred_off[beg + i + 1];
  if (A.base[u] == A.qseq[max_acj]) {
    for (int k = k0; k < k1; k++) {
      const int p = A.pred_rank[k] - beg;
      if (p < 0) continue;
      max_slp = A.Sl[p] > max_slp ? A.Sl[p] : max_slp;
      if (A.Mp[p] + 1 == A.Mp[i]) {
        A.Sl[i] = max(A.Sl[i], A.Sl[p] + 1);
        is_match = true;
      }
    }
  }
  if (!is_match && max_slp < 30) {
    A.Ow[i] += max_slp + 1;
    A.Ow[i] = min(A.Ow[i], 5 * w);
  }
  if (A.Sl[i] >= 30) {
    // hlen indexed by the local row number (align.cpp:1544-1545)
    A.Ol[i] = max_acj - (A.hlen[i] - A.hlen[beg]);
    A.Or[i] = max_acj - (A.hlen[i] - A.hlen[beg]);
    A.Ow[i] = 0;
    for (int e = A.out_head[u]; e != -1; e = A.oe_next[e]) {
      const int s = A.rank[A.oe_dst[e]] - beg;
      if (s >= n) continue;
      A.Ol[s] = A.Ol[i];
      A.Or[s] = A.Or[i];
    }
  }
  for (int e = A.out_head[u]; e != -1; e = A.oe_next[e]) {
    const int s = A.rank[A.oe_dst[e]] - beg;
    if (s >= n) continue;
    A.Ol[s] = min(A.Ol[s], A.Ol[i]);
    A.Or[s] = max(A.Or[s], A.Or[i]);
  }
  for (int e = A.out_head[u]; e != -1; e = A.oe_next[e]) {
    const int s = A.rank[A.oe_dst[e]] - beg;
    if (s >= n) continue;
    A.Pl[s] = min(A.Pl[s], max_acj);
    A.Pr[s] = max(A.Pr[s], max_acj + 1);
    A.Ow[s] = A.Ow[i] ? max(A.Ow[s], A.Ow[i]) : A.Ow[i];
  }
}

// LDS ring shape per workgroup size: RING rows of WR columns (int32 H and D).
template <int NT>
struct RingShape {
  // 64: four rows of 344 columns, 14.3 KB in all, so that eleven workgroups share a CU (the
  // DP is latency-bound: more windows in flight beat the predecessor reads the fifth to eighth
  // most recent rows would serve from LDS). 128/256: the larger batched tiles, sized so that
  // the sort of a graph of up to ~5,400 / ~8,000 nodes runs in LDS while leaving room for
  // small-tile workgroups on the same CU (39 / 63 KB). 512: one window alone (minipoa's own
  // run), wide bands.
  static constexpr int RING = NT == 64 ? 4 : RING_ROWS;
  static constexpr int WR = NT == 64 ? 344 : NT == 128 ? 512 : NT == 256 ? 896 : 2048;
  // longest query held in LDS (64: queries of up to 1,500 bases)
  static constexpr int QCAP = NT == 64 ? 1536 : NT <= 256 ? 4096 : 24576;
};

template <int NT, int CPL, bool QLDS, class T, class RingT>
__device__ __forceinline__ void poa_rows(const KernelArgs& a, Arena& A, WinShared& S, RingT& R,
                                         bool adaptive, int w, int full_len, int64_t t0);

// One poa() call on the device graph: rows [beg_i, end_i] of the current rank order. NARROW:
// the int16 form may be taken (batched windows; the other kernels keep int32 only).
template <int NT, int CPL, bool QLDS, bool NARROW, class RingT>
__device__ __forceinline__ void poa_call(const KernelArgs& a, Arena& A, WinShared& S, RingT& R, bool adaptive) {
  const DpScores& sc = a.scores;
  const int n = S.n_rows, m = S.m, beg = S.beg_i;
  const int tid = threadIdx.x;
  int w = sc.b;
  if (sc.f) w += m / sc.f;
  const int64_t t0 = now();
  const int full_len = (m / Q + 2) * Q;
  if (tid == 0) S.call_logical = 0;
  if (!adaptive) {
    if (tid == 0) S.row_wmax = 0;
    __syncthreads();
    if (!plan_static<NT>(A, S, sc, a.caps, w)) {
      if (tid == 0) S.status = MPOA_E_OVERFLOW;
      __syncthreads();
      return;
    }
  } else {
    // Full-width rows (align.cpp:1385-1390); adaptive state reset (align.cpp:1377).
    if ((int64_t)full_len * n > a.caps.dp) {
      if (tid == 0) S.status = MPOA_E_OVERFLOW;
      __syncthreads();
      return;
    }
    for (int r = tid; r < n; r += NT) {
      A.row_off[r] = r * full_len;
      A.Mp[r] = 0; A.Sl[r] = 0; A.Pl[r] = m; A.Pr[r] = 0; A.Ol[r] = 0; A.Or[r] = 0; A.Ow[r] = 0;
    }
    if (tid == 0) {
      A.row_off[n] = n * full_len;
      adaptive_band_row(A, S, 0, w);
    }
    __syncthreads();
  }
  // The int16 form when the call's static bound allows it: every path score is at most
  // match per query column consumed (gap scores are not positive, a mismatch scores at most a
  // match), so H, D and I stay at or below match * m; the lower side is checked on the
  // finite H values as they are computed (a finite D or I is at least a finite H plus the
  // gap-open score). Adaptive calls stay int32: a row without a finite cell picks its argmax
  // among minipoa's NEG_INF offsets, which the int16 form does not keep.
  if constexpr (NARROW && QLDS) {
    const bool narrow = !adaptive && !(a.forms & MPOA_FORM_WIDE) && sc.o1 <= 0 && sc.e1 <= 0 &&
                        sc.o1 >= -1024 && sc.match >= 0 && sc.mismatch <= sc.match &&
                        (int64_t)sc.match * m <= 32767;
    if (narrow) {
      poa_rows<NT, CPL, QLDS, int16_t>(a, A, S, R, adaptive, w, full_len, t0);
      return;
    }
  }
  poa_rows<NT, CPL, QLDS, int>(a, A, S, R, adaptive, w, full_len, t0);
}

// The DP rows, the score and the backtrack of one poa() call whose plan is filled, with
// H, D and I stored as T.
template <int NT, int CPL, bool QLDS, class T, class RingT>
__device__ __forceinline__ void poa_rows(const KernelArgs& a, Arena& A, WinShared& S, RingT& R,
                                         bool adaptive, int w, int full_len, int64_t t0) {
  constexpr int RING = RingT::kRing, WR = RingT::kWR;
  const DpScores& sc = a.scores;
  const int n = S.n_rows, m = S.m, beg = S.beg_i;
  const int tid = threadIdx.x;
  DpMats<T> X;
  X.H = reinterpret_cast<T*>(A.H);
  X.D = reinterpret_cast<T*>(A.D);
  X.I = reinterpret_cast<T*>(A.I);
  X.RH = reinterpret_cast<T*>(R.H);
  X.RD = reinterpret_cast<T*>(R.D);
  // columns of each of the ring's H and D arrays in the storage type
  constexpr int RCOLS = RING * WR * (int)(sizeof(int) / sizeof(T));
  // LDS row cache: slots of sw columns (H and D each), as many as fit the LDS arrays (at most
  // 64, one per lane of the directory), sw the widest row of the call (static band) or a
  // typical adaptive band. A row takes a slot when a later row reads it and it fits; it keeps
  // the slot until its last reader (the largest successor rank, in its row record) is done.
  // With no slot free, it takes the slot whose row is read last, if that is read after its
  // own last reader, and never a slot the row itself reads. Rows without a slot are read from
  // HBM.
  const int lane = tid & 63;
  int sw;
  if (!adaptive) {
    sw = S.row_wmax;
  } else {
    sw = min(full_len - Q, (2 * w + 128 + Q - 1) / Q * Q);
  }
  sw = max(sw, Q);
  const int ns = min(64, RCOLS / sw);
  int meta_row = -1, meta_bsbe = 0, meta_last = -1;
  int scan_ph = 0;  // half of S.scan the next row scan writes
  uint32_t fmin = 0xffffffffu;  // int16 form: smallest finite H of this thread, less FINITE_CUT
  // Row 0 (align.cpp:1415-1425): M, D = NEG_INF over blocks [Bs0, Be0) from the row start;
  // M[0][0] = 0. I of row 0 is never read; it is set to NEG_INF. Bs0 is always 0.
  {
    const int be0 = A.Be[0];
    const int len0 = be0 * Q, w0 = (be0 - 1) * Q;
    const int last0 = A.rowrec[12LL * beg + 11];
    T* H0 = X.H + A.row_off[0];
    T* D0 = X.D + A.row_off[0];
    T* I0 = X.I + A.row_off[0];
    const bool cache0 = ns > 0 && w0 <= sw && last0 > beg;
    const T zero = store_of<T>(0), ninf = store_of<T>(NEG_INF);
    if (tid == 0) S.fmin = 0xffffffffu;
    for (int j = tid; j < len0; j += NT) {
      H0[j] = j == 0 ? zero : ninf;
      D0[j] = ninf;
      I0[j] = ninf;
      if (cache0 && j < w0) {
        X.RH[j] = j == 0 ? zero : ninf;
        X.RD[j] = ninf;
      }
    }
    if (cache0 && lane == 0) {
      meta_row = 0;
      meta_bsbe = be0 << 16;
      meta_last = last0 - beg;
    }
  }
  __syncthreads();
  const int64_t t1 = now();
  if (tid == 0) S.ticks[PH_PLAN] += t1 - t0;
  int64_t cells = 0, ecells = 0;
  int wmax = 0;
  for (int i = 1; i < n; i++) {
    // Row records come through LDS, 32 rows at a time: one wait on HBM per chunk instead of
    // one per row (a load's wait on gfx9 also waits for every store issued before it), and
    // a second, shorter one (no store is pending then) for the chunk's fourth to sixth
    // predecessors.
    if ((i - 1) % 32 == 0) {
      lds_barrier<NT>();
      if (tid < 32 && i + tid < n) {
        const int4* src = reinterpret_cast<const int4*>(A.rowrec + 12LL * (beg + i + tid));
        const int4 ra = src[0], rb = src[1];
        int4 rc = src[2];
        if (ra.w > 3) {
          int x[3];
#pragma unroll
          for (int j = 0; j < 3; j++) {
            x[j] = 0;
            if (3 + j < ra.w)
              x[j] = ((A.pred_rank[ra.z + 3 + j] - beg) << 3) | (A.pred_base[ra.z + 3 + j] & 7);
          }
          rc.x = x[0];
          rc.y = x[1];
          rc.z = x[2];
        }
        R.rec[3 * tid] = ra;
        R.rec[3 * tid + 1] = rb;
        R.rec[3 * tid + 2] = rc;
      }
      lds_barrier<NT>();
    }
    const int ci = (i - 1) % 32;
    RowRec rr = rowrec_of(R.rec[3 * ci], R.rec[3 * ci + 1], R.rec[3 * ci + 2]);
    int bs, be;
    if (adaptive) {
      if (tid == 0) {
        adaptive_band_row(A, S, i, w);
        S.row_bs = A.Bs[i];
        S.row_be = A.Be[i];
      }
      lds_barrier<NT>();
      bs = S.row_bs;
      be = S.row_be;
      rr.row_off = i * full_len;
    } else {
      bs = rr.bsbe & 0xffff;
      be = rr.bsbe >> 16;
    }
    const int W = (be - bs - 1) * Q;
    if (W < Q) {
      if (tid == 0) S.status = MPOA_E_BAND;
      __syncthreads();
      return;
    }
    // Slots this row reads, and whether a predecessor comes from HBM: with several waves, its
    // stores must have completed first (one wave sees its own stores in order).
    uint64_t pmask = 0;
    bool far = false;
    {
      const int n3 = min(rr.np, 3);
      for (int k = 0; k < n3; k++) {
        const int p = (k == 0 ? rr.p0 : k == 1 ? rr.p1 : rr.p2) - beg;
        if (p < 0) continue;
        const int sl = cache_slot(meta_row, p);
        if (sl >= 0) pmask |= 1ull << sl;
        else far = true;
      }
      if (rr.np > 3) {
        const int n6 = min(rr.np, 6);
        for (int k = 3; k < n6; k++) {
          const int p = (k == 3 ? rr.w0 : k == 4 ? rr.w1 : rr.w2) >> 3;
          if (p < 0) continue;
          const int sl = cache_slot(meta_row, p);
          if (sl >= 0) pmask |= 1ull << sl;
          else far = true;
        }
        if (NT > 64) {
          for (int k = 6; k < rr.np; k++) {
            const int p = A.pred_rank[rr.k0 + k] - beg;
            drain_vmem();
            if (p >= 0) far |= cache_slot(meta_row, p) < 0;
          }
        }
      }
      if (NT > 64 && far) __syncthreads();
    }
    // A slot for this row: a free one, else (predecessors after the sixth unknown here: no
    // eviction then) the one whose row is read last.
    const int last_i = rr.last < 0 ? -1 : rr.last - beg;
    int slot_i = -1;
    if (W <= sw && last_i > i) {
      const uint64_t freem = __ballot(lane < ns && meta_last < i);
      if (freem) {
        slot_i = (int)__builtin_ctzll(freem);
      } else if (rr.np <= 6) {
        // the latest last reader among the slots this row does not read: a DPP wave maximum
        // (shuffles would be six LDS round trips on every row once the cache is full)
        const int v = (lane < ns && !((pmask >> lane) & 1)) ? meta_last : -1;
        const int mx = __builtin_amdgcn_readlane(wave_incl_max(v), 63);
        if (mx > last_i) slot_i = (int)__builtin_ctzll(__ballot(v == mx));
      }
    }
    int np_in = 0;
    const int mj = dp_row<NT, CPL, QLDS>(A, S, R, X, sc, i, rr, bs, be, adaptive, &np_in,
                                         meta_row, meta_bsbe, sw, slot_i, scan_ph, fmin);
    if (slot_i >= 0 && lane == slot_i) {
      meta_row = i;
      meta_bsbe = bs | (be << 16);
      meta_last = last_i;
    }
    lds_barrier<NT>();
    cells += W;
    ecells += (int64_t)np_in * W;
    wmax = max(wmax, W);
    if (adaptive) {
      if (tid == 0) adaptive_update(A, S, i, mj, w);
      lds_barrier<NT>();
    }
  }
  if constexpr (sizeof(T) == 2) {
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) fmin = min(fmin, (uint32_t)__shfl_xor((int)fmin, s, 64));
    if (lane == 0) atomicMin(&S.fmin, fmin);
  }
  __syncthreads();
  const int64_t t2 = now();
  if (tid == 0) {
    S.ticks[PH_DP] += t2 - t1;
    S.cells += cells;
    S.edge_cells += ecells;
    S.call_cells = cells;
    S.call_ecells = ecells;
    S.logical += S.call_logical;
    S.call_width = wmax;
    S.score = widen(X.H[A.row_off[n - 1] + (m - A.Bs[n - 1] * Q)]);
    S.narrow_bad = false;
    if constexpr (sizeof(T) == 2) {
      // Every finite H at least -32767 - o1 keeps every finite D and I at -32767 or above,
      // so no finite value was stored as N16; a NEG_INF-derived end cell would send the
      // backtrack through minipoa's NEG_INF offsets. Either failing: the window is re-run
      // in int32.
      const int64_t lo = (int64_t)S.fmin + FINITE_CUT;
      S.narrow_bad = S.fmin == 0xffffffffu || lo < -32767LL - sc.o1 || S.score <= FINITE_CUT;
      if (!S.narrow_bad) S.narrow_cells += cells;
    }
  }
  __syncthreads();
  // The backtrack's move log goes to LDS (the ring's rows are dead now) when the query fits: a
  // global store there would make the walk's next loads wait for it as well.
  int* log = m <= 2 * RING * WR ? reinterpret_cast<int*>(&R) : A.res_from;  // H, D first
  // Each row's heaviest in-window predecessor, for the backtrack's chain runs, after the log in
  // LDS when both fit there (every thread, one row each).
  int* hp = nullptr;
  {
    const int at = (m + 63) / 64 * 64;
    if (log != A.res_from && at + n <= 2 * RING * WR) hp = reinterpret_cast<int*>(&R) + at;
  }
  if (hp && !(a.forms & MPOA_FORM_BT_RANKS)) {
    for (int r = tid; r < n; r += NT) hp[r] = heaviest_pred(A, beg, r);
    __syncthreads();
  } else {
    hp = nullptr;
  }
  if (__builtin_amdgcn_readfirstlane(tid >> 6) == 0) {  // wave 0 (a wave-uniform branch)
    const int rc = S.narrow_bad ? MPOA_E_OVERFLOW
                   : (a.forms & MPOA_FORM_BT_LANE) ? 1
                                                   : backtrack_wave(A, S, sc, X, log, hp);
    if (tid == 0) {
      // A row with more than 63 predecessors: the one-lane backtrack, which writes the
      // res_t entries itself.
      S.bt_serial = rc == 1;
      S.bt_rc = rc == 1 ? backtrack(A, S, sc, X) : rc;
    }
  }
  __syncthreads();
  if (!S.bt_serial && S.bt_rc == MPOA_OK) resolve_res<NT>(A, S, log);
  if (tid == 0) {
    if (S.bt_rc != MPOA_OK) S.status = S.bt_rc;
    S.ticks[PH_BT] += now() - t2;
  }
  __syncthreads();
}

// Query slice src[0..qlen): sentinel N, the codes, N padding to col_size
// (align.cpp:1362-1373); plain in A.qseq and shifted one column right in A.qsh and, when it
// fits, R.q (entry c is the character column c consumes in the M-source, seq[c-1]).
template <int NT, class RingT>
__device__ __forceinline__ void load_query(Arena& A, WinShared& S, RingT& R, const uint8_t* src, int qlen) {
  const int tid = threadIdx.x;
  const int m = qlen + 1;
  const int col_size = (m + 1 + Q - 1) / Q * Q;
  const bool q_lds = col_size + 2 * Q <= RingT::kQ;
  for (int j = tid; j < col_size + 2 * Q; j += NT) {
    const uint8_t cq = j >= 1 && j <= qlen ? (uint8_t)code_of(src[j - 1]) : (uint8_t)CODE_N;
    const uint8_t cs = j >= 2 && j <= qlen + 1 ? (uint8_t)code_of(src[j - 2]) : (uint8_t)CODE_N;
    A.qseq[j] = cq;
    A.qsh[j] = cs;
    if (q_lds) R.q[j] = cs;
  }
  if (tid == 0) {
    S.qlen = qlen;
    S.m = m;
    S.q_lds = q_lds;
    S.q_cap = col_size + 2 * Q;
  }
  __syncthreads();
}

// First read: the source, then one insertion per base (align.cpp:1747-1753).
template <int NT>
__device__ __forceinline__ void trivial_res(Arena& A, int m) {
  for (int j = threadIdx.x; j < m; j += NT) {
    A.res_from[j] = j == 0 ? 0 : -1;
    A.res_base[j] = j == 0 ? (uint8_t)CODE_N : A.qseq[j];
    A.res_aligned[j] = -1;
  }
  __syncthreads();
}

// graph::build_consensus (graph.cpp:294-350) for -S seeding, into the consensus arrays: every
// node keeps the predecessor of largest in-weight, ties to the larger accumulated score, then
// the first in its in-list; the path back from the sink, in order, with each node's rank.
// minipoa visits the nodes in FIFO order, but a node's choice reads only its predecessors' final
// scores, so any topological order gives the same path: here the rank order of the sort just
// done, with the DP plan's predecessor CSR, by wave 0 in blocks of 64 ranks. The choice, the
// largest (weight, score) and then the lowest in-list position, does not depend on the order
// the candidates are met in, so each lane first reduces its rank's predecessors in earlier
// blocks (final scores in t2, read with device-coherent loads after each block's stores are
// waited for) to one candidate and keeps those inside the block (up to HR_IN: two in registers,
// the rest in LDS). Ranks with none inside settle at once; the others settle one at a time in
// ascending rank order, with the block's scores in one register across the lanes (readlane /
// writelane). A block holding a rank with more in-block predecessors settles in rounds instead,
// each lane re-reading its in-list until all of its in-block predecessors have settled. Choices
// go to t1; the path goes to t0 and is turned into the consensus arrays by all threads.
constexpr int HR_IN = 8;

template <int NT>
__device__ __forceinline__ void heaviest_ranked(Arena& A, WinShared& S, uint8_t* lds, int lds_bytes,
                                                int forms) {
  const int n = S.n_rows;
  int* score = A.t2;
  int* pre = A.t1;
  int* path = A.t0;
  int* bval = reinterpret_cast<int*>(lds);  // rounds: this block's scores, INT_MIN until settled
  int2* bin = reinterpret_cast<int2*>(bval + 64);  // in-block predecessors past the first two
  const int cap = 2 + max(0, min(HR_IN - 2, (lds_bytes - 256) / (64 * 8)));
  const int xcap = cap - 2;
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) {
    const int lane = threadIdx.x & 63;
    for (int b = 0; b < n; b += 64) {
      const int r = b + lane;
      const bool act = r < n;
      const int k0 = act ? A.pred_off[r] : 0, k1 = act ? A.pred_off[r + 1] : 0;
      int done = !act || A.rank_id[r] == 0;  // the source scores 0
      // the best earlier-block predecessor (in-list position relative to k0); in-block ones as
      // (offset | in-list position << 6, w)
      int ew = -1, ep = -1, eps = 0, ek = INT_MAX, nin = 0;
      int2 e0 = make_int2(0, 0), e1 = make_int2(0, 0);
      bool reread = false;
      if (!done) {
        // the first four predecessors' loads are issued together, then their scores'
        constexpr int PF = 4;
        int pp[PF], pw[PF], ps4[PF];
#pragma unroll
        for (int i = 0; i < PF; i++) {
          pp[i] = k0 + i < k1 ? A.pred_rank[k0 + i] : INT_MAX;
          pw[i] = k0 + i < k1 ? A.pred_w[k0 + i] : 0;
        }
#pragma unroll
        for (int i = 0; i < PF; i++)
          ps4[i] = pp[i] < b ? __hip_atomic_load(score + pp[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        for (int k = k0; k < k1; k++) {
          const int i = k - k0;
          int p, w, ps = 0;
          if (i < PF) {
            p = pp[0]; w = pw[0]; ps = ps4[0];
#pragma unroll
            for (int t = 1; t < PF; t++)
              if (i == t) { p = pp[t]; w = pw[t]; ps = ps4[t]; }
          } else {
            p = A.pred_rank[k];
            w = A.pred_w[k];
            if (p < b) ps = __hip_atomic_load(score + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          if (p < b) {
            if (w > ew || (w == ew && ps > eps)) {
              ew = w; ep = p; eps = ps; ek = i;
            }
          } else if (nin < cap) {
            const int2 e = make_int2((p - b) | i << 6, w);
            if (nin == 0) e0 = e;
            else if (nin == 1) e1 = e;
            else bin[lane * xcap + nin - 2] = e;
            nin++;
          } else {
            reread = true;
          }
        }
      }
      if (forms & MPOA_FORM_CONS_ROUNDS) reread = true;
      int sc = 0, pr = -1;
      if (!done && nin == 0) {
        sc = eps + ew;
        pr = ep;
        done = 1;
      }
      if (!__any(reread)) {
        uint64_t todo = __ballot(!done);
        while (todo) {
          const int j = __builtin_ctzll(todo);
          todo &= todo - 1;
          int bw = __builtin_amdgcn_readlane(ew, j), bp = __builtin_amdgcn_readlane(ep, j);
          int bps = __builtin_amdgcn_readlane(eps, j), bk = __builtin_amdgcn_readlane(ek, j);
          const int cnt = __builtin_amdgcn_readlane(nin, j);
          for (int c = 0; c < cnt; c++) {
            int ex, ey;
            if (c == 0) {
              ex = __builtin_amdgcn_readlane(e0.x, j);
              ey = __builtin_amdgcn_readlane(e0.y, j);
            } else if (c == 1) {
              ex = __builtin_amdgcn_readlane(e1.x, j);
              ey = __builtin_amdgcn_readlane(e1.y, j);
            } else {
              const int2 e = bin[j * xcap + c - 2];
              ex = __builtin_amdgcn_readfirstlane(e.x);
              ey = __builtin_amdgcn_readfirstlane(e.y);
            }
            const int o = ex & 63, k = ex >> 6;
            const int ps = __builtin_amdgcn_readlane(sc, o);
            if (ey > bw || (ey == bw && (ps > bps || (ps == bps && k < bk)))) {
              bw = ey; bp = b + o; bps = ps; bk = k;
            }
          }
          if (lane == j) {
            sc = bps + bw;
            pr = bp;
          }
        }
      } else {
        bval[lane] = done ? sc : INT_MIN;
        __builtin_amdgcn_wave_barrier();
        for (int round = 0; round < 64 && !__all(done); round++) {
          if (!done) {
            bool ready = true;
            int bw = -1, bp = -1, bps = 0;
            for (int k = k0; k < k1; k++) {
              const int p = A.pred_rank[k], w = A.pred_w[k];
              int ps;
              if (p < b) {
                ps = __hip_atomic_load(score + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              } else {
                ps = bval[p - b];
                if (ps == INT_MIN) {
                  ready = false;
                  break;
                }
              }
              if (w > bw || (w == bw && ps > bps)) {
                bw = w; bp = p; bps = ps;
              }
            }
            if (ready) {
              sc = bps + bw;
              pr = bp;
              done = 1;
            }
          }
          bval[lane] = done ? sc : INT_MIN;  // every lane's reads of this round come first
          __builtin_amdgcn_wave_barrier();
        }
      }
      if (act) {
        score[r] = sc;
        pre[r] = pr;
      }
      drain_vmem();
    }
    // The path back from the sink, in descending ranks (a choice always has a lower rank): the
    // wave loads the choices and node ids of the 64-rank block holding the current rank at once
    // and follows the path inside it across the lanes.
    int len = 0;
    int x = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(pre + A.rank[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    while (x >= 0 && x < n && len < n) {  // bounded even if a choice were not of lower rank
      const int b = x & ~63;
      const bool in = b + lane < n;
      const int pv = in ? __hip_atomic_load(pre + b + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : -1;
      const int id = in ? A.rank_id[b + lane] : 0;
      while (x >= b && x < b + 64 && len < n) {
        if (__builtin_amdgcn_readlane(id, x - b) == 0) {  // the source
          x = -1;
          break;
        }
        if (lane == 0) path[len] = x;
        len++;
        x = __builtin_amdgcn_readlane(pv, x - b);
      }
    }
    if (lane == 0) S.cons_len = len;
  }
  __syncthreads();
  const int len = S.cons_len;
  for (int at = threadIdx.x; at < len; at += NT) {
    const int x = path[len - 1 - at];
    const int id = A.rank_id[x];
    A.cons_id[at] = id;
    A.cons_ch[at] = A.base[id];
    A.cons_rank[at] = x;
  }
  __syncthreads();
}

// add_path then the re-sort of one read (main.cpp:118-120): op=1 after the last read, else
// op=0 (with hlen/tlen for the static or adaptive band) and, when seeding, the consensus
// the next read is seeded against. lds / lds_bytes: LDS scratch for the sort.
template <int NT>
__device__ __forceinline__ void commit_read(const KernelArgs& a, Arena& A, WinShared& S, int r, int m,
                            int path0, bool last, bool seeding, uint8_t* lds, int lds_bytes) {
  const int tid = threadIdx.x;
  const bool keep_path = a.scores.result != 0;
  const bool with_len = a.scores.f > 0;
  if (keep_path) {
    // positions add_path does not record (a repeated node) print as node 1 - 1 = 0, as
    // output_gfa's zero-initialised path does (graph.cpp:365)
    for (int j = tid; j < m - 1; j += NT) A.path[path0 + j] = 1;
    __syncthreads();
  }
  {
    const int64_t ta = now();
    if ((a.forms & MPOA_FORM_ADD_LANE) || !add_path_parallel<NT>(A, S, a.caps, keep_path, r, m, path0, A.t2, A.t1)) {
      if (tid == 0) add_path(A, S, a.caps, keep_path, r, m, path0);
    }
    __syncthreads();
    if (tid == 0) S.ticks[PH_ADD] += now() - ta;
  }
  if (S.status != MPOA_OK) return;
  const int64_t tb = now();
  if (last) {
    // the aligned-group order of the last read (op=1) only places the MSA columns and the GFA
    // records; the consensus output does not read it
    if (tid == 0 && a.scores.result != 0) S.msa_cols = topsort1(A, S);
  } else {
    topsort0<NT>(A, S, a.caps, with_len, lds, lds_bytes, a.forms);
    if (seeding) heaviest_ranked<NT>(A, S, lds, lds_bytes, a.forms);
  }
  if (tid == 0) S.ticks[PH_SORT] += now() - tb;
  __syncthreads();
}

__device__ __forceinline__ void put_str(uint8_t* o, int64_t& at, const char* t) {
  while (*t) o[at++] = (uint8_t)*t++;
}
__device__ __forceinline__ void put_uint(uint8_t* o, int64_t& at, unsigned v) {
  char d[10];
  int n = 0;
  do {
    d[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) o[at++] = (uint8_t)d[--n];
}

// graph::output_gfa (graph.cpp:358-433), one lane: segments and links in FIFO BFS order from
// the source (a node's links in its in-list order, the source's left out), then one path line
// per read named by its 1-based index. Segment ids are node ids - 1. Returns the text length,
// or -1 when the capacity is short. line_at (optional) receives each path line's offset.
__device__ int64_t gfa_text(const KernelArgs& a, Arena& A, WinShared& S, const WinDesc& wd) {
  uint8_t* o = a.out + wd.out0;
  const int64_t cap = wd.out_cap;
  const int N = S.n_nodes;
  if (N <= 2) return 0;
  int64_t at = 0;
  put_str(o, at, "H\tVN:Z:1.0\n");
  int* deg = A.t0;
  int* queue = A.t1;
  for (int v = 0; v < N; v++) deg[v] = A.in_cnt[v];
  int qh = 0, qt = 0;
  queue[qt++] = 0;
  while (qh < qt) {
    const int cur = queue[qh++];
    if (cur == 1) break;
    if (cur != 0) {
      if (at + 32 > cap) return -1;
      put_str(o, at, "S\t");
      put_uint(o, at, (unsigned)(cur - 1));
      o[at++] = '\t';
      o[at++] = char_of(A.base[cur]);
      o[at++] = '\n';
      for (int e = A.in_head[cur]; e != -1; e = A.ie_next[e]) {
        const int pre = A.ie_src[e];
        if (pre == 0) continue;
        if (at + 48 > cap) return -1;
        put_str(o, at, "L\t");
        put_uint(o, at, (unsigned)(pre - 1));
        put_str(o, at, "\t+\t");
        put_uint(o, at, (unsigned)(cur - 1));
        put_str(o, at, "\t+\t0M\n");
      }
    }
    for (int e = A.out_head[cur]; e != -1; e = A.oe_next[e]) {
      const int t = A.oe_dst[e];
      if (--deg[t] == 0) queue[qt++] = t;
    }
  }
  int64_t p0 = 0;
  for (int r = 0; r < wd.n_reads; r++) {
    const int qlen = (int)(a.read_off[wd.read0 + r + 1] - a.read_off[wd.read0 + r]);
    if (a.gfa_lines) a.gfa_lines[r] = at;
    if (at + 32 > cap) return -1;
    put_str(o, at, "P\t");
    put_uint(o, at, (unsigned)(r + 1));
    o[at++] = '\t';
    for (int j = 0; j < qlen; j++) {
      if (at + 16 > cap) return -1;
      put_uint(o, at, (unsigned)(A.path[p0 + j] - 1));
      o[at++] = '+';
      if (j != qlen - 1) o[at++] = ',';
      else put_str(o, at, "\t*\n");
    }
    p0 += qlen;
  }
  if (a.gfa_lines) a.gfa_lines[wd.n_reads] = at;
  return at;
}

// Output text of a finished window and its result record.
template <int NT>
__device__ __forceinline__ void window_outputs(const KernelArgs& a, Arena& A, WinShared& S, int w,
                                               uint8_t* lds, int lds_bytes) {
  const WinDesc& wd = a.win[w];
  const int tid = threadIdx.x;
  const int64_t t_out = now();
  if (S.status == MPOA_OK) {
    if (a.scores.result == 0) {
      // graph::output_consensus (graph.cpp:294-357): the last read left the graph unsorted, so
      // the wave sort gives the rank order and predecessor CSR the blocked heaviest-bundle walk
      // reads (any topological order gives minipoa's FIFO result), then the bases are written.
      topsort0<NT>(A, S, a.caps, false, lds, lds_bytes, a.forms);
      heaviest_ranked<NT>(A, S, lds, lds_bytes, a.forms);
      const int len = S.cons_len;
      if (len > wd.out_cap) {
        if (tid == 0) S.status = MPOA_E_OVERFLOW;
      } else {
        uint8_t* o = a.out + wd.out0;
        for (int at = tid; at < len; at += NT) o[at] = char_of(A.cons_ch[at]);
      }
      __syncthreads();
    } else if (a.scores.result == 1 || a.scores.result == 3) {
      // Row-column MSA (graph.cpp:237-251): rows = reads, columns = group ranks 1..R-2. With
      // result 3 one more row follows: the consensus (graph.cpp:294-357) with each base in
      // its node's column.
      const int R = S.msa_cols;
      const int L = R >= 2 ? R - 2 : 0;
      const bool cons_row = a.scores.result == 3;
      const int64_t need = (int64_t)(wd.n_reads + (cons_row ? 1 : 0)) * L;
      if (need > wd.out_cap) {
        if (tid == 0) S.status = MPOA_E_OVERFLOW;
      } else {
        uint8_t* o = a.out + wd.out0;
        for (int64_t x = tid; x < need; x += NT) o[x] = '-';
        __syncthreads();
        int64_t p0 = 0;
        for (int r = 0; r < wd.n_reads; r++) {
          const int qlen = (int)(a.read_off[wd.read0 + r + 1] - a.read_off[wd.read0 + r]);
          for (int pos = tid; pos < qlen; pos += NT) {
            const int v = A.path[p0 + pos];
            if (v >= 2) {
              const int rk = A.rank[A.par[v]];
              if (rk >= 1 && rk <= L) o[(int64_t)r * L + rk - 1] = char_of(A.base[v]);
            }
          }
          p0 += qlen;
        }
        if (cons_row) {
          // the columns are the aligned-group ranks of the op 1 sort; the consensus needs the
          // op 0 order of the final graph, which overwrites them: keep them in A.Mp (DP
          // scratch) first
          for (int v = tid; v < S.n_nodes; v += NT) A.Mp[v] = A.rank[v];
          __syncthreads();
          topsort0<NT>(A, S, a.caps, false, lds, lds_bytes, a.forms);
          heaviest_ranked<NT>(A, S, lds, lds_bytes, a.forms);
          uint8_t* row = o + (int64_t)wd.n_reads * L;
          for (int x = tid; x < L; x += NT) row[x] = '-';
          __syncthreads();
          for (int at = tid; at < S.cons_len; at += NT) {
            const int id = A.cons_id[at];
            const int rk = A.Mp[A.par[id]];
            if (rk >= 1 && rk <= L) row[rk - 1] = char_of(A.base[id]);
          }
        }
        __syncthreads();
        if (tid == 0) S.cons_len = L;
      }
      __syncthreads();
    } else if (a.scores.result == 2) {
      if (tid == 0) {
        const int64_t len = gfa_text(a, A, S, wd);
        if (len < 0 || len > INT32_MAX) S.status = MPOA_E_OVERFLOW;
        else S.cons_len = (int)len;
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    S.ticks[PH_OUT] += now() - t_out;
    WinResult& res = a.result[w];
    for (int k = 0; k < PH_N; k++) res.ticks[k] = S.ticks[k];
    res.t_end = now();
    res.status = S.status;
    res.out_len = S.status == MPOA_OK ? S.cons_len : 0;
    res.cells = S.cells;
    res.edge_cells = S.edge_cells;
    res.logical = S.logical;
    res.narrow_cells = S.narrow_cells;
    res.nodes = S.n_nodes;
    res.edges = S.n_in_e;
  }
  __syncthreads();
}

template <int NT>
__device__ __forceinline__ void window_begin(const KernelArgs& a, Arena& A, WinShared& S, int w) {
  if (threadIdx.x == 0) {
    a.result[w].t_begin = now();
    S.status = MPOA_OK;
    S.n_nodes = 0; S.n_in_e = 0; S.n_out_e = 0;
    S.cells = 0; S.edge_cells = 0; S.logical = 0; S.narrow_cells = 0;
    S.cons_len = 0;
    for (int k = 0; k < PH_N; k++) S.ticks[k] = 0;
    new_node(A, S, a.caps, CODE_N);  // source 0 (graph.cpp:129-135)
    new_node(A, S, a.caps, CODE_N);  // sink 1
  }
  __syncthreads();
}

// The whole per-window pipeline (no seeding): every read against the whole graph.
template <int NT, int CPL, class RingT>
__device__ __forceinline__ void run_window(const KernelArgs& a, Arena& A, WinShared& S, RingT& R, int w) {
  const WinDesc& wd = a.win[w];
  const int tid = threadIdx.x;
  const bool keep_path = a.scores.result != 0;
  window_begin<NT>(a, A, S, w);
  int path0 = 0;
  for (int r = 0; r < wd.n_reads; r++) {
    const int64_t q0 = a.read_off[wd.read0 + r], q1 = a.read_off[wd.read0 + r + 1];
    const int qlen = (int)(q1 - q0);
    const int m = qlen + 1;
    if (qlen > a.caps.qmax || (keep_path && path0 + qlen > a.caps.path)) {
      if (tid == 0) S.status = MPOA_E_OVERFLOW;
      __syncthreads();
      break;
    }
    load_query<NT>(A, S, R, a.read_seq + q0, qlen);
    if (tid == 0) S.read = r;
    if (S.n_nodes <= 2) {
      trivial_res<NT>(A, m);
    } else {
      if (tid == 0) {
        S.beg_i = A.rank[0];
        S.end_i = A.rank[1];
        S.n_rows = S.end_i - S.beg_i + 1;
      }
      __syncthreads();
      const bool adaptive = a.scores.ab_band != 0 && a.scores.f > 0;
      if (S.q_lds) poa_call<NT, CPL, true, true>(a, A, S, R, adaptive);
      else poa_call<NT, CPL, false, true>(a, A, S, R, adaptive);
      if (tid == 0 && a.call_stats && a.call_off[w] + r - 1 < a.call_off[w + 1]) {
        CallStat& cs = a.call_stats[a.call_off[w] + r - 1];
        cs.score = S.score;
        cs.rows = S.n_rows;
        cs.qlen = qlen;
        cs.width = S.call_width;
        cs.cells = S.call_cells;
        cs.edge_cells = S.call_ecells;
        cs.logical = S.call_logical;
        cs.nodes = S.n_nodes;
        cs.edges = S.n_in_e;
      }
    }
    __syncthreads();
    if (S.status != MPOA_OK) break;
    // The sort's LDS scratch is the whole ring: the query and row records are rewritten
    // for every read.
    commit_read<NT>(a, A, S, r, m, path0, r + 1 == wd.n_reads, false,
                    reinterpret_cast<uint8_t*>(&R), (int)sizeof(R));
    path0 += qlen;
    __syncthreads();
    if (S.status != MPOA_OK) break;
  }
  window_outputs<NT>(a, A, S, w, reinterpret_cast<uint8_t*>(&R), (int)sizeof(R));
}

// ---------------------------------------------------------------------------------------
// Seeding (-S) windows: alignment() splits each read at the anchors of its chain against the
// consensus (align.cpp:1763-1827); the sub-windows are independent DP calls on the same
// graph. The host seeds each read; k_segments runs every sub-window of the read in its own
// workgroup, and k_commit splices the anchor bases, adds the path, re-sorts and builds the
// consensus the next read is seeded against. The graph stays in slot 0 of the graph pools,
// and the window state between launches in SeedArgs::state.
// ---------------------------------------------------------------------------------------

#include "seed_dev.h"

// poa() on a -S sub-window of query length -1 (two anchors overlapping by one base in the
// read, align.cpp:1771-1790): minipoa runs its DP with m = 0, a band of one block of N columns
// on every row (Ms 0, Me 0), and its backtrack from column 0 takes deletions only, so the
// call adds no res entry. Its score is H[n-1][0]: row 0 has H 0, and every later row takes the
// largest of its predecessors' H + o1 and D + e1 (the D source, align.cpp:1484-1506) and of
// NEG_INF plus the match or mismatch score against the column's N (the M source's block-0
// diagonal, :1467-1483). One lane; h and d hold n ints each.
__device__ int null_call_score(const Arena& A, const DpScores& sc, int beg, int n, int* h,
                               int* d) {
  h[0] = 0;
  d[0] = NEG_INF;
  for (int i = 1; i < n; i++) {
    const RowRec rr = load_rowrec(A.rowrec, beg + i);
    int hm = NEG_INF, dv = NEG_INF;
    for (int k = 0; k < rr.np; k++) {
      const int prank = k == 0 ? rr.p0 : k == 1 ? rr.p1 : k == 2 ? rr.p2 : A.pred_rank[rr.k0 + k];
      const int p = prank - beg;
      if (p < 0) continue;
      const int b = k < 3 ? (rr.pk >> (3 * k)) & 7 : A.pred_base[rr.k0 + k] & 7;
      const int pbase = p == 0 ? CODE_N : b;
      hm = max(hm, NEG_INF + (pbase == CODE_N ? sc.match : sc.mismatch));
      dv = max(dv, max(d[p] + sc.e1, h[p] + sc.o1));
    }
    d[i] = dv;
    h[i] = max(max(hm, dv), NEG_INF);
  }
  return h[n - 1];
}

__device__ void state_load(WinShared& S, const WinState& st) {
  S.n_nodes = st.n_nodes; S.n_in_e = st.n_in_e; S.n_out_e = st.n_out_e;
  S.status = st.status; S.msa_cols = st.msa_cols; S.cons_len = st.cons_len;
  S.cells = st.cells; S.edge_cells = st.edge_cells; S.logical = st.logical;
  for (int k = 0; k < PH_N; k++) S.ticks[k] = st.ticks[k];
}
__device__ void state_store(const WinShared& S, WinState& st) {
  st.n_nodes = S.n_nodes; st.n_in_e = S.n_in_e; st.n_out_e = S.n_out_e;
  st.status = S.status; st.msa_cols = S.msa_cols; st.cons_len = S.cons_len;
  st.cells = S.cells; st.edge_cells = S.edge_cells; st.logical = S.logical;
  for (int k = 0; k < PH_N; k++) st.ticks[k] = S.ticks[k];
}

template <int NT, int CPL>
__global__ __launch_bounds__(NT) void k_segments(SeedArgs s) {
  __shared__ WinShared S;
  __shared__ Ring<RingShape<NT>::RING, RingShape<NT>::WR, RingShape<NT>::QCAP> R;
  using RingT = Ring<RingShape<NT>::RING, RingShape<NT>::WR, RingShape<NT>::QCAP>;
  const int tid = threadIdx.x;
  if (s.use_dev && (int)blockIdx.x >= s.dev.n_seg[s.read]) return;
  const SegDesc d = s.segs[blockIdx.x];
  const Arena G = carve(s.graph, 0);
  Arena A = carve(s.seg, blockIdx.x);
  if (d.dp_len > 0) {  // the sub-window's own share of the pool (device seeding)
    A.H = s.seg.dp_pool + d.dp0;
    A.D = A.H + d.dp_len;
    A.I = A.D + d.dp_len;
  }
  // The graph is read-only here; the DP plan, band state, query and matrices are private.
  A.par = G.par; A.aligned = G.aligned; A.ind = G.ind; A.rank = G.rank;
  A.in_head = G.in_head; A.in_tail = G.in_tail; A.out_head = G.out_head;
  A.out_tail = G.out_tail; A.in_cnt = G.in_cnt; A.last_read = G.last_read;
  A.out_cnt = G.out_cnt; A.ie_src = G.ie_src; A.ie_w = G.ie_w; A.ie_next = G.ie_next;
  A.oe_dst = G.oe_dst; A.oe_w = G.oe_w; A.oe_next = G.oe_next; A.rank_id = G.rank_id;
  A.hlen = G.hlen; A.tlen = G.tlen; A.pred_off = G.pred_off; A.pred_rank = G.pred_rank;
  A.pred_w = G.pred_w; A.pred_base = G.pred_base; A.base = G.base;
  A.res_from = G.res_from + d.res0;
  A.res_aligned = G.res_aligned + d.res0;
  A.res_base = G.res_base + d.res0;
  SegResult& out = s.seg_res[blockIdx.x];
  if (tid == 0) {
    S.status = MPOA_OK;
    S.n_nodes = s.state->n_nodes;
    S.n_in_e = s.state->n_in_e;
    S.cells = 0;
    S.edge_cells = 0;
    for (int k = 0; k < PH_N; k++) S.ticks[k] = 0;
    S.beg_i = G.rank[d.beg_id];
    S.end_i = G.rank[d.end_id];
    S.n_rows = S.end_i - S.beg_i + 1;
    S.read = s.read;
  }
  __syncthreads();
  if (S.n_rows <= 0 || d.qlen + 1 > s.seg.caps.qmax + 1 || S.n_rows > s.seg.caps.nodes) {
    if (tid == 0) {
      out.status = S.n_rows <= 0 ? MPOA_E_EMPTY_WINDOW : MPOA_E_OVERFLOW;
      out.score = 0; out.rows = S.n_rows; out.qlen = d.qlen;
      out.cells = 0; out.edge_cells = 0; out.width = 0; out.logical = 0;
    }
    return;
  }
  // Row records of the window's ranks: the predecessor fields come from the graph; the
  // DP plan overwrites its own fields in this private copy.
  for (int e = tid; e < 12 * S.n_rows; e += NT) {
    const int64_t k = 12LL * S.beg_i + e;
    A.rowrec[k] = G.rowrec[k];
  }
  __syncthreads();
  if (d.qlen == -1) {  // anchors overlapping by one base: no DP, no res entry
    if (tid == 0) {
      const int64_t cap = d.dp_len > 0 ? 3 * d.dp_len : 3 * s.seg.caps.dp;
      out.status = 2LL * S.n_rows <= cap ? MPOA_OK : MPOA_E_OVERFLOW;
      out.score = out.status == MPOA_OK
                      ? null_call_score(A, s.graph.scores, S.beg_i, S.n_rows, A.H, A.H + S.n_rows)
                      : 0;
      out.rows = S.n_rows; out.qlen = -1; out.width = 0;
      out.cells = 0; out.edge_cells = 0; out.logical = 0;
      for (int k = 0; k < PH_N; k++) out.ticks[k] = 0;
    }
    return;
  }
  const WinDesc& wd = s.graph.win[0];
  const int64_t q0 = s.graph.read_off[wd.read0 + s.read] + d.q0;
  load_query<NT>(A, S, R, s.graph.read_seq + q0, d.qlen);
  KernelArgs a = s.seg;
  a.scores = s.graph.scores;
  if (d.dp_len > 0) a.caps.dp = d.dp_len;
  // -S always uses the adaptive band (align.cpp:1780, 1808, 1826).
  const bool adaptive = a.scores.f > 0;
  if (S.q_lds) poa_call<NT, CPL, true, false, RingT>(a, A, S, R, adaptive);
  else poa_call<NT, CPL, false, false, RingT>(a, A, S, R, adaptive);
  if (tid == 0) {
    out.status = S.status;
    out.score = S.score;
    out.rows = S.n_rows;
    out.qlen = d.qlen;
    out.width = S.call_width;
    out.cells = S.call_cells;
    out.edge_cells = S.call_ecells;
    out.logical = S.call_logical;
    for (int k = 0; k < PH_N; k++) out.ticks[k] = S.ticks[k];
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_commit(SeedArgs s) {
  __shared__ WinShared S;
  __shared__ int4 lds4[(156 * 1024) / 16];
  const int tid = threadIdx.x;
  const KernelArgs& a = s.graph;
  Arena A = carve(a, 0);
  const WinDesc& wd = a.win[0];
  const int r = s.read;
  if (r == 0) {
    window_begin<NT>(a, A, S, 0);
  } else {
    if (tid == 0) state_load(S, *s.state);
    __syncthreads();
  }
  const int64_t q0 = a.read_off[wd.read0 + r], q1 = a.read_off[wd.read0 + r + 1];
  const int qlen = (int)(q1 - q0), m = qlen + 1;
  const bool keep_path = a.scores.result != 0;
  const int n_seg = s.use_dev ? s.dev.n_seg[r] : s.n_seg;
  const int n_splice = s.use_dev ? s.dev.n_splice[r] : s.n_splice;
  if (S.status == MPOA_OK) {
    if (qlen > a.caps.qmax || (keep_path && s.path0 + qlen > a.caps.path)) {
      if (tid == 0) S.status = MPOA_E_OVERFLOW;
    } else if (n_seg == 0) {
      for (int j = tid; j < m; j += NT) {
        A.res_from[j] = j == 0 ? 0 : -1;
        A.res_base[j] = j == 0 ? (uint8_t)CODE_N : (uint8_t)code_of(a.read_seq[q0 + j - 1]);
        A.res_aligned[j] = -1;
      }
    } else {
      if (tid == 0) {
        for (int k = 0; k < n_seg; k++) {
          const SegResult& g = s.seg_res[k];
          if (g.status != MPOA_OK && S.status == MPOA_OK) S.status = g.status;
          S.cells += g.cells;
          S.edge_cells += g.edge_cells;
          S.logical += g.logical;
          for (int t = 0; t < PH_N; t++) S.ticks[t] += g.ticks[t];
        }
      }
      // Anchor bases between sub-windows: existing consensus nodes (align.cpp:1794-1797).
      for (int k = tid; k < n_splice; k += NT) {
        const int pos = s.splice[2 * k], id = s.splice[2 * k + 1];
        A.res_from[pos] = id;
        A.res_base[pos] = A.base[id];
        A.res_aligned[pos] = -1;
      }
    }
  }
  __syncthreads();
  if (S.status == MPOA_OK) {
    commit_read<NT>(a, A, S, r, m, s.path0, r + 1 == s.n_reads, true,
                    reinterpret_cast<uint8_t*>(lds4), (int)sizeof(lds4));
  }
  if (s.use_dev && tid == 0) {
    s.dev.read_graph[2 * r] = S.n_nodes;
    s.dev.read_graph[2 * r + 1] = S.n_in_e;
  }
  if (s.use_dev && S.status == MPOA_OK && r + 1 < s.n_reads) {
    seed_next<NT>(a, wd, s.dev, r + 1, r + 1, A, S, reinterpret_cast<uint8_t*>(lds4),
                  (int)sizeof(lds4));
  } else if (S.status == MPOA_OK && r + 1 < s.n_reads) {
    if (S.cons_len > s.cons_cap) {
      if (tid == 0) S.status = MPOA_E_OVERFLOW;
    } else {
      for (int p = tid; p < S.cons_len; p += NT) {
        s.cons_out[3 * p] = A.cons_id[p];
        s.cons_out[3 * p + 1] = A.cons_rank[p];
        s.cons_out[3 * p + 2] = A.cons_ch[p];
      }
    }
    __syncthreads();
  }
  if (r + 1 == s.n_reads || S.status != MPOA_OK)
    window_outputs<NT>(a, A, S, 0, reinterpret_cast<uint8_t*>(lds4), (int)sizeof(lds4));
  if (tid == 0) state_store(S, *s.state);
}

// A seeded (-S) window of a batch in one workgroup (minipoa's loop, main.cpp:110-121, with
// alignment()'s -S branch, align.cpp:1752-1797): the reads' minimizers, then read by read its
// sub-windows one after another (each a poa() call on a rank window of the graph, writing its
// stretch of the read's res), the anchor bases between them, the commit, and the seeding of
// the next read against the new consensus. lds: the workgroup's LDS, shared by the DP ring,
// the sort and the seeding, which never run at once.
template <int NT, int CPL, class RingT>
__device__ void run_window_seeded(const KernelArgs& a, Arena& A, WinShared& S, RingT& R,
                                  uint8_t* lds, int lds_bytes, const SeedDev& d, int w) {
  const WinDesc& wd = a.win[w];
  const int tid = threadIdx.x;
  const bool keep_path = a.scores.result != 0;
  window_begin<NT>(a, A, S, w);
  int* sh = reinterpret_cast<int*>(lds + lds_bytes - 64);
  bool ok = true;
  for (int r = 0; r < wd.n_reads && ok; r++)
    ok = seed_read_minimizers<NT>(a, wd, d, r, lds, lds_bytes - 64, sh);
  if (tid == 0) {
    if (!ok) S.status = MPOA_E_OVERFLOW;
    d.n_seg[0] = 0;
    d.n_splice[0] = 0;
  }
  __syncthreads();
  int path0 = 0;
  for (int r = 0; r < wd.n_reads && S.status == MPOA_OK; r++) {
    const int64_t q0 = a.read_off[wd.read0 + r], q1 = a.read_off[wd.read0 + r + 1];
    const int qlen = (int)(q1 - q0), m = qlen + 1;
    if (qlen > a.caps.qmax || (keep_path && path0 + qlen > a.caps.path)) {
      if (tid == 0) S.status = MPOA_E_OVERFLOW;
      __syncthreads();
      break;
    }
    if (tid == 0) S.read = r;
    const int n_seg = d.n_seg[0], n_splice = d.n_splice[0];
    if (n_seg == 0) {  // a graph without bases takes the read as a new path
      load_query<NT>(A, S, R, a.read_seq + q0, qlen);
      trivial_res<NT>(A, m);
    } else {
      const bool adaptive = a.scores.f > 0;  // -S always bands adaptively (align.cpp:1780)
      for (int g = 0; g < n_seg && S.status == MPOA_OK; g++) {
        const SegDesc sd = d.segs[g];
        Arena As = A;
        As.res_from += sd.res0;
        As.res_aligned += sd.res0;
        As.res_base += sd.res0;
        if (tid == 0) {
          S.beg_i = A.rank[sd.beg_id];
          S.end_i = A.rank[sd.end_id];
          S.n_rows = S.end_i - S.beg_i + 1;
        }
        __syncthreads();
        if (sd.qlen == -1) {  // anchors overlapping by one base: no DP, no res entry
          if (tid == 0) {
            if (2LL * S.n_rows > 3 * a.caps.dp) S.status = MPOA_E_OVERFLOW;
            else null_call_score(A, a.scores, S.beg_i, S.n_rows, A.H, A.H + S.n_rows);
          }
          __syncthreads();
          continue;
        }
        load_query<NT>(As, S, R, a.read_seq + q0 + sd.q0, sd.qlen);
        if (S.q_lds) poa_call<NT, CPL, true, false, RingT>(a, As, S, R, adaptive);
        else poa_call<NT, CPL, false, false, RingT>(a, As, S, R, adaptive);
        __syncthreads();
      }
      if (S.status == MPOA_OK) {
        // anchor bases between the sub-windows: existing consensus nodes (align.cpp:1794-1797)
        for (int k = tid; k < n_splice; k += NT) {
          const int pos = d.splice[2 * k], id = d.splice[2 * k + 1];
          A.res_from[pos] = id;
          A.res_base[pos] = A.base[id];
          A.res_aligned[pos] = -1;
        }
      }
      __syncthreads();
    }
    if (S.status != MPOA_OK) break;
    commit_read<NT>(a, A, S, r, m, path0, r + 1 == wd.n_reads, true, lds, lds_bytes);
    path0 += qlen;
    __syncthreads();
    if (S.status != MPOA_OK) break;
    if (r + 1 < wd.n_reads) seed_next<NT>(a, wd, d, r + 1, 0, A, S, lds, lds_bytes);
  }
  window_outputs<NT>(a, A, S, w, lds, lds_bytes);
}

// The anchors of read s.read on a given consensus (mpoa_seed_anchors_device): the seeding of
// seed_next on a graph-less consensus, stopping at the picked anchors.
template <int NT>
__global__ __launch_bounds__(NT) void k_seed_anchors(SeedArgs s) {
  __shared__ WinShared S;
  __shared__ int4 lds4[(156 * 1024) / 16];
  Arena A{};
  A.cons_ch = const_cast<uint8_t*>(s.dev.cons_codes);
  if (threadIdx.x == 0) {
    S.status = MPOA_OK;
    S.n_nodes = 3;  // any graph with bases
    S.cons_len = s.cons_cap;
  }
  __syncthreads();
  seed_next<NT>(s.graph, s.graph.win[0], s.dev, s.read, 0, A, S,
                reinterpret_cast<uint8_t*>(lds4), (int)sizeof(lds4));
  if (threadIdx.x == 0 && S.status != MPOA_OK) *s.dev.n_pick_out = -1;
}

// LDS of the seeded-batch tiles: the 256-thread tile takes what one CU has (every window
// fits its seeding); the 64-thread tile a fifth, so that several windows share a CU (a window
// whose seeding outgrows it re-runs on the 256-thread tile).
template <int NT>
struct SeededLds {
  static constexpr int bytes = NT == 64 ? 30 * 1024 : 156 * 1024;
};

template <int NT, int CPL>
__global__ __launch_bounds__(NT) void k_seeded(KernelArgs a) {
  constexpr int SEEDED_LDS = SeededLds<NT>::bytes;
  __shared__ WinShared S;
  __shared__ __align__(16) uint8_t lds[SEEDED_LDS];
  using RingT = Ring<RingShape<NT>::RING, RingShape<NT>::WR, RingShape<NT>::QCAP>;
  static_assert(sizeof(RingT) <= SEEDED_LDS, "the DP ring shares the seeding LDS");
  RingT& R = *reinterpret_cast<RingT*>(lds);
  Arena A = carve(a, blockIdx.x);
  const SeedDev d = a.seed[blockIdx.x];
  for (;;) {
    if (threadIdx.x == 0) S.job = atomicAdd(a.queue, 1);
    __syncthreads();
    const int job = S.job;
    __syncthreads();
    if (job >= a.n_jobs) break;
    run_window_seeded<NT, CPL>(a, A, S, R, lds, SEEDED_LDS, d, a.order[job]);
  }
}

// The call-level operations on a graph kept in device memory (GraphOpArgs): graph::init, one
// poa() call on a rank window, graph::add_path of a given res, graph::topsort op 0 or 1 (with
// the consensus the next read is seeded against), and the consensus, MSA or GFA output. The
// same device functions as the batched pipeline, one step per launch.
template <int NT, int CPL>
__global__ __launch_bounds__(NT) void k_graph_op(GraphOpArgs g) {
  __shared__ WinShared S;
  __shared__ Ring<RingShape<NT>::RING, RingShape<NT>::WR, RingShape<NT>::QCAP> R;
  using RingT = Ring<RingShape<NT>::RING, RingShape<NT>::WR, RingShape<NT>::QCAP>;
  const KernelArgs& a = g.a;
  const int tid = threadIdx.x;
  Arena A = carve(a, 0);
  uint8_t* lds = reinterpret_cast<uint8_t*>(&R);
  if (g.op == GOP_INIT) {
    window_begin<NT>(a, A, S, 0);  // graph::init: source 0 and sink 1 (graph.cpp:129-135)
    if (tid == 0) {
      state_store(S, *g.state);
      g.out->status = S.status;
    }
    return;
  }
  if (tid == 0) state_load(S, *g.state);
  __syncthreads();
  if (g.op == GOP_POA) {
    // poa() (align.cpp:1353-1737) on ranks [rank(beg_id), rank(end_id)]; the graph is read
    // only (the DP plan, band state and matrices are scratch), so the state is not stored
    if (tid == 0) {
      S.status = MPOA_OK;
      S.read = 0;
      S.beg_i = A.rank[g.beg_id];
      S.end_i = A.rank[g.end_id];
      S.n_rows = S.end_i - S.beg_i + 1;
      S.cells = 0;
      S.edge_cells = 0;
      for (int k = 0; k < PH_N; k++) S.ticks[k] = 0;
    }
    __syncthreads();
    if (S.n_rows <= 0 || g.qlen > a.caps.qmax || S.n_rows > a.caps.nodes) {
      if (tid == 0) {
        g.out->status = S.n_rows <= 0 ? MPOA_E_EMPTY_WINDOW : MPOA_E_OVERFLOW;
        g.out->rows = S.n_rows;
        g.out->qlen = g.qlen;
      }
      return;
    }
    if (g.qlen == -1) {  // a -S sub-window of anchors overlapping by one base: no res entry
      if (tid == 0) {
        GraphOpOut& o = *g.out;
        o.status = 2LL * S.n_rows <= 3 * a.caps.dp ? MPOA_OK : MPOA_E_OVERFLOW;
        o.score = o.status == MPOA_OK
                      ? null_call_score(A, a.scores, S.beg_i, S.n_rows, A.H, A.H + S.n_rows)
                      : 0;
        o.rows = S.n_rows; o.qlen = -1; o.width = 0;
        o.cells = 0; o.edge_cells = 0; o.logical = 0;
      }
      return;
    }
    load_query<NT>(A, S, R, g.q, g.qlen);
    if (S.q_lds) poa_call<NT, CPL, true, false, RingT>(a, A, S, R, g.adaptive != 0);
    else poa_call<NT, CPL, false, false, RingT>(a, A, S, R, g.adaptive != 0);
    if (tid == 0) {
      GraphOpOut& o = *g.out;
      o.status = S.status;
      o.score = S.score;
      o.rows = S.n_rows;
      o.qlen = g.qlen;
      o.width = S.call_width;
      o.cells = S.call_cells;
      o.edge_cells = S.call_ecells;
      o.logical = S.call_logical;
    }
    return;
  }
  if (g.op == GOP_ADD_PATH) {
    // graph::add_path(m, seq_id, res, sink 1) (graph.cpp:170-235)
    const bool keep_path = a.scores.result != 0;
    const int m = g.n_res;
    if (S.status == MPOA_OK) {
      if (keep_path) {
        for (int j = tid; j < m - 1; j += NT) A.path[g.path0 + j] = 1;
        __syncthreads();
      }
      if ((a.forms & MPOA_FORM_ADD_LANE) ||
          !add_path_parallel<NT>(A, S, a.caps, keep_path, g.seq_id, m, g.path0, A.t2, A.t1)) {
        if (tid == 0) add_path(A, S, a.caps, keep_path, g.seq_id, m, g.path0);
      }
      __syncthreads();
    }
  } else if (g.op == GOP_ADD_NODES) {
    // graph::add_node per segment base (graph.cpp:7-12, read_gfa's S lines,
    // handle_input.cpp:103-121): ids in order, each its own column head, no edges yet
    const int n0 = S.n_nodes, cnt = g.qlen;
    if (S.status == MPOA_OK && n0 + cnt <= a.caps.nodes) {
      for (int k = tid; k < cnt; k += NT) {
        const int v = n0 + k, b = g.q[k];
        A.base[v] = (uint8_t)b;
        A.par[v] = v;
        for (int c = 0; c < NB; c++) A.aligned[v * NB + c] = -1;
        A.aligned[v * NB + b] = v;
        A.ind[v] = 0;
        A.in_head[v] = A.in_tail[v] = A.out_head[v] = A.out_tail[v] = -1;
        A.in_cnt[v] = 0;
        A.out_cnt[v] = 0;
        A.last_read[v] = -1;
      }
      __syncthreads();
      if (tid == 0) S.n_nodes = n0 + cnt;
    } else if (tid == 0 && S.status == MPOA_OK) {
      S.status = MPOA_E_OVERFLOW;
    }
    __syncthreads();
  } else if (g.op == GOP_TOPSORT) {
    // graph::topsort(para, op) (graph.cpp:21-128); with seeding, build_consensus
    if (S.status == MPOA_OK) {
      if (g.sort_op == 0) {
        topsort0<NT>(A, S, a.caps, a.scores.f > 0, lds, (int)sizeof(R), a.forms);
        if (tid == 0) S.msa_cols = S.n_nodes;  // an MSA now has one column per rank
        if (g.seeding) {
          heaviest_ranked<NT>(A, S, lds, (int)sizeof(R), a.forms);
          if (S.cons_len > g.cons_cap) {
            if (tid == 0) S.status = MPOA_E_OVERFLOW;
          } else {
            for (int p = tid; p < S.cons_len; p += NT) {
              g.cons_out[3 * p] = A.cons_id[p];
              g.cons_out[3 * p + 1] = A.cons_rank[p];
              g.cons_out[3 * p + 2] = A.cons_ch[p];
            }
          }
        }
      } else if (tid == 0) {
        S.msa_cols = topsort1(A, S);
      }
      __syncthreads();
    }
  } else if (g.op == GOP_OUTPUT) {
    // graph::output_consensus / output_rc_msa / output_gfa (graph.cpp:237-433)
    window_outputs<NT>(a, A, S, 0, lds, (int)sizeof(R));
  }
  if (tid == 0) {
    state_store(S, *g.state);
    g.out->status = S.status;
  }
}

// The one-wavefront tile asks for three waves per SIMD (168 VGPRs at most): its LDS admits
// eleven workgroups per CU, and the compiler's own occupancy target (two) would let the
// register count grow past what three need. The larger tiles are held to two by their LDS.
template <int NT, int CPL>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 64 ? 3 : 1, 8)))
k_windows(KernelArgs a) {
  __shared__ WinShared S;
  __shared__ Ring<RingShape<NT>::RING, RingShape<NT>::WR, RingShape<NT>::QCAP> R;
  Arena A = carve(a, blockIdx.x);
  for (;;) {
    if (threadIdx.x == 0) S.job = atomicAdd(a.queue, 1);
    __syncthreads();
    const int job = S.job;
    __syncthreads();
    if (job >= a.n_jobs) break;
    run_window<NT, CPL>(a, A, S, R, a.order[job]);
  }
}

// Resets a work-queue counter (a kernel rather than a memset: it goes into graphs).
__global__ void k_fill(int* p, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Holds one wavefront for `ticks` of the wall clock (at most 65,536 sleeps of 64 x 64 cycles,
// ~0.1 s): in a graph it delays the launch that follows it on its branch, the way a host sleep
// does between eager launches.
__global__ void k_hold(int64_t ticks) {
  const int64_t t0 = now();
  for (int i = 0; i < (1 << 16) && now() - t0 < ticks; i++) __builtin_amdgcn_s_sleep(64);
}

// ---------------------------------------------------------------------------------------
// Allocation-free form (mpoa_align_reads_noalloc): one window whose work-queue counter,
// processing order, descriptor, result record and slot arenas all live in `scratch`, a fixed
// device array (each device holds its own copy until the process ends). Host code never uses
// the array's name: it takes the address hipGetSymbolAddress returns for the current device
// (mpoa_scratch_address) and passes pointers into it to the kernels below and to k_windows.
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) uint8_t scratch[MPOA_NOALLOC_SCRATCH_BYTES];

// Writes the window descriptor and resets the processing order and the queue counter: kernel
// arguments, not host copies, so that the whole sequence allocates nothing and may be captured.
__global__ void k_noalloc_setup(WinDesc* wd, WinDesc v, int* order, int* queue) {
  if (threadIdx.x == 0) {
    *wd = v;
    order[0] = 0;
    queue[0] = 0;
  }
}

// Copies the window's status and output size in bytes (out_len units of `unit` bytes: 1 for
// the consensus, n_reads for the MSA's columns) to the caller's device words.
__global__ void k_noalloc_finish(const WinResult* r, int64_t* d_res, int unit) {
  if (threadIdx.x == 0) {
    d_res[0] = r->status;
    d_res[1] = r->status == MPOA_OK ? (int64_t)r->out_len * unit : 0;
  }
}

}  // namespace mpoa

int mpoa_scratch_address(int device, void** out) {
  // one lookup per device, with that device current (the caller sets it); later calls reuse it
  static std::mutex mu;
  static void* addr[64] = {};
  if (device < 0 || device >= 64 || !out) return -1;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(mpoa::scratch)) != hipSuccess || !p) return -1;
    addr[device] = p;
  }
  *out = addr[device];
  return 0;
}

int mpoa_launch_noalloc(const mpoa::KernelArgs& a, const mpoa::WinDesc& wd, int nt, int cpl,
                        int64_t* d_res, hipStream_t stream) {
  const int unit = a.scores.result == 1 ? wd.n_reads : 1;
  using namespace mpoa;
  hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(64), 0, stream, const_cast<WinDesc*>(a.win),
                     wd, const_cast<int*>(a.order), a.queue);
  if (hipGetLastError() != hipSuccess) return -1;
  const int rc = mpoa_launch_windows(a, 1, nt, cpl, stream);
  if (rc != 0) return rc;
  hipLaunchKernelGGL(k_noalloc_finish, dim3(1), dim3(64), 0, stream, a.result, d_res, unit);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

#define MPOA_TILES(X) \
  X(64, 1) X(64, 2) X(64, 4) X(64, 8) X(128, 2) X(128, 4) X(256, 1) X(256, 2) X(256, 4) X(512, 2) \
  X(512, 4)

bool mpoa_tile_supported(int nt, int cpl) {
#define MPOA_T(NT_, CPL_) if (nt == NT_ && cpl == CPL_) return true;
  MPOA_TILES(MPOA_T)
#undef MPOA_T
  return false;
}

int mpoa_tile_occupancy(int nt, int cpl) {
  using namespace mpoa;
  int blocks = 0;
#define MPOA_O(NT_, CPL_)                                                                    \
  if (nt == NT_ && cpl == CPL_)                                                             \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_windows<NT_, CPL_>, NT_,  \
                                                        0) == hipSuccess ? blocks : 0;
  MPOA_TILES(MPOA_O)
#undef MPOA_O
  return 0;
}

// Host-side launch (pipeline.h).
int mpoa_launch_windows(const mpoa::KernelArgs& a, int n_slots, int nt, int cpl,
                        hipStream_t stream) {
  using namespace mpoa;
#define MPOA_L(NT_, CPL_)                                                                     \
  if (nt == NT_ && cpl == CPL_) {                                                            \
    hipLaunchKernelGGL((k_windows<NT_, CPL_>), dim3(n_slots), dim3(NT_), 0, stream, a);     \
    return hipGetLastError() == hipSuccess ? 0 : -1;                                         \
  }
  MPOA_TILES(MPOA_L)
#undef MPOA_L
  return -2;
}

int mpoa_graph_add_windows(hipGraph_t g, const hipGraphNode_t* deps, size_t n_deps,
                           const mpoa::KernelArgs& a, int n_slots, int nt, int cpl,
                           hipGraphNode_t* last) {
  using namespace mpoa;
  const void* fn = nullptr;
#define MPOA_G(NT_, CPL_) \
  if (nt == NT_ && cpl == CPL_) fn = reinterpret_cast<const void*>(&k_windows<NT_, CPL_>);
  MPOA_TILES(MPOA_G)
#undef MPOA_G
  if (!fn) return -2;
  int* queue = a.queue;
  int zero = 0, one = 1;
  void* fill_args[] = {&queue, &zero, &one};
  hipKernelNodeParams fp{};
  fp.func = reinterpret_cast<void*>(&k_fill);
  fp.gridDim = dim3(1);
  fp.blockDim = dim3(64);
  fp.kernelParams = fill_args;
  hipGraphNode_t fill;
  if (hipGraphAddKernelNode(&fill, g, deps, n_deps, &fp) != hipSuccess) return -1;
  KernelArgs args = a;
  void* win_args[] = {&args};
  hipKernelNodeParams wp{};
  wp.func = const_cast<void*>(fn);
  wp.gridDim = dim3(n_slots);
  wp.blockDim = dim3(nt);
  wp.kernelParams = win_args;
  return hipGraphAddKernelNode(last, g, &fill, 1, &wp) == hipSuccess ? 0 : -1;
}

int mpoa_graph_add_hold(hipGraph_t g, const hipGraphNode_t* deps, size_t n_deps, int64_t ticks,
                        hipGraphNode_t* node) {
  void* args[] = {&ticks};
  hipKernelNodeParams p{};
  p.func = reinterpret_cast<void*>(&mpoa::k_hold);
  p.gridDim = dim3(1);
  p.blockDim = dim3(64);
  p.kernelParams = args;
  return hipGraphAddKernelNode(node, g, deps, n_deps, &p) == hipSuccess ? 0 : -1;
}

int mpoa_launch_segments(const mpoa::SeedArgs& s, int nt, int cpl, hipStream_t stream) {
  using namespace mpoa;
  // device seeding: as many workgroups as a read may have sub-windows; the extra ones return
  const int grid = s.use_dev ? SEED_SEGS : s.n_seg;
  if (grid <= 0) return 0;
#define MPOA_S(NT_, CPL_)                                                                     \
  if (nt == NT_ && cpl == CPL_) {                                                            \
    hipLaunchKernelGGL((k_segments<NT_, CPL_>), dim3(grid), dim3(NT_), 0, stream, s);        \
    return hipGetLastError() == hipSuccess ? 0 : -1;                                         \
  }
  MPOA_S(64, 4) MPOA_S(128, 4) MPOA_S(256, 2)
#undef MPOA_S
  return -2;
}

int mpoa_launch_seed_reads(const mpoa::SeedArgs& s, hipStream_t stream) {
  using namespace mpoa;
  if (s.n_reads <= 0) return 0;
  hipLaunchKernelGGL((k_seed_reads<256>), dim3(s.n_reads), dim3(256), 0, stream, s);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mpoa_launch_seeded(const mpoa::KernelArgs& a, int n_slots, int nt, hipStream_t stream) {
  using namespace mpoa;
  if (n_slots <= 0) return 0;
  if (nt == 64) hipLaunchKernelGGL((k_seeded<64, 4>), dim3(n_slots), dim3(64), 0, stream, a);
  else if (nt == 256) hipLaunchKernelGGL((k_seeded<256, 2>), dim3(n_slots), dim3(256), 0, stream, a);
  else return -2;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mpoa_seeded_occupancy(int nt) {
  using namespace mpoa;
  int blocks = 0;
  hipError_t e = nt == 64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_seeded<64, 4>, 64, 0)
               : nt == 256 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_seeded<256, 2>, 256, 0)
                           : hipErrorInvalidValue;
  return e == hipSuccess ? blocks : 0;
}

int mpoa_launch_seed_anchors(const mpoa::SeedArgs& s, hipStream_t stream) {
  using namespace mpoa;
  hipLaunchKernelGGL((k_seed_anchors<256>), dim3(1), dim3(256), 0, stream, s);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mpoa_launch_commit(const mpoa::SeedArgs& s, hipStream_t stream) {
  using namespace mpoa;
  hipLaunchKernelGGL((k_commit<256>), dim3(1), dim3(256), 0, stream, s);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mpoa_launch_graph_op(const mpoa::GraphOpArgs& g, hipStream_t stream) {
  using namespace mpoa;
  hipLaunchKernelGGL((k_graph_op<512, 2>), dim3(1), dim3(512), 0, stream, g);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
