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

#include "pipeline.h"

namespace mpoa {

constexpr int Q = 8;                      // band quantum (src/align.cpp:13,19)
constexpr int NEG_INF = (int)0xc0c0c0c0;  // src/align.cpp:22
constexpr int NB = 5;
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

// Per-slot arena: carved from the int pool / byte pool / DP pool at fixed offsets.
struct Arena {
  uint8_t* base;
  int *par, *aligned, *ind, *rank, *in_head, *in_tail, *out_head, *out_tail, *in_cnt, *last_read;
  int *ie_src, *ie_w, *ie_next, *oe_dst, *oe_w, *oe_next;
  int *rank_id, *t0, *t1, *t2, *hlen, *tlen;
  int *pred_off, *pred_rank, *pred_w, *Bs, *Be, *row_off;
  uint8_t* pred_base;
  int *Mp, *Sl, *Pl, *Pr, *Ol, *Or, *Ow;
  int *res_from, *res_aligned;
  uint8_t *res_base, *qseq;
  int* path;
  int *H, *D, *I;
};

__device__ Arena carve(const KernelArgs& a, int slot) {
  const SlotCaps& c = a.caps;
  Arena A;
  int* p = a.int_pool + (int64_t)slot * c.ints_per_slot;
  auto take = [&](int64_t n) { int* r = p; p += n; return r; };
  const int N = c.nodes, E = c.edges;
  A.par = take(N); A.aligned = take((int64_t)NB * N); A.ind = take(N); A.rank = take(N);
  A.in_head = take(N); A.in_tail = take(N); A.out_head = take(N); A.out_tail = take(N);
  A.in_cnt = take(N); A.last_read = take(N);
  A.ie_src = take(E); A.ie_w = take(E); A.ie_next = take(E);
  A.oe_dst = take(E); A.oe_w = take(E); A.oe_next = take(E);
  A.rank_id = take(N); A.t0 = take(N); A.t1 = take(N); A.t2 = take(N);
  A.hlen = take(N); A.tlen = take(N);
  A.pred_off = take(N + 1); A.pred_rank = take(E); A.pred_w = take(E);
  A.Bs = take(N); A.Be = take(N); A.row_off = take(N + 1);
  A.Mp = take(N); A.Sl = take(N); A.Pl = take(N); A.Pr = take(N);
  A.Ol = take(N); A.Or = take(N); A.Ow = take(N);
  A.res_from = take(c.qmax + 2); A.res_aligned = take(c.qmax + 2);
  A.path = take(c.path);
  uint8_t* b = a.byte_pool + (int64_t)slot * c.bytes_per_slot;
  A.base = b; b += N;
  A.pred_base = b; b += E;
  A.res_base = b; b += c.qmax + 2;
  A.qseq = b;
  int* dp = a.dp_pool + (int64_t)slot * 3 * c.dp;
  // The int words taken above never exceed slot_ints(N, E, qmax, path) (pipeline.h).
  A.H = dp; A.D = dp + c.dp; A.I = dp + 2 * c.dp;
  return A;
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
  int64_t total;  // DP storage of the current call
  int job;
  int scan[16];
  int row_bs, row_be;
};

// ---------------------------------------------------------------------------------------
// Graph edits (one lane). Follow graph::add_adj / node_t::add_in_adj / add_out_adj
// (src/graph.cpp:13-20, include/node.h:23-61) and graph::add_path (graph.cpp:170-235).
// ---------------------------------------------------------------------------------------
__device__ int new_node(Arena& A, WinShared& S, const SlotCaps& c, int base) {
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
  A.last_read[id] = -1;
  return id;
}

__device__ void add_adj(Arena& A, WinShared& S, const SlotCaps& c, bool keep_path, int seq_id,
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

__device__ void add_path(Arena& A, WinShared& S, const SlotCaps& c, bool keep_path, int seq_id,
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

// graph::topsort op=0 (src/graph.cpp:59-123): LIFO Kahn order seeded with the zero
// in-degree nodes in ascending id order; then, in rank order, the DP plan (predecessor CSR in
// node.in order, with rank, weight and base) and hlen (1 + hlen of the max-in-weight
// predecessor, first wins); then tlen in reverse rank order (1 + tlen of the max-out-weight
// successor; the sink is 1).
__device__ void topsort0(Arena& A, WinShared& S, const SlotCaps& c, bool with_len) {
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
    int u = stk[--sp];
    A.rank[u] = r;
    A.rank_id[r] = u;
    r++;
    for (int e = A.out_head[u]; e != -1; e = A.oe_next[e]) {
      int v = A.oe_dst[e];
      if (--deg[v] == 0) stk[sp++] = v;
    }
  }
  // plan + hlen
  int k = 0;
  A.hlen[0] = 0;
  for (int rr = 0; rr < r; rr++) {
    int u = A.rank_id[rr];
    A.pred_off[rr] = k;
    int wmax = -1, max_pre = -1;
    for (int e = A.in_head[u]; e != -1; e = A.ie_next[e]) {
      int src = A.ie_src[e], w = A.ie_w[e], pr = A.rank[src];
      A.pred_rank[k] = pr;
      A.pred_w[k] = w;
      A.pred_base[k] = A.base[src];
      k++;
      if (w > wmax) { wmax = w; max_pre = pr; }
    }
    if (with_len && max_pre != -1) A.hlen[rr] = A.hlen[max_pre] + 1;
  }
  A.pred_off[r] = k;
  if (with_len) {
    A.tlen[A.rank[1]] = 1;
    for (int rr = r - 1; rr >= 0; rr--) {
      int u = A.rank_id[rr];
      int wmax = -1, max_suc = -1;
      for (int e = A.out_head[u]; e != -1; e = A.oe_next[e])
        if (A.oe_w[e] > wmax) { wmax = A.oe_w[e]; max_suc = A.rank[A.oe_dst[e]]; }
      if (max_suc != -1) A.tlen[rr] = A.tlen[max_suc] + 1;
    }
  }
}

// graph::topsort op=1 (src/graph.cpp:24-57): LIFO order over aligned-node groups; members
// inherit their column head's rank. Returns the number of groups.
__device__ int topsort1(Arena& A, WinShared& S) {
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

// graph::build_consensus (src/graph.cpp:294-350): FIFO BFS from the source; each node keeps
// the predecessor with the largest in-weight, ties to the larger accumulated score. Writes the
// consensus bases to out (reversed trace) and returns its length, or -1 if over cap.
__device__ int consensus(Arena& A, WinShared& S, uint8_t* out, int cap) {
  const int N = S.n_nodes;
  int* deg = A.t0;
  int* queue = A.t1;
  int* score = A.t2;
  int* pre = A.hlen;
  for (int v = 0; v < N; v++) {
    deg[v] = A.in_cnt[v];
    score[v] = 0;
    pre[v] = 0;
  }
  int qh = 0, qt = 0;
  queue[qt++] = 0;
  while (qh < qt) {
    int u = queue[qh++];
    if (u != 0) {
      int wmax = -1, max_pre = -1;
      for (int e = A.in_head[u]; e != -1; e = A.ie_next[e]) {
        int v = A.ie_src[e], w = A.ie_w[e];
        if (w > wmax || (w == wmax && score[v] > score[max_pre])) {
          wmax = w;
          max_pre = v;
        }
      }
      pre[u] = max_pre;
      score[u] = score[max_pre] + wmax;
    }
    if (u == 1) break;
    for (int e = A.out_head[u]; e != -1; e = A.oe_next[e]) {
      int v = A.oe_dst[e];
      if (--deg[v] == 0) queue[qt++] = v;
    }
  }
  int len = 0;
  for (int id = pre[1]; id != 0; id = pre[id]) len++;
  if (len > cap) return -1;
  int at = len;
  for (int id = pre[1]; id != 0; id = pre[id]) out[--at] = char_of(A.base[id]);
  return len;
}

// ---------------------------------------------------------------------------------------
// Banded DP (src/align.cpp:1357-1514).
// ---------------------------------------------------------------------------------------

// Exclusive prefix max of v over the workgroup's threads (NT threads, NT/64 waves), and the
// maximum over all threads in *total.
template <int NT>
__device__ __forceinline__ int block_excl_max(int v, WinShared& S, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const int u = __shfl_up(incl, s, 64);
    if (lane >= s) incl = max(incl, u);
  }
  int excl = __shfl_up(incl, 1, 64);
  if (lane == 0) excl = SCAN_ID;
  if (NT == 64) {
    *total = __shfl(incl, 63, 64);
    return excl;
  }
  if (lane == 63) S.scan[wave] = incl;
  __syncthreads();
  int before = SCAN_ID, all = SCAN_ID;
#pragma unroll
  for (int w2 = 0; w2 < NT / 64; w2++) {
    const int t = S.scan[w2];
    if (w2 < wave) before = max(before, t);
    all = max(all, t);
  }
  __syncthreads();
  *total = all;
  return max(excl, before);
}

// One DP row i (local) for rows in rank window [beg_i, end_i]: M/D sources from every
// in-window predecessor under the per-block band conditions, then the I/H scan computed as
// an exact prefix max: I[j] = j*e + max(NEG_INF, max_{k<j} (G[k] + o - (k+1)*e)),
// G = max(M-source, D-source), H = max(G, I). This equals the serial recurrence
// I[j] = max(I[j-1]+e, H[j-1]+o) because o <= e (gap_open <= 0) and int32 never wraps here.
template <int NT, int CPL>
__device__ void dp_row(const Arena& A, WinShared& S, const DpScores& sc, int i) {
  const int tid = threadIdx.x;
  const int Bsi = A.Bs[i], Bei = A.Be[i];
  const int c0 = Bsi * Q, W = (Bei - Bsi - 1) * Q, c1 = c0 + W;
  const int rowi = A.row_off[i];
  int* Hi = A.H + rowi - c0;  // indexed by absolute column
  int* Di = A.D + rowi - c0;
  int* Ii = A.I + rowi - c0;
  const int k0 = A.pred_off[S.beg_i + i], k1 = A.pred_off[S.beg_i + i + 1];
  const int o = sc.o1, e = sc.e1;
  int carry = SCAN_ID;
  for (int pb = c0; pb < c1; pb += NT * CPL) {
    const int cb = pb + tid * CPL;
    int h[CPL], d[CPL];
#pragma unroll
    for (int l = 0; l < CPL; l++) h[l] = d[l] = NEG_INF;
    for (int k = k0; k < k1; k++) {
      const int p = A.pred_rank[k] - S.beg_i;
      if (p < 0) continue;
      const int pbase = p == 0 ? CODE_N : A.pred_base[k];
      const int Bsp = A.Bs[p], Bep = A.Be[p];
      const int pofs = Bsp * Q;
      const int* Hp = A.H + A.row_off[p] - pofs;
      const int* Dp = A.D + A.row_off[p] - pofs;
      const int mlo = max(c0, Bsp == 0 ? 0 : (Bsp + 1) * Q), mhi = min(c1, (Bep - 1) * Q);
      const int dlo = max(c0, Bsp * Q), dhi = min(c1, (Bep - 1) * Q);
#pragma unroll
      for (int l = 0; l < CPL; l++) {
        const int c = cb + l;
        if (c >= mlo && c < mhi) {
          const int diag = c == c0 ? NEG_INF : Hp[c - 1];
          const int qc = c == 0 ? CODE_N : A.qseq[c - 1];
          h[l] = max(h[l], diag + (pbase == qc ? sc.match : sc.mismatch));
        }
        if (c >= dlo && c < dhi) d[l] = max(d[l], max(Dp[c] + e, Hp[c] + o));
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
    int run = max(block_excl_max<NT>(lane_max, S, &total), carry);
#pragma unroll
    for (int l = 0; l < CPL; l++) {
      const int c = cb + l;
      if (c < c1) {
        const int j = c - c0;
        const int iv = j * e + max(NEG_INF, run);
        Hi[c] = max(g[l], iv);
        Di[c] = d[l];
        Ii[c] = iv;
      }
      run = max(run, t[l]);
    }
    carry = max(carry, total);
  }
  // NEG_INF sentinel block after the band (align.cpp:1505-1506,1515).
  if (tid < Q) Hi[c1 + tid] = Di[c1 + tid] = Ii[c1 + tid] = NEG_INF;
}

// Static band per row (src/align.cpp:1378-1394) and ragged row offsets; all lanes.
// Returns false on arena overflow.
template <int NT>
__device__ bool plan_static(const Arena& A, WinShared& S, const DpScores& sc, const SlotCaps& cap,
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
      } else {
        bs = 0;
        be = m / Q + 2;
      }
      A.Bs[r] = bs;
      A.Be[r] = be;
      len = (be - bs) * Q;
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
    if (r < n) A.row_off[r] = (int)(base + before + v - len);
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
__device__ int backtrack(const Arena& A, WinShared& S, const DpScores& sc) {
  const int n = S.n_rows, m = S.m, beg = S.beg_i;
  const int match = sc.match, mismatch = sc.mismatch, o1 = sc.o1, e1 = sc.e1;
  auto Hat = [&](int r, int j) { return A.H[A.row_off[r] + j]; };
  auto Dat = [&](int r, int j) { return A.D[A.row_off[r] + j]; };
  auto Iat = [&](int r, int j) { return A.I[A.row_off[r] + j]; };
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
        const int pbase = p == 0 ? CODE_N : A.pred_base[k];
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
        const int pbase = p == 0 ? CODE_N : A.pred_base[k];
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

// Adaptive band of row i before its DP (src/align.cpp:1433-1444); one lane.
__device__ void adaptive_band_row(const Arena& A, WinShared& S, int i, int w) {
  const int beg = S.beg_i, end = S.end_i, m = S.m;
  const int aci = beg + i;
  const int lo = A.hlen[aci] - A.hlen[beg], hi = m - A.tlen[aci] + A.tlen[end];
  const int ow = i == 0 ? 0 : A.Ow[i];
  const int Ms = max(0, min(min(A.Pl[i], lo + A.Ol[i]), hi) - w - ow);
  const int Me = min(m, max(max(A.Pr[i], lo + A.Or[i]), hi) + w + ow);
  A.Bs[i] = Ms / Q;
  A.Be[i] = Me / Q + 2;
}

// Adaptive-band update after row i (src/align.cpp:1516-1568); one lane, argmax given.
__device__ void adaptive_update(const Arena& A, WinShared& S, int i, int max_j, int w) {
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

// First argmax of H over the row's computed columns (strict >, from column 0).
template <int NT>
__device__ int row_argmax(const Arena& A, WinShared& S, int i) {
  const int W = (A.Be[i] - A.Bs[i] - 1) * Q;
  const int* Hi = A.H + A.row_off[i];
  int best = SCAN_ID, bj = 0x7fffffff;
  for (int j = threadIdx.x; j < W; j += NT) {
    const int v = Hi[j];
    if (v > best) { best = v; bj = j; }
  }
  // reduce (value desc, index asc)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    const int ov = __shfl_xor(best, s, 64), oj = __shfl_xor(bj, s, 64);
    if (ov > best || (ov == best && oj < bj)) { best = ov; bj = oj; }
  }
  if (NT > 64) {
    __syncthreads();
    if (lane == 0) { S.scan[2 * wave] = best; S.scan[2 * wave + 1] = bj; }
    __syncthreads();
    best = S.scan[0]; bj = S.scan[1];
    for (int w2 = 1; w2 < NT / 64; w2++) {
      const int ov = S.scan[2 * w2], oj = S.scan[2 * w2 + 1];
      if (ov > best || (ov == best && oj < bj)) { best = ov; bj = oj; }
    }
    __syncthreads();
  }
  return bj == 0x7fffffff ? 0 : bj;
}

// One poa() call on the device graph: rows [beg_i, end_i] of the current rank order.
template <int NT, int CPL>
__device__ void poa_call(const KernelArgs& a, Arena& A, WinShared& S, bool adaptive) {
  const DpScores& sc = a.scores;
  const int n = S.n_rows, m = S.m;
  int w = sc.b;
  if (sc.f) w += m / sc.f;
  if (!adaptive) {
    if (!plan_static<NT>(A, S, sc, a.caps, w)) {
      if (threadIdx.x == 0) S.status = MPOA_E_OVERFLOW;
      __syncthreads();
      return;
    }
  } else {
    // Full-width rows (align.cpp:1385-1390); adaptive state reset (align.cpp:1377).
    const int len = (m / Q + 2) * Q;
    if ((int64_t)len * n > a.caps.dp) {
      if (threadIdx.x == 0) S.status = MPOA_E_OVERFLOW;
      __syncthreads();
      return;
    }
    for (int r = threadIdx.x; r < n; r += NT) {
      A.row_off[r] = r * len;
      A.Mp[r] = 0; A.Sl[r] = 0; A.Pl[r] = m; A.Pr[r] = 0; A.Ol[r] = 0; A.Or[r] = 0; A.Ow[r] = 0;
    }
    if (threadIdx.x == 0) {
      A.row_off[n] = n * len;
      adaptive_band_row(A, S, 0, w);
    }
    __syncthreads();
  }
  // Row 0 (align.cpp:1415-1425): M, D = NEG_INF over blocks [Bs0, Be0) from the row start;
  // M[0][0] = 0. I of row 0 is never read; it is set to NEG_INF.
  {
    const int len0 = A.Be[0] * Q;
    int* H0 = A.H + A.row_off[0];
    int* D0 = A.D + A.row_off[0];
    int* I0 = A.I + A.row_off[0];
    for (int j = threadIdx.x; j < len0; j += NT) {
      H0[j] = j == 0 ? 0 : NEG_INF;
      D0[j] = NEG_INF;
      I0[j] = NEG_INF;
    }
  }
  __syncthreads();
  int64_t cells = 0, ecells = 0;
  for (int i = 1; i < n; i++) {
    if (adaptive) {
      if (threadIdx.x == 0) adaptive_band_row(A, S, i, w);
      __syncthreads();
    }
    if (A.Be[i] - A.Bs[i] - 1 < 1) {
      if (threadIdx.x == 0) S.status = MPOA_E_BAND;
      __syncthreads();
      return;
    }
    dp_row<NT, CPL>(A, S, sc, i);
    __syncthreads();
    if (threadIdx.x == 0) {
      const int Wd = (A.Be[i] - A.Bs[i] - 1) * Q;
      int np = 0;
      for (int k = A.pred_off[S.beg_i + i]; k < A.pred_off[S.beg_i + i + 1]; k++)
        np += A.pred_rank[k] - S.beg_i >= 0;
      cells += Wd;
      ecells += (int64_t)np * Wd;
    }
    if (adaptive) {
      const int mj = row_argmax<NT>(A, S, i);
      if (threadIdx.x == 0) adaptive_update(A, S, i, mj, w);
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    S.cells += cells;
    S.edge_cells += ecells;
    S.score = A.H[A.row_off[n - 1] + (m - A.Bs[n - 1] * Q)];
    const int rc = backtrack(A, S, sc);
    if (rc != MPOA_OK) S.status = rc;
  }
  __syncthreads();
}

// The whole per-window pipeline.
template <int NT, int CPL>
__device__ void run_window(const KernelArgs& a, Arena& A, WinShared& S, int w) {
  const WinDesc& wd = a.win[w];
  const int tid = threadIdx.x;
  const bool keep_path = a.scores.result != 0;
  const bool with_len = a.scores.f > 0;
  if (tid == 0) {
    S.status = MPOA_OK;
    S.n_nodes = 0; S.n_in_e = 0; S.n_out_e = 0;
    S.cells = 0; S.edge_cells = 0;
    new_node(A, S, a.caps, CODE_N);  // source 0 (graph.cpp:129-135)
    new_node(A, S, a.caps, CODE_N);  // sink 1
  }
  __syncthreads();
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
    // Query: sentinel N, codes, N padding to col_size (align.cpp:1362-1373).
    const int col_size = (m + 1 + Q - 1) / Q * Q;
    for (int j = tid; j < col_size + Q; j += NT)
      A.qseq[j] = j >= 1 && j <= qlen ? (uint8_t)code_of(a.read_seq[q0 + j - 1]) : (uint8_t)CODE_N;
    if (tid == 0) { S.read = r; S.qlen = qlen; S.m = m; }
    __syncthreads();
    if (S.n_nodes <= 2) {
      // First read: source then insertions (align.cpp:1747-1753).
      for (int j = tid; j < m; j += NT) {
        A.res_from[j] = j == 0 ? 0 : -1;
        A.res_base[j] = j == 0 ? (uint8_t)CODE_N : A.qseq[j];
        A.res_aligned[j] = -1;
      }
    } else {
      if (tid == 0) {
        S.beg_i = A.rank[0];
        S.end_i = A.rank[1];
        S.n_rows = S.end_i - S.beg_i + 1;
      }
      __syncthreads();
      poa_call<NT, CPL>(a, A, S, a.scores.ab_band != 0 && a.scores.f > 0);
      if (tid == 0 && a.call_stats && r - 1 < a.call_stats_cap) {
        CallStat& cs = a.call_stats[(int64_t)w * a.call_stats_stride + (r - 1)];
        cs.score = S.score;
        cs.rows = S.n_rows;
        cs.qlen = qlen;
      }
    }
    __syncthreads();
    if (S.status != MPOA_OK) break;
    if (tid == 0) {
      add_path(A, S, a.caps, keep_path, r, m, path0);
      if (S.status == MPOA_OK) {
        if (r + 1 == wd.n_reads) S.msa_cols = topsort1(A, S);
        else topsort0(A, S, a.caps, with_len);
      }
    }
    path0 += qlen;
    __syncthreads();
    if (S.status != MPOA_OK) break;
  }
  // Outputs.
  if (S.status == MPOA_OK) {
    if (a.scores.result == 0) {
      if (tid == 0) {
        const int len = consensus(A, S, a.out + wd.out0, wd.out_cap);
        if (len < 0) S.status = MPOA_E_OVERFLOW;
        else S.cons_len = len;
      }
      __syncthreads();
    } else if (a.scores.result == 1) {
      // Row-column MSA (graph.cpp:237-251): rows = reads, columns = group ranks 1..R-2.
      const int R = S.msa_cols;
      const int L = R >= 2 ? R - 2 : 0;
      const int64_t need = (int64_t)wd.n_reads * L;
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
              o[(int64_t)r * L + rk - 1] = char_of(A.base[v]);
            }
          }
          p0 += qlen;
        }
        if (tid == 0) S.cons_len = L;
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    WinResult& res = a.result[w];
    res.status = S.status;
    res.out_len = S.status == MPOA_OK ? S.cons_len : 0;
    res.cells = S.cells;
    res.edge_cells = S.edge_cells;
    res.nodes = S.n_nodes;
    res.edges = S.n_in_e;
  }
  __syncthreads();
}

template <int NT, int CPL>
__global__ __launch_bounds__(NT) void k_windows(KernelArgs a) {
  __shared__ WinShared S;
  Arena A = carve(a, blockIdx.x);
  for (;;) {
    if (threadIdx.x == 0) S.job = atomicAdd(a.queue, 1);
    __syncthreads();
    const int job = S.job;
    __syncthreads();
    if (job >= a.n_jobs) break;
    run_window<NT, CPL>(a, A, S, a.order[job]);
  }
}

}  // namespace mpoa

// Host-side launch (pipeline.h).
int mpoa_launch_windows(const mpoa::KernelArgs& a, int n_slots, int nt, int cpl,
                        hipStream_t stream) {
  using namespace mpoa;
#define MPOA_L(NT_, CPL_)                                                                     \
  if (nt == NT_ && cpl == CPL_) {                                                            \
    hipLaunchKernelGGL((k_windows<NT_, CPL_>), dim3(n_slots), dim3(NT_), 0, stream, a);     \
    return hipGetLastError() == hipSuccess ? 0 : -1;                                         \
  }
  MPOA_L(64, 1) MPOA_L(64, 2) MPOA_L(64, 4) MPOA_L(64, 8)
  MPOA_L(128, 2) MPOA_L(128, 4)
  MPOA_L(256, 1) MPOA_L(256, 2) MPOA_L(256, 4)
  MPOA_L(512, 2)
#undef MPOA_L
  return -2;
}
