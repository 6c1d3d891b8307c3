// CPU oracle for minipoa's banded POA path. TEST INFRASTRUCTURE ONLY: the checker that
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg compare the HIP product
// against. Nothing in poa_amd/ links or calls it.
//
// Plain scalar C++ restating the reference semantics; every function cites the
// NCl3-lhd/POA lines it follows. int32 arithmetic throughout, exactly as the reference
// (it has no narrower path): NEG_INF = 0xc0c0c0c0, no saturation.
//
// Parity status of each part (how it is pinned) is recorded in the tests under tests/
// (golden vectors produced by the upstream binary, closed forms and brute force).
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

namespace orc {

constexpr int Q = 8;                        // band quantum: reg_size = 32/sizeof(int), align.cpp:13,19
constexpr int NEG_INF = (int)0xc0c0c0c0;    // align.cpp:22
constexpr int NB = 5;                       // alphabet size para->m, handle_input.cpp:51
constexpr int CODE_N = 4;
constexpr int M_OP = 1, D_OP = 2, I_OP = 4, ALL_OP = 7;  // align.cpp:23

// nt4_table (sequence.cpp:6-23): ACGT/acgt -> 0..3, U/u -> 3, codes 0..3 map to themselves,
// everything else -> 4 (N).
static inline uint8_t code_of(unsigned char c) {
  switch (c) {
    case 0: case 1: case 2: case 3: return c;
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return CODE_N;
  }
}
// nt256_table (sequence.cpp:27-43) restricted to the codes a node can hold.
static inline char char_of(uint8_t code) { return "ACGTN-"[code < 6 ? code : 4]; }

struct Params {
  int match, mismatch, gap_open, gap_ext, b, f, k, mm_w, poa_w, bw, result;
  bool ab_band, seeding, progressive;
};

// initPara (handle_input.cpp:45-77): mismatch/gap scores forced non-positive; the
// `match < 0` branch negates para->m (then overwritten), so match is used as typed.
static Params normalise(const orc_params& p) {
  Params P;
  P.match = p.match;
  P.mismatch = p.mismatch > 0 ? -p.mismatch : p.mismatch;
  P.gap_open = p.gap_open > 0 ? -p.gap_open : p.gap_open;
  P.gap_ext = p.gap_ext > 0 ? -p.gap_ext : p.gap_ext;
  P.b = p.band_b;
  P.f = p.band_f;
  P.ab_band = p.ab_band != 0;
  P.seeding = p.seeding != 0;
  P.k = p.k;
  P.mm_w = p.mm_w;
  P.poa_w = p.poa_w;
  P.bw = 1000;  // main.cpp:65
  P.result = p.result;
  P.progressive = p.progressive != 0;
  return P;
}

// node_t (include/node.h:5-68)
struct Node {
  int id, rank, par_id;
  uint8_t base;
  int ind;
  std::vector<int> in, in_w, out, out_w;
  int aligned[NB];
  std::vector<int> ids, idp;
  Node(int id_, uint8_t base_) : id(id_), rank(0), par_id(id_), base(base_), ind(0) {
    for (int c = 0; c < NB; c++) aligned[c] = -1;
    aligned[base_] = id_;
  }
  // node.h:23-44
  void add_in(int seq_id, int from, int pos) {
    ind++;
    bool found = false;
    for (size_t i = 0; i < in.size(); i++)
      if (in[i] == from) { in_w[i]++; found = true; break; }
    if (!found) { in.push_back(from); in_w.push_back(1); }
    if (ids.empty() || ids.back() != seq_id) { ids.push_back(seq_id); idp.push_back(pos); }
  }
  // node.h:45-61
  void add_out(int seq_id, int to) {
    bool found = false;
    for (size_t i = 0; i < out.size(); i++)
      if (out[i] == to) { out_w[i]++; found = true; break; }
    if (!found) { out.push_back(to); out_w.push_back(1); }
    if (ids.empty() || ids.back() != seq_id) ids.push_back(seq_id);
  }
};

// res_t (include/result.h:3-13)
struct Res {
  int from;
  uint8_t base;
  int aligned_id;
};

struct Graph {
  std::vector<Node> node;
  std::vector<int> rank;        // rank -> node id
  std::vector<int> hlen, tlen;  // by rank
  bool is_topsorted = false;
  std::string cons;
  std::vector<int> cons_pos_to_id;

  // graph::init (graph.cpp:129-135): source 0 and sink 1, both N.
  void init() {
    node.clear();
    node.emplace_back(0, CODE_N);
    node.emplace_back(1, CODE_N);
    rank.clear();
    hlen.clear();
    tlen.clear();
    is_topsorted = false;
  }
  int n_edges() const {
    size_t e = 0;
    for (const Node& v : node) e += v.out.size();
    return (int)e;
  }
  // graph::add_adj (graph.cpp:13-20)
  void add_adj(int seq_id, int from, int to, int pos) {
    if (from == to) return;
    node[from].add_out(seq_id, to);
    node[to].add_in(seq_id, from, pos);
  }
  // graph::add_path (graph.cpp:170-235)
  void add_path(int seq_id, const std::vector<Res>& res, int sink_id) {
    int anchor = -1, pos = 0;
    for (const Res& r : res) {
      int cur = r.from;
      if (cur < 0) {
        if (r.aligned_id >= 0) {
          int existing = node[r.aligned_id].aligned[r.base];
          if (existing == -1) {
            cur = (int)node.size();
            node.emplace_back(cur, r.base);
            node[r.aligned_id].aligned[r.base] = cur;
            node[cur].par_id = r.aligned_id;
          } else {
            cur = existing;
          }
        } else {
          cur = (int)node.size();
          node.emplace_back(cur, r.base);
        }
      }
      if (anchor != -1) add_adj(seq_id, anchor, cur, pos++);
      anchor = cur;
    }
    if (sink_id != -1) add_adj(seq_id, anchor, sink_id, pos);
    is_topsorted = false;
  }
  // graph::build_consensus (graph.cpp:294-350): FIFO BFS from the source; each node keeps
  // the predecessor with the largest in-weight, ties to the larger accumulated score.
  void build_consensus() {
    size_t N = node.size();
    std::vector<int> deg(N), pre(N, 0), score(N, 0);
    for (size_t i = 0; i < N; i++) deg[i] = (int)node[i].in.size();
    std::deque<int> q;
    q.push_back(0);
    while (!q.empty()) {
      int u = q.front();
      q.pop_front();
      const Node& cur = node[u];
      if (u != 0) {
        int wmax = -1, max_pre = -1;
        for (size_t k = 0; k < cur.in.size(); k++) {
          int v = cur.in[k];
          if (cur.in_w[k] > wmax || (cur.in_w[k] == wmax && score[v] > score[max_pre])) {
            wmax = cur.in_w[k];
            max_pre = v;
          }
        }
        pre[u] = max_pre;
        score[u] = score[max_pre] + wmax;
      }
      if (u == 1) break;
      for (int v : cur.out)
        if (--deg[v] == 0) q.push_back(v);
    }
    cons.clear();
    cons_pos_to_id.clear();
    for (int id = pre[1]; id != 0; id = pre[id]) {
      cons.push_back(char_of(node[id].base));
      cons_pos_to_id.push_back(id);
    }
    std::reverse(cons.begin(), cons.end());
    std::reverse(cons_pos_to_id.begin(), cons_pos_to_id.end());
  }
  // graph::topsort (graph.cpp:21-128)
  void topsort(const Params& P, int op) {
    size_t N = node.size();
    std::vector<int> stk, deg(N, 0);
    rank.clear();
    if (op == 1) {
      // Order over aligned-node groups (column heads), LIFO (graph.cpp:24-57).
      for (size_t i = 0; i < N; i++) {
        node[i].rank = -1;
        if (node[i].par_id != node[i].id) continue;
        int sum = 0;
        for (int c = 0; c < NB; c++)
          if (node[i].aligned[c] >= 0) sum += (int)node[node[i].aligned[c]].in.size();
        deg[i] = sum;
        if (sum == 0) stk.push_back((int)i);
      }
      while (!stk.empty()) {
        int u = stk.back();
        stk.pop_back();
        node[u].rank = (int)rank.size();
        rank.push_back(u);
        for (int c = 0; c < NB; c++) {
          int s = node[u].aligned[c];
          if (s < 0) continue;
          node[s].rank = node[u].rank;
          for (int t : node[s].out) {
            int v = node[t].par_id;
            if (--deg[v] == 0) stk.push_back(v);
          }
        }
      }
      is_topsorted = true;
      return;
    }
    // Kahn order, LIFO stack seeded in ascending id order (graph.cpp:59-78).
    for (size_t i = 0; i < N; i++) {
      deg[i] = (int)node[i].in.size();
      if (deg[i] == 0) stk.push_back((int)i);
    }
    while (!stk.empty()) {
      int u = stk.back();
      stk.pop_back();
      node[u].rank = (int)rank.size();
      rank.push_back(u);
      for (int v : node[u].out)
        if (--deg[v] == 0) stk.push_back(v);
    }
    is_topsorted = true;
    if (P.f > 0) {
      // hlen: 1 + hlen of the max-in-weight predecessor (first wins); vector resized, not
      // cleared (graph.cpp:84-99).
      hlen.resize(N);
      for (size_t i = 0; i < rank.size(); i++) {
        const Node& cur = node[rank[i]];
        int wmax = -1, max_pre = -1;
        for (size_t k = 0; k < cur.in.size(); k++)
          if (cur.in_w[k] > wmax) { wmax = cur.in_w[k]; max_pre = node[cur.in[k]].rank; }
        if (max_pre != -1) hlen[cur.rank] = hlen[max_pre] + 1;
      }
      // tlen: 1 + tlen of the max-out-weight successor; the sink is 1 (graph.cpp:101-116).
      tlen.resize(N);
      tlen[node[1].rank] = 1;
      for (int i = (int)rank.size() - 1; i >= 0; i--) {
        const Node& cur = node[rank[i]];
        int wmax = -1, max_suc = -1;
        for (size_t k = 0; k < cur.out.size(); k++)
          if (cur.out_w[k] > wmax) { wmax = cur.out_w[k]; max_suc = node[cur.out[k]].rank; }
        if (max_suc != -1) tlen[cur.rank] = tlen[max_suc] + 1;
      }
    }
    if (P.seeding) build_consensus();
  }
};

struct CallOut {
  int score = 0;
  int qlen = 0;
  int ab_band = 0;
  int64_t cells = 0, edge_cells = 0, logical = 0;
  int max_width = 0;
  int rows = 0;
};

struct DpDump {
  std::vector<int> Bs, Be;
  std::vector<int64_t> off;
  std::vector<int> H, D, I;
};

static inline int calj(int acj, int Bs) { return acj - Bs * Q; }  // align.cpp:25-27

// poa() (align.cpp:1353-1737): banded affine DP of query q[0..qlen) (codes or ASCII) against
// the graph rows of rank beg..end, then the weighted backtrack. Appends the alignment to res.
static int poa(const Params& P, const Graph& G, int beg_id, int end_id, const uint8_t* q,
               int qlen, bool ab_band, std::vector<Res>& res, CallOut* co, DpDump* dump) {
  const std::vector<Node>& node = G.node;
  const std::vector<int>& rank = G.rank;
  const int beg_i = node[beg_id].rank, end_i = node[end_id].rank;
  const int n = end_i - beg_i + 1, m = qlen + 1;
  // a sub-window that ends before it begins in ranks: minipoa writes to an empty vector
  if (n <= 0) return ORC_E_EMPTY_WINDOW;

  // Query: sentinel N, the codes, N padding to col_size (align.cpp:1362-1373). Overlapping -S
  // anchors give qlen < 0: minipoa runs on down to qlen = -16 and its vector of negative size
  // throws std::length_error below that.
  const int col_size = (m + 1 + Q - 1) / Q * Q;
  if (col_size < 0) return ORC_E_EMPTY_WINDOW;
  std::vector<uint8_t> seq(col_size, CODE_N);
  for (int j = 0; j < qlen; j++) seq[1 + j] = code_of(q[j]);

  const int match = P.match, mismatch = P.mismatch, e1 = P.gap_ext, o1 = P.gap_open + e1;
  int w = P.b;
  if (P.f) w += m / P.f;
  const bool adaptive = P.f > 0 && ab_band;

  std::vector<int> Ms(n, m + 1), Me(n, 0), Bs(n), Be(n);
  std::vector<int> Mp(n, 0), Sl(n, 0), Pl(n, m), Pr(n, 0), Ol(n, 0), Or(n, 0), Ow(n, 0);
  std::vector<int64_t> off(n + 1, 0);
  const std::vector<int>& hlen = G.hlen;
  const std::vector<int>& tlen = G.tlen;

  // Static band or full width (align.cpp:1378-1394).
  for (int i = 0; i < n; i++) {
    int aci = beg_i + i;
    if (P.f > 0 && !ab_band) {
      int lo = hlen[aci] - hlen[beg_i], hi = m - tlen[aci] + tlen[end_i];
      Ms[i] = std::max(0, std::min(lo, hi) - w);
      Me[i] = std::min(m, std::max(lo, hi) + w);
    } else {
      Ms[i] = 0;
      Me[i] = m;
    }
    Bs[i] = Ms[i] / Q;
    Be[i] = Me[i] / Q + 2;
    off[i + 1] = off[i] + (int64_t)(Be[i] - Bs[i]) * Q;
  }
  // Grow-only per-thread DP buffer, as aligned_buff_t (mem_alloc_utils.h:10-45). Every
  // entry the DP or the backtrack reads is written first; I of row 0 is never read and is
  // set to NEG_INF.
  thread_local std::vector<int> pool;
  if ((int64_t)pool.size() < 3 * off[n]) pool.resize(3 * off[n]);
  int* const H = pool.data();
  int* const D = H + off[n];
  int* const I = D + off[n];
  for (int64_t j = off[0]; j < off[1]; j++) I[j] = NEG_INF;

  if (adaptive) {  // row 0's adaptive band (align.cpp:1410-1414)
    int hi = m - tlen[beg_i] + tlen[end_i];
    Ms[0] = std::max(0, std::min({Pl[0], hlen[beg_i] - hlen[beg_i] + Ol[0], hi}) - w);
    Me[0] = std::min(m, std::max({Pr[0], hlen[beg_i] - hlen[beg_i] + Or[0], hi}) + w);
    Bs[0] = Ms[0] / Q;
    Be[0] = Me[0] / Q + 2;
  }
  // Row 0: M and D are NEG_INF over blocks [Bs0, Be0) addressed from the row start, then
  // M[0][0] = 0 (align.cpp:1415-1425). Bs0 is always 0 (lo = 0 makes min(...) <= 0).
  for (int bid = Bs[0]; bid < Be[0]; bid++)
    for (int l = 0; l < Q; l++) H[off[0] + bid * Q + l] = D[off[0] + bid * Q + l] = NEG_INF;
  H[off[0]] = 0;

  CallOut stat;
  stat.rows = n;
  stat.qlen = qlen;
  stat.ab_band = adaptive ? 1 : 0;
  for (int i = 1; i < n; i++) {
    const int aci = beg_i + i;
    const Node& cur = node[rank[aci]];
    if (adaptive) {  // align.cpp:1433-1444
      int lo = hlen[aci] - hlen[beg_i], hi = m - tlen[aci] + tlen[end_i];
      Ms[i] = std::max(0, std::min({Pl[i], lo + Ol[i], hi}) - w - Ow[i]);
      Me[i] = std::min(m, std::max({Pr[i], lo + Or[i], hi}) + w + Ow[i]);
      Bs[i] = Ms[i] / Q;
      Be[i] = Me[i] / Q + 2;
    }
    const int block_num = Be[i] - Bs[i] - 1;
    if (block_num < 1) return ORC_E_BAND;
    const int W = block_num * Q;
    int* Hi = &H[off[i]];
    int* Di = &D[off[i]];
    int* Ii = &I[off[i]];
    // First visited predecessor starts from NEG_INF; with none, the row stays NEG_INF; the
    // block after the band is a NEG_INF sentinel (align.cpp:1461-1506).
    for (int j = 0; j < W + Q; j++) Hi[j] = Di[j] = NEG_INF;
    int n_pre = 0;
    for (size_t k = 0; k < cur.in.size(); k++) {
      const Node& pre = node[cur.in[k]];
      const int p = pre.rank - beg_i;
      if (p < 0) continue;
      n_pre++;
      const int pbase = p == 0 ? CODE_N : pre.base;
      const int* Hp = &H[off[p]];
      const int* Dp = &D[off[p]];
      // The per-block conditions of align.cpp:1467 and :1484 select contiguous block
      // ranges, walked here column by column. Row i computes columns [c0, c1).
      const int c0 = Bs[i] * Q, c1 = c0 + W, pofs = Bs[p] * Q;
      // M-source blocks: Bs[p] <= max(0, acBid-1) and acBid <= Be[p]-2. Lane 0 of block 0
      // takes NEG_INF as its diagonal (align.cpp:1476); at column 0 the query char is N.
      {
        const int blo = Bs[p] == 0 ? 0 : Bs[p] + 1;
        int lo = std::max(c0, blo * Q);
        const int hi = std::min(c1, (Be[p] - 1) * Q);
        if (lo == c0 && lo < hi) {
          const int qc = c0 == 0 ? CODE_N : seq[c0 - 1];
          Hi[0] = std::max(Hi[0], NEG_INF + (pbase == qc ? match : mismatch));
          lo++;
        }
        const int* hp = Hp - pofs - 1;
        const uint8_t* sq = seq.data() - 1;
        int* ho = Hi - c0;
        for (int c = lo; c < hi; c++) {
          const int v = hp[c] + (pbase == sq[c] ? match : mismatch);
          ho[c] = std::max(ho[c], v);
        }
      }
      // D-source blocks: Bs[p] <= acBid < Be[p]-1.
      {
        const int lo = std::max(c0, Bs[p] * Q), hi = std::min(c1, (Be[p] - 1) * Q);
        const int* hp = Hp - pofs;
        const int* dp = Dp - pofs;
        int* dout = Di - c0;
        for (int c = lo; c < hi; c++)
          dout[c] = std::max(dout[c], std::max(dp[c] + e1, hp[c] + o1));
      }
    }
    // In-row I/H scan (align.cpp:1507-1514).
    {
      int iv = NEG_INF, hv = std::max({Hi[0], Di[0], iv});
      Ii[0] = iv;
      Hi[0] = hv;
      for (int j = 1; j < W; j++) {
        iv = std::max(iv + e1, hv + o1);
        hv = std::max(std::max(Hi[j], Di[j]), iv);
        Ii[j] = iv;
        Hi[j] = hv;
      }
    }
    for (int j = W; j < W + Q; j++) Ii[j] = NEG_INF;

    stat.cells += W;
    stat.edge_cells += (int64_t)n_pre * W;
    stat.logical += Me[i] - Ms[i] + 1;
    stat.max_width = std::max(stat.max_width, W);

    if (adaptive) {  // adaptive-band update (align.cpp:1516-1568)
      int max_j = 0;
      for (int j = 1; j < W; j++) max_j = Hi[j] > Hi[max_j] ? j : max_j;
      const int max_acj = Bs[i] * Q + max_j;
      int max_slp = 0;
      bool is_match = false;
      Mp[i] = max_acj;
      if (cur.base == seq[max_acj]) {
        for (size_t k = 0; k < cur.in.size(); k++) {
          int p = node[cur.in[k]].rank - beg_i;
          if (p < 0) continue;
          max_slp = Sl[p] > max_slp ? Sl[p] : max_slp;
          if (Mp[p] + 1 == Mp[i]) {
            Sl[i] = std::max(Sl[i], Sl[p] + 1);
            is_match = true;
          }
        }
      }
      if (!is_match && max_slp < 30) {
        Ow[i] += max_slp + 1;
        Ow[i] = std::min(Ow[i], 5 * w);
      }
      if (Sl[i] >= 30) {
        // hlen indexed by the local row number here (align.cpp:1544-1545).
        Ol[i] = max_acj - (hlen[i] - hlen[beg_i]);
        Or[i] = max_acj - (hlen[i] - hlen[beg_i]);
        Ow[i] = 0;
        for (int t : cur.out) {
          int s = node[t].rank - beg_i;
          if (s >= n) continue;
          Ol[s] = Ol[i];
          Or[s] = Or[i];
        }
      }
      for (int t : cur.out) {
        int s = node[t].rank - beg_i;
        if (s >= n) continue;
        Ol[s] = std::min(Ol[s], Ol[i]);
        Or[s] = std::max(Or[s], Or[i]);
      }
      for (int t : cur.out) {
        int s = node[t].rank - beg_i;
        if (s >= n) continue;
        Pl[s] = std::min(Pl[s], max_acj);
        Pr[s] = std::max(Pr[s], max_acj + 1);
        Ow[s] = Ow[i] ? std::max(Ow[s], Ow[i]) : Ow[i];
      }
    }
  }

  auto Hat = [&](int r, int j) -> int { return H[off[r] + j]; };
  auto Dat = [&](int r, int j) -> int { return D[off[r] + j]; };
  auto Iat = [&](int r, int j) -> int { return I[off[r] + j]; };

  int i = n - 1, acj = m;
  stat.score = Hat(i, calj(acj, Bs[i]));
  if (co) *co = stat;
  if (dump) {
    dump->Bs = Bs;
    dump->Be = Be;
    dump->off = off;
    dump->H.assign(H, H + off[n]);
    dump->D.assign(D, D + off[n]);
    dump->I.assign(I, I + off[n]);
  }

  // Backtrack (align.cpp:1572-1736).
  std::vector<Res> out;
  out.reserve(m + 1);
  int op = ALL_OP;
  auto emit_sub = [&](const Node& pre, int p) {
    const int pbase = p == 0 ? CODE_N : pre.base;
    const uint8_t qc = seq[acj - 1];
    if (pbase == qc) {
      out.push_back({pre.id, pre.base, -1});
    } else {
      const Node& par = node[pre.par_id];
      if (par.aligned[qc] != -1) out.push_back({par.aligned[qc], qc, -1});
      else out.push_back({-1, qc, pre.par_id});
    }
  };
  while (i > 0 || acj > 0) {
    const int j = calj(acj, Bs[i]);
    const int aci = beg_i + i;
    if (acj > (Be[i] - 1) * Q || acj < Bs[i] * Q) return ORC_E_BAND;
    const Node& cur = node[rank[aci]];
    const int h = Hat(i, j);
    // 1. match-only move with the in-weight rule (align.cpp:1591-1626)
    if ((op & M_OP) && acj > 0) {
      int bk = -1;
      for (size_t k = 0; k < cur.in.size(); k++) {
        const Node& pre = node[cur.in[k]];
        int p = pre.rank - beg_i;
        if (p < 0) continue;
        int pbase = p == 0 ? CODE_N : pre.base;
        if (pbase != seq[acj - 1]) continue;
        int pj = calj(acj, Bs[p]);
        int bn = Be[p] - Bs[p] - 1;
        if (pj - 1 >= 0 && pj - 1 < bn * Q && h == Hat(p, pj - 1) + match &&
            (bk == -1 || cur.in_w[k] > cur.in_w[bk]))
          bk = (int)k;
      }
      if (bk != -1 && cur.in_w[bk] >= cur.ind / 10) {
        const Node& pre = node[cur.in[bk]];
        int p = pre.rank - beg_i;
        emit_sub(pre, p);
        op = ALL_OP;
        i = p;
        acj--;
        continue;
      }
    }
    // 2. deletion: open first, then extend (align.cpp:1627-1667)
    if (op & D_OP) {
      if (op == D_OP || h == Dat(i, j)) {
        int bk = -1, bop = 0;
        for (int pass = 0; pass < 2 && bk == -1; pass++) {
          for (size_t k = 0; k < cur.in.size(); k++) {
            const Node& pre = node[cur.in[k]];
            int p = pre.rank - beg_i;
            if (p < 0) continue;
            if (Bs[p] <= acj / Q && acj / Q < Be[p] - 1) {
              int pj = calj(acj, Bs[p]);
              int src = pass == 0 ? Hat(p, pj) + o1 : Dat(p, pj) + e1;
              if (Dat(i, j) == src && (bk == -1 || cur.in_w[k] > cur.in_w[bk])) {
                bk = (int)k;
                bop = pass == 0 ? (M_OP | I_OP) : D_OP;
              }
            }
          }
        }
        if (bk != -1) {
          i = node[cur.in[bk]].rank - beg_i;
          op = bop;
          continue;
        }
      }
    }
    // 3. insertion: open first, then extend (align.cpp:1668-1688)
    if ((op & I_OP) && acj > 0) {
      if (j - 1 >= 0) {
        if (op == I_OP || h == Iat(i, j)) {
          if (Iat(i, j) == Hat(i, j - 1) + o1) {
            out.push_back({-1, seq[acj - 1], -1});
            op = M_OP | D_OP;
            acj--;
            continue;
          }
          if (Iat(i, j) == Iat(i, j - 1) + e1) {
            out.push_back({-1, seq[acj - 1], -1});
            op = I_OP;
            acj--;
            continue;
          }
        }
      }
    }
    // 4. any substitution, no in-weight rule (align.cpp:1689-1728)
    if ((op & M_OP) && acj > 0) {
      int bk = -1;
      for (size_t k = 0; k < cur.in.size(); k++) {
        const Node& pre = node[cur.in[k]];
        int p = pre.rank - beg_i;
        if (p < 0) continue;
        int pbase = p == 0 ? CODE_N : pre.base;
        int pj = calj(acj, Bs[p]);
        int bn = Be[p] - Bs[p] - 1;
        if (pj - 1 >= 0 && pj - 1 < bn * Q &&
            h == Hat(p, pj - 1) + (pbase == seq[acj - 1] ? match : mismatch) &&
            (bk == -1 || cur.in_w[k] > cur.in_w[bk]))
          bk = (int)k;
      }
      if (bk != -1) {
        op = ALL_OP;
        const Node& pre = node[cur.in[bk]];
        int p = pre.rank - beg_i;
        emit_sub(pre, p);
        i = p;
        acj--;
        continue;
      }
    }
    return ORC_E_BACKTRACK;
  }
  std::reverse(out.begin(), out.end());
  res.insert(res.end(), out.begin(), out.end());
  return ORC_OK;
}

// ---------------------------------------------------------------------------------------
// Seeding (-S): minimizers, anchors, local chains, global chain.
// ---------------------------------------------------------------------------------------

struct mm128 {
  uint64_t x, y;
};

// hash64 (minimizer.cpp:11-21), an invertible 64-bit integer hash.
static inline uint64_t hash64(uint64_t key, uint64_t mask) {
  key = (~key + (key << 21)) & mask;
  key = key ^ key >> 24;
  key = ((key + (key << 3)) + (key << 8)) & mask;
  key = key ^ key >> 14;
  key = ((key + (key << 2)) + (key << 4)) & mask;
  key = key ^ key >> 28;
  key = (key + (key << 31)) & mask;
  return key;
}

// mm_sketch (minimizer.cpp:60-135) with is_hpc = 0 and the strand fixed to 0 (:96):
// x = hash<<8 | span, y = rid<<32 | last_pos<<1. Appends to p.
static void mm_sketch(const char* str, int len, int w, int k, uint32_t rid,
                      std::vector<mm128>& p) {
  const uint64_t shift1 = 2 * (uint64_t)(k - 1), mask = (1ULL << 2 * k) - 1;
  uint64_t kmer[2] = {0, 0};
  int l = 0, buf_pos = 0, min_pos = 0, kmer_span = 0;
  mm128 buf[256];
  const mm128 none = {UINT64_MAX, UINT64_MAX};
  mm128 mn = none;
  for (int j = 0; j < w; j++) buf[j] = none;
  for (int i = 0; i < len; ++i) {
    int c = code_of((uint8_t)str[i]);
    mm128 info = none;
    if (c < 4) {
      kmer_span = l + 1 < k ? l + 1 : k;
      kmer[0] = (kmer[0] << 2 | (uint64_t)c) & mask;
      kmer[1] = (kmer[1] >> 2) | (3ULL ^ (uint64_t)c) << shift1;
      ++l;
      if (l >= k && kmer_span < 256) {
        info.x = hash64(kmer[0], mask) << 8 | (uint64_t)kmer_span;
        info.y = (uint64_t)rid << 32 | (uint32_t)i << 1;
      }
    } else {
      l = 0;
      kmer_span = 0;
    }
    buf[buf_pos] = info;
    if (l == w + k - 1 && mn.x != UINT64_MAX) {
      for (int j = buf_pos + 1; j < w; ++j)
        if (mn.x == buf[j].x && buf[j].y != mn.y) p.push_back(buf[j]);
      for (int j = 0; j < buf_pos; ++j)
        if (mn.x == buf[j].x && buf[j].y != mn.y) p.push_back(buf[j]);
    }
    if (info.x <= mn.x) {
      if (l >= w + k && mn.x != UINT64_MAX) p.push_back(mn);
      mn = info;
      min_pos = buf_pos;
    } else if (buf_pos == min_pos) {
      if (l >= w + k - 1 && mn.x != UINT64_MAX) p.push_back(mn);
      mn.x = UINT64_MAX;
      for (int j = buf_pos + 1; j < w; ++j)
        if (mn.x >= buf[j].x) { mn = buf[j]; min_pos = j; }
      for (int j = 0; j <= buf_pos; ++j)
        if (mn.x >= buf[j].x) { mn = buf[j]; min_pos = j; }
      if (l >= w + k - 1 && mn.x != UINT64_MAX) {
        for (int j = buf_pos + 1; j < w; ++j)
          if (mn.x == buf[j].x && mn.y != buf[j].y) p.push_back(buf[j]);
        for (int j = 0; j <= buf_pos; ++j)
          if (mn.x == buf[j].x && mn.y != buf[j].y) p.push_back(buf[j]);
      }
    }
    if (++buf_pos == w) buf_pos = 0;
  }
  if (mn.x != UINT64_MAX) p.push_back(mn);
}

// In-place MSD radix sort on .x, 8 bits per pass, insertion sort at <= 64 elements: the
// klib algorithm minipoa sorts minimizers, anchors and chains with (ksort.h:101-151). It is
// not stable, so the order of equal keys is part of the semantics and the same permutation
// steps are followed.
static void rs_insertsort(mm128* beg, mm128* end) {
  for (mm128* i = beg + 1; i < end; ++i) {
    if (i->x < (i - 1)->x) {
      mm128 tmp = *i;
      mm128* j = i;
      for (; j > beg && tmp.x < (j - 1)->x; --j) *j = *(j - 1);
      *j = tmp;
    }
  }
}
static void rs_sort(mm128* beg, mm128* end, int n_bits, int s) {
  struct Bucket {
    mm128 *b, *e;
  };
  const int size = 1 << n_bits, msk = size - 1;
  Bucket bk[256];
  Bucket* be = bk + size;
  for (Bucket* k = bk; k != be; ++k) k->b = k->e = beg;
  for (mm128* i = beg; i != end; ++i) ++bk[i->x >> s & msk].e;
  for (Bucket* k = bk + 1; k != be; ++k) k->e += (k - 1)->e - beg, k->b = (k - 1)->e;
  for (Bucket* k = bk; k != be;) {
    if (k->b != k->e) {
      Bucket* l = bk + (k->b->x >> s & msk);
      if (l != k) {
        mm128 tmp = *k->b, swp;
        do {
          swp = tmp;
          tmp = *l->b;
          *l->b++ = swp;
          l = bk + (tmp.x >> s & msk);
        } while (l != k);
        *k->b++ = tmp;
      } else {
        ++k->b;
      }
    } else {
      ++k;
    }
  }
  bk[0].b = beg;
  for (Bucket* k = bk + 1; k != be; ++k) k->b = (k - 1)->e;
  if (s) {
    s = s > n_bits ? s - n_bits : 0;
    for (Bucket* k = bk; k != be; ++k) {
      if (k->e - k->b > 64) rs_sort(k->b, k->e, n_bits, s);
      else if (k->e - k->b > 1) rs_insertsort(k->b, k->e);
    }
  }
}
static void radix_sort(mm128* beg, mm128* end) {
  if (end - beg <= 64) rs_insertsort(beg, end);
  else rs_sort(beg, end, 8, 56);
}

// mg_log2 (include/mmpriv.h:12-20): a quadratic fit of log2 on the mantissa.
static inline float mg_log2(float x) {
  union {
    float f;
    uint32_t i;
  } z = {x};
  float log_2 = (float)(int)((z.i >> 23) & 255) - 128;
  z.i &= ~(255u << 23);
  z.i += 127u << 23;
  log_2 += (-0.34484843f * z.f + 2.02466578f) * z.f - 0.67487759f;
  return log_2;
}

constexpr int CH_INF = 1000000000;  // lchain.cpp:7

// comput_sc (lchain.cpp:28-63) with is_equal = false and is_cdna = 0, as minipoa calls it.
static inline int32_t comput_sc(const mm128* ai, const mm128* aj, int32_t max_dist_x,
                                float chn_pen_gap, float chn_pen_skip) {
  int32_t dq = (int32_t)ai->y - (int32_t)aj->y;
  if (dq <= 0 || dq > max_dist_x) return -CH_INF;
  int32_t dr = (int32_t)(ai->x - aj->x);
  int32_t dd = dr > dq ? dr - dq : dq - dr;
  int32_t dg = dr < dq ? dr : dq;
  int32_t q_span = aj->y >> 32 & 0xff;
  int32_t sc = q_span < dg ? q_span : dg;
  if (dd || dg > q_span) {
    float lin_pen = chn_pen_gap * (float)dd + chn_pen_skip * (float)dg;
    float log_pen = dd >= 1 ? mg_log2((float)(dd + 1)) : 0.0f;
    if (dr == 0) ++sc;
    else sc -= (int)(lin_pen < log_pen ? lin_pen : log_pen);
  }
  return sc;
}

// mg_chain_bk_end (lchain.cpp:9-26)
static int64_t chain_bk_end(int32_t max_drop, const mm128* z, const int32_t* f, const int64_t* p,
                            int32_t* t, int64_t k) {
  int64_t i = z[k].y, end_i = -1, max_i = i;
  int32_t max_s = 0;
  if (i < 0 || t[i] != 0) return i;
  do {
    int32_t s;
    t[i] = 2;
    end_i = i = p[i];
    s = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
    if (s > max_s) max_s = s, max_i = i;
    else if (max_s - s > max_drop) break;
  } while (i >= 0 && t[i] == 0);
  for (i = z[k].y; i >= 0 && i != end_i; i = p[i]) t[i] = 0;
  return max_i;
}

// mg_chain_backtrack (lchain.cpp:66-135). `t` here has one extra leading slot so that
// t[-1] reads 0: upstream indexes t[p[i]] with p[i] = -1, which lands on the high half of
// kalloc's size header, always 0 for these sizes.
static std::vector<mm128> chain_backtrack(int64_t n, const int32_t* f, const mm128* a,
                                          const int64_t* p, int32_t* v, int32_t* t,
                                          int32_t min_cnt, int32_t min_sc, int32_t max_drop,
                                          int32_t* n_v_) {
  std::vector<mm128> u;
  *n_v_ = 0;
  std::memset(t, 0, n * sizeof(int32_t));
  int64_t n_z = 0;
  for (int64_t i = n - 1; i >= 0; i--) {
    if (t[p[i]] == 0 && f[i] >= 0 && f[i] >= min_sc) {
      ++n_z;
      t[i] = 2;
    }
    if (p[i] >= 0) t[p[i]] = 1;
  }
  if (n_z == 0) return u;
  std::vector<mm128> z(n_z);
  for (int64_t i = 0, k = 0; i < n; ++i)
    if (t[i] == 2) z[k].x = (uint64_t)(int64_t)f[i], z[k++].y = (uint64_t)i;
  radix_sort(z.data(), z.data() + n_z);

  std::memset(t, 0, n * 4);
  int64_t n_v = 0;
  for (int64_t k = n_z - 1; k >= 0; --k) {
    if (t[z[k].y] == 0) {
      int64_t n_v0 = n_v, end_i, i;
      int32_t sc;
      end_i = chain_bk_end(max_drop, z.data(), f, p, t, k);
      for (i = (int64_t)z[k].y; i != end_i; i = p[i]) v[n_v++] = (int32_t)i, t[i] = 1;
      sc = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
      if (sc >= min_sc && n_v > n_v0 && n_v - n_v0 >= min_cnt) {
        uint64_t end_id = z[k].y;
        uint64_t strand = a[end_id].x >> 63, tpos = (uint32_t)a[end_id].x,
                 qpos = (uint32_t)a[end_id].y;
        u.push_back({strand << 63 | tpos << 32 | qpos, (uint64_t)n_v0 << 32 | (uint64_t)n_v});
      } else {
        n_v = n_v0;
      }
    }
  }
  *n_v_ = (int32_t)n_v;
  return u;
}

// mg_lchain_dp (lchain.cpp:158-238) + compact_a (:136-156) with minipoa's arguments.
// Returns the chained anchors (chain after chain, each in ascending order) and the chains.
static std::vector<mm128> lchain_dp(int max_dist_x, int max_dist_y, int bw, int max_skip,
                                    int max_iter, int min_cnt, int min_sc, float chn_pen_gap,
                                    float chn_pen_skip, const std::vector<mm128>& a,
                                    std::vector<mm128>& u) {
  (void)max_dist_y;
  u.clear();
  const int64_t n = (int64_t)a.size();
  std::vector<mm128> b;
  if (n == 0) return b;
  const int32_t max_drop = bw;
  if (max_dist_x < bw) max_dist_x = bw;
  std::vector<int64_t> p(n);
  std::vector<int32_t> f(n), v(n), tbuf(n + 1, 0);
  int32_t* t = tbuf.data() + 1;  // t[-1] == 0, see chain_backtrack
  int64_t st = 0, max_ii = -1;
  for (int64_t i = 0; i < n; ++i) {
    int64_t max_j = -1;
    int32_t max_f = a[i].y >> 32 & 0xff, n_skip = 0, n_iter = 0;
    while (st < i && (a[i].x >> 32 != a[st].x >> 32 || a[i].x > a[st].x + max_dist_x)) ++st;
    if (i - st > 5000) st = i - 5000;
    int64_t j;
    for (j = i - 1; j >= st; --j) {
      int32_t sc = comput_sc(&a[i], &a[j], max_dist_x, chn_pen_gap, chn_pen_skip);
      if (sc == -CH_INF) continue;
      sc += f[j];
      if (sc > max_f) {
        max_f = sc, max_j = j;
        if (n_skip > 0) --n_skip;
      } else {
        if (sc != max_f) {
          if (++n_iter > max_iter) break;
        }
        if (t[j] == (int32_t)i) {
          if (++n_skip > max_skip) break;
        }
      }
      if (p[j] >= 0) t[p[j]] = (int32_t)i;
    }
    f[i] = max_f, p[i] = max_j;
    if (max_ii < 0 || (a[i].x - a[max_ii].x <= (uint64_t)(int64_t)max_dist_x && f[max_ii] < f[i]))
      max_ii = i;
  }
  int32_t n_v = 0;
  u = chain_backtrack(n, f.data(), a.data(), p.data(), v.data(), t, min_cnt, min_sc, max_drop,
                      &n_v);
  if (u.empty()) return b;
  b.reserve(n_v);
  for (const mm128& c : u) {
    int32_t s0 = (int32_t)(c.y >> 32), e0 = (int32_t)c.y;
    for (int32_t j2 = e0 - 1; j2 >= s0; j2--) b.push_back(a[v[j2]]);
  }
  return b;
}

// get_local_chain_score (lchain.cpp:242-253)
static int local_chain_score(int end_tpos_j, int end_qpos_j, int start_i, int end_i,
                             const mm128* a, const int* score) {
  int l = start_i, r = end_i;
  while (l + 1 < r) {
    int mid = (l + r) / 2;
    int tpos_mid = (uint32_t)a[mid].x, qpos_mid = (uint32_t)a[mid].y;
    if (tpos_mid <= end_tpos_j || qpos_mid <= end_qpos_j) l = mid;
    else r = mid;
  }
  if (r == end_i) return -CH_INF;
  return score[end_i - 1] - score[r];
}

// chain_dp (lchain.cpp:261-386): best global chain over local chains, then anchors picked
// back-to-front at least min_w apart on both sequences.
static std::vector<mm128> global_chain(const std::vector<mm128>& lchains, std::vector<mm128>& a,
                                       int min_w, int tlen, int qlen, int max_dist_x,
                                       float chn_pen_gap, float chn_pen_skip) {
  const int n_l = (int)lchains.size();
  std::vector<int32_t> f(a.size());
  std::vector<int> chain_score(n_l), pre_chain(n_l);
  int global_max_score = -CH_INF, global_max_i = -1;
  for (int i = 0, st = 0; i < n_l; i++) {
    uint64_t xi = lchains[i].x, yi = lchains[i].y;
    int strand = (int)(xi >> 63), start_i = (int)(yi >> 32), end_i = (int32_t)yi;
    int start_tpos_i = (int32_t)a[start_i].x, start_qpos_i = (int32_t)a[start_i].y;
    f[start_i] = a[start_i].y >> 32 & 0xff;
    for (int j = start_i + 1; j < end_i; j++)
      f[j] = f[j - 1] + comput_sc(&a[j], &a[j - 1], max_dist_x, chn_pen_gap, chn_pen_skip);
    int max_j = -1, max_score = f[end_i - 1];
    while (st < i) {
      if ((int)(lchains[st].x >> 63) != strand) ++st;
      else break;
    }
    for (int j = i - 1; j >= st; --j) {
      uint64_t xj = lchains[j].x, yj = lchains[j].y;
      int end_tpos_j = (int)((xj >> 32) & 0x7fffffff), end_qpos_j = (int32_t)xj;
      int start_j = (int)(yj >> 32);
      int start_tpos_j = (int32_t)a[start_j].x;
      if (start_tpos_i < start_tpos_j) continue;
      int score1;
      if (start_tpos_i > end_tpos_j && start_qpos_i > end_qpos_j)
        score1 = chain_score[j] + f[end_i - 1];
      else
        score1 = chain_score[j] +
                 local_chain_score(end_tpos_j, end_qpos_j, start_i, end_i, a.data(), f.data());
      if (score1 > max_score) {
        max_score = score1;
        max_j = j;
      }
    }
    chain_score[i] = max_score;
    pre_chain[i] = max_j;
    if (max_score > global_max_score) {
      global_max_score = max_score;
      global_max_i = i;
    }
  }
  std::vector<mm128> out;
  if (global_max_i < 0) return a;  // upstream returns leaving the anchors untouched
  int cur_i = global_max_i, pre_i = pre_chain[global_max_i];
  uint64_t cur_y = lchains[cur_i].y;
  int last_tpos = tlen, last_qpos = qlen;
  int i = (int)(uint32_t)cur_y - 1;
  int filter_num = 0;
  while (pre_i != -1) {
    uint64_t pre_x = lchains[pre_i].x, pre_y = lchains[pre_i].y;
    int pre_end_tpos = (int)((pre_x >> 32) & 0x7fffffff), pre_end_qpos = (int32_t)pre_x;
    while (i >= 0) {
      int cur_tpos = (int)(uint32_t)a[i].x, cur_qpos = (int32_t)a[i].y;
      if (cur_tpos > pre_end_tpos && cur_qpos > pre_end_qpos) {
        if (last_tpos - cur_tpos >= min_w && last_qpos - cur_qpos >= min_w) {
          if (filter_num < min_w / 10) a[i].y |= 1ULL << 40;  // MM_SEED_BAND_MODE_MASK
          filter_num = 0;
          out.push_back(a[i]);
          last_tpos = cur_tpos, last_qpos = cur_qpos;
        } else {
          filter_num++;
        }
      } else {
        break;
      }
      i--;
      if ((uint64_t)(int64_t)i == (cur_y >> 32)) break;
    }
    cur_i = pre_i, pre_i = pre_chain[pre_i], cur_y = pre_y;
    i = (int)(uint32_t)cur_y - 1;
  }
  while (i >= 0) {
    int cur_tpos = (int)(uint32_t)a[i].x, cur_qpos = (int32_t)a[i].y;
    if (last_tpos - cur_tpos >= min_w && last_qpos - cur_qpos >= min_w) {
      if (filter_num < min_w / 10) a[i].y |= 1ULL << 40;
      filter_num = 0;
      out.push_back(a[i]);
      last_tpos = cur_tpos, last_qpos = cur_qpos;
    } else {
      filter_num++;
    }
    if ((uint64_t)(int64_t)i == (cur_y >> 32)) break;
    i--;
  }
  std::reverse(out.begin(), out.end());
  return out;
}

// minimizer_t state (minimizer.cpp:142-194): every read's minimizers sorted by hash; the
// consensus sketch is appended at slot n_reads per call.
struct Seeder {
  int n_reads = 0;
  std::vector<mm128> sorted;   // all reads, then the consensus
  std::vector<int64_t> mm_h;   // n_reads + 2 offsets
  void init(const Params& P, int n, const char* seq, const int64_t* off) {
    n_reads = n;
    sorted.clear();
    mm_h.assign(n + 2, 0);
    for (int i = 0; i < n; i++) {
      mm_sketch(seq + off[i], (int)(off[i + 1] - off[i]), P.mm_w, P.k, (uint32_t)i, sorted);
      mm_h[i + 1] = (int64_t)sorted.size();
    }
    for (int i = 0; i < n; i++) radix_sort(sorted.data() + mm_h[i], sorted.data() + mm_h[i + 1]);
  }
  // collect_anchors_bycons (minimizer.cpp:537-570) + collect_anchors (:404-442) +
  // dp_chaining (:493-535).
  std::vector<mm128> anchors_bycons(const Params& P, int qid, int qlen, const std::string& cons) {
    const int n = n_reads;
    sorted.resize(mm_h[n]);
    mm_sketch(cons.data(), (int)cons.size(), P.mm_w, P.k, (uint32_t)n, sorted);
    mm_h[n + 1] = (int64_t)sorted.size();
    radix_sort(sorted.data() + mm_h[n], sorted.data() + mm_h[n + 1]);

    std::vector<mm128> anchors;
    int64_t i = mm_h[n], j = mm_h[qid], jj = 0;
    const int64_t ie = mm_h[n + 1], je = mm_h[qid + 1];
    while (i < ie && j < je) {
      uint64_t xi = sorted[i].x, xj = sorted[j].x;
      if (xi == xj) {
        int64_t ii = i;
        for (; ii < ie; ++ii) {
          if (sorted[ii].x != xi) break;
          uint64_t yi = sorted[ii].y;
          for (jj = j; jj < je; ++jj) {
            if (sorted[jj].x != xj) break;
            uint64_t yj = sorted[jj].y;
            uint32_t qspan = (uint32_t)(sorted[jj].x & 0xff);
            mm128 an;
            an.x = (yi & 0x7fffffff00000000ULL) | ((uint32_t)yi >> 1);
            an.y = (uint64_t)qspan << 32 | ((uint32_t)yj >> 1);
            anchors.push_back(an);
          }
        }
        i = ii, j = jj;
      } else if (xi < xj) {
        ++i;
      } else {
        ++j;
      }
    }
    radix_sort(anchors.data(), anchors.data() + anchors.size());

    const int tlen = (int)cons.size();
    int min_w = std::min(20000, (tlen + 24 - 1) / 24);
    if (P.poa_w) min_w = std::min(min_w, P.poa_w + P.k);
    const int max_dis = P.bw;
    const float chn_pen_gap = (float)(0.01 * P.k), chn_pen_skip = (float)(0.0 * 0.01 * P.k);
    std::vector<mm128> lchains;
    std::vector<mm128> chained = lchain_dp(max_dis, max_dis, P.bw, 25, 50, 3, 100, chn_pen_gap,
                                           chn_pen_skip, anchors, lchains);
    if (lchains.empty()) return {};
    radix_sort(lchains.data(), lchains.data() + lchains.size());
    return global_chain(lchains, chained, min_w, tlen, qlen, max_dis, chn_pen_gap, chn_pen_skip);
  }
};

// minimizer_t::get_guide_tree (minimizer.cpp:195-386): the order reads are added in with -p.
// Minimizer hits between every two reads (the smaller count per shared minimizer, summed) over
// all minimizers sorted by hash; each read's own count; the Jaccard similarity of every ordered
// pair weighted by 0.51 and 0.49 of the two lengths over the longest; the pair of the largest
// similarity (first in row-major order) opens the order, then each next read is the one whose
// similarities from the reads already placed, summed in their order, are largest (first wins).
static std::vector<int> guide_tree(const Params& P, int n, const char* seq, const int64_t* off) {
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  if (n < 2) return ord;
  std::vector<mm128> mm;
  std::vector<int> len(n);
  int maxl = 0;
  for (int i = 0; i < n; i++) {
    len[i] = (int)(off[i + 1] - off[i]);
    maxl = std::max(maxl, len[i]);
    mm_sketch(seq + off[i], len[i], P.mm_w, P.k, (uint32_t)i, mm);
  }
  if (mm.empty()) return ord;
  radix_sort(mm.data(), mm.data() + mm.size());
  // hit[i*(i+1)/2 + j], j <= i: shared minimizers of reads i and j; j == i: read i's total
  std::vector<long> hit((size_t)n * (n + 1) / 2, 0);
  std::vector<int> cnt(n);
  size_t g = 0;
  for (size_t t = 1; t <= mm.size(); t++) {
    if (t < mm.size() && mm[t].x == mm[g].x) continue;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t u = g; u < t; u++) {
      const int r = (int)(mm[u].y >> 32);
      cnt[r]++;
      hit[(size_t)r * (r + 1) / 2 + r]++;
    }
    for (int r1 = 0; r1 < n - 1; r1++)
      for (int r2 = r1 + 1; r2 < n; r2++) hit[(size_t)r2 * (r2 + 1) / 2 + r1] += std::min(cnt[r1], cnt[r2]);
    g = t;
  }
  std::vector<double> jac((size_t)n * n, 0.0);
  double best = -1.0;
  int bi = -1, bj = -1;
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) {
      if (i == j) continue;
      const long sh = i > j ? hit[(size_t)i * (i + 1) / 2 + j] : hit[(size_t)j * (j + 1) / 2 + i];
      const long tot = hit[(size_t)i * (i + 1) / 2 + i] + hit[(size_t)j * (j + 1) / 2 + j] - sh;
      double v = 0;
      if (tot > 0) {
        v = (0.0 + sh) / tot;
        v = v * ((0.51 * len[i] / maxl) + (0.49 * len[j] / maxl));
      }
      jac[(size_t)i * n + j] = v;
      if (v > best) {
        best = v;
        bi = i;
        bj = j;
      }
    }
  }
  ord[0] = bi;
  ord[1] = bj;
  for (int in = 2; in < n; in++) {
    double top = -1.0;
    int pick = n;
    for (int r2 = 0; r2 < n; r2++) {
      double v = 0.0;
      for (int t = 0; t < in; t++) {
        if (ord[t] == r2) {
          v = -1.0;
          break;
        }
        v += jac[(size_t)ord[t] * n + r2];
      }
      if (v > top) {
        top = v;
        pick = r2;
      }
    }
    ord[in] = pick;
  }
  return ord;
}

// alignment() (align.cpp:1741-1828). The serial (-t 1) and OpenMP forms give the same
// segments (q_span is always k); the serial form is followed.
static int alignment(const Params& P, Graph& G, Seeder* S, int rid, const char* qseq, int qlen,
                     std::vector<Res>& res, std::vector<CallOut>* calls) {
  std::vector<uint8_t> t(qlen);
  for (int j = 0; j < qlen; j++) t[j] = code_of((uint8_t)qseq[j]);
  res.clear();
  if (G.node.size() <= 2) {
    res.push_back({0, CODE_N, -1});
    for (int j = 0; j < qlen; j++) res.push_back({-1, t[j], -1});
    return ORC_OK;
  }
  CallOut co;
  if (!P.seeding) {
    int rc = poa(P, G, 0, 1, t.data(), qlen, P.ab_band, res, &co, nullptr);
    if (calls) calls->push_back(co);
    return rc;
  }
  if (!G.is_topsorted) G.topsort(P, 0);
  std::vector<mm128> anchors = S->anchors_bycons(P, rid, qlen, G.cons);
  if (anchors.empty()) {
    int rc = poa(P, G, 0, 1, t.data(), qlen, true, res, &co, nullptr);
    if (calls) calls->push_back(co);
    return rc;
  }
  int beg_id = 0, beg_qpos = 0, j = 0;
  for (const mm128& an : anchors) {
    int q_span = (int)(an.y >> 32 & 0xff);
    int end_tpos = (int)(uint32_t)an.x - q_span + 1, end_qpos = (int)(uint32_t)an.y - q_span + 1;
    int end_id = G.cons_pos_to_id[end_tpos];
    int sub_len = end_qpos - beg_qpos;
    int rc = poa(P, G, beg_id, end_id, t.data() + j, sub_len, true, res, &co, nullptr);
    if (calls) calls->push_back(co);
    if (rc != ORC_OK) return rc;
    j += sub_len;
    for (int k = 0; k < q_span; k++, j++) {
      end_id = G.cons_pos_to_id[end_tpos + k];
      if (k != q_span - 1) res.push_back({G.node[end_id].id, G.node[end_id].base, -1});
    }
    beg_id = end_id;
    beg_qpos = j;
  }
  int rc = poa(P, G, beg_id, 1, t.data() + j, qlen - beg_qpos, true, res, &co, nullptr);
  if (calls) calls->push_back(co);
  return rc;
}

struct Text {
  std::string s;
  void put(const char* x) { s += x; }
  void put(const std::string& x) { s += x; }
  void put(char c) { s.push_back(c); }
  void put_int(long v) { s += std::to_string(v); }
};

// output_rc_msa (graph.cpp:237-251)
static void output_msa(Graph& G, const Params& P, int n_reads, const std::vector<std::string>& name,
                       const std::vector<std::string>& comment, Text& o,
                       const std::vector<int>* rid_to_ord = nullptr) {
  if (!G.is_topsorted) G.topsort(P, 1);
  const size_t L = G.rank.size() >= 2 ? G.rank.size() - 2 : 0;
  std::vector<std::string> rows(n_reads, std::string(L, '-'));
  for (size_t i = 2; i < G.node.size(); i++) {
    int rk = G.node[G.node[i].par_id].rank;
    for (int id : G.node[i].ids) rows[id][rk - 1] = char_of(G.node[i].base);
  }
  for (int i = 0; i < n_reads; i++) {
    if (name.empty() || name[i].empty()) {
      o.put('>');
      o.put_int(i + 1);
      o.put('\n');
    } else {
      o.put('>');
      o.put(name[i]);
      o.put(' ');
      o.put(comment.empty() ? std::string() : comment[i]);
      o.put('\n');
    }
    // read i's row is that of the read added at position rid_to_ord[i] (graph.cpp:249)
    o.put(rows[rid_to_ord ? (*rid_to_ord)[i] : i]);
    o.put('\n');
  }
}

// The row-column MSA (-r 1) followed by a consensus row: the consensus of build_consensus
// (graph.cpp:294-357), each base at the column of its node's aligned group (the op 1 rank that
// places the node's base in the read rows, graph.cpp:240-245), '-' elsewhere. minipoa has no
// such output; its rows are the -r 1 rows and its ungapped row the -r 0 consensus.
static void output_msa_cons(Graph& G, const Params& P, int n_reads,
                            const std::vector<std::string>& name,
                            const std::vector<std::string>& comment, Text& o,
                            const std::vector<int>* rid_to_ord = nullptr) {
  output_msa(G, P, n_reads, name, comment, o, rid_to_ord);
  const size_t L = G.rank.size() >= 2 ? G.rank.size() - 2 : 0;
  std::string row(L, '-');
  G.build_consensus();
  for (size_t p = 0; p < G.cons_pos_to_id.size(); p++) {
    const int id = G.cons_pos_to_id[p];
    const int rk = G.node[G.node[id].par_id].rank;
    if (rk >= 1 && (size_t)rk <= L) row[rk - 1] = char_of(G.node[id].base);
  }
  o.put(">consensus sequence\n");
  o.put(row);
  o.put('\n');
}

// output_gfa (graph.cpp:358-433)
static void output_gfa(const Graph& G, int n_reads, const std::vector<int>& read_len,
                       const std::vector<std::string>& name, Text& o) {
  if (G.node.size() <= 2) return;
  std::vector<int> deg(G.node.size());
  for (size_t i = 0; i < G.node.size(); i++) deg[i] = (int)G.node[i].in.size();
  std::vector<std::vector<int>> paths(n_reads);
  for (int i = 0; i < n_reads; i++) paths[i].assign(read_len[i], 0);
  o.put("H\tVN:Z:1.0\n");
  std::deque<int> q;
  q.push_back(0);
  while (!q.empty()) {
    int cur = q.front();
    q.pop_front();
    if (cur == 1) break;
    const Node& v = G.node[cur];
    if (cur != 0) {
      o.put("S\t");
      o.put_int(cur - 1);
      o.put('\t');
      o.put(char_of(v.base));
      o.put('\n');
      for (int pre : v.in) {
        if (pre == 0) continue;
        o.put("L\t");
        o.put_int(pre - 1);
        o.put("\t+\t");
        o.put_int(cur - 1);
        o.put("\t+\t0M\n");
      }
      for (size_t j = 0; j < v.ids.size(); j++) paths[v.ids[j]][v.idp[j]] = cur - 1;
    }
    for (int t : v.out)
      if (--deg[t] == 0) q.push_back(t);
  }
  for (int i = 0; i < n_reads; i++) {
    o.put("P\t");
    if (!name.empty() && !name[i].empty()) o.put(name[i]);
    else o.put_int(i + 1);
    o.put('\t');
    for (int j = 0; j < read_len[i]; j++) {
      o.put_int(paths[i][j]);
      o.put('+');
      if (j != read_len[i] - 1) o.put(',');
      else o.put("\t*\n");
    }
  }
}

struct WindowRun {
  Graph G;
  std::vector<CallOut> calls;
  std::vector<int> call_read;
  std::vector<int> call_nodes, call_edges;
};

// main() (main.cpp:84-126) for one window: reads in input order, alignment -> add_path ->
// topsort (op=1 after the last read), then the selected output. With partial >= 0 only reads
// [0, partial) are added and every topsort is op=0, the state before a further read.
static int run_window(const Params& P, int n_reads, const char* seq, const int64_t* off,
                      const std::vector<std::string>& name,
                      const std::vector<std::string>& comment, Text* o, WindowRun& W,
                      int partial = -1) {
  Graph& G = W.G;
  G.init();
  Seeder S;
  if (P.seeding) S.init(P, n_reads, seq, off);
  // -p: reads in guide-tree order, read ord[i] added as the i-th (main.cpp:100-116)
  if (P.progressive && P.result == 2) return ORC_E_ARG;
  const std::vector<int> ord = P.progressive ? guide_tree(P, n_reads, seq, off) : [&] {
    std::vector<int> id(n_reads);
    for (int i = 0; i < n_reads; i++) id[i] = i;
    return id;
  }();
  std::vector<int> rid_to_ord(n_reads);
  for (int i = 0; i < n_reads; i++) rid_to_ord[ord[i]] = i;
  std::vector<Res> res;
  int last = partial >= 0 ? partial : n_reads;
  for (int i = 0; i < last; i++) {
    size_t before = W.calls.size();
    int nodes = (int)G.node.size(), edges = G.n_edges();
    const int rid = ord[i];
    int rc = alignment(P, G, &S, rid, seq + off[rid], (int)(off[rid + 1] - off[rid]), res,
                       &W.calls);
    for (size_t c = before; c < W.calls.size(); c++) {
      W.call_read.push_back(i);
      W.call_nodes.push_back(nodes);
      W.call_edges.push_back(edges);
    }
    if (rc != ORC_OK) return rc;
    G.add_path(i, res, 1);
    G.topsort(P, partial < 0 && i + 1 == n_reads ? 1 : 0);
  }
  if (!o) return ORC_OK;
  if (P.result == 0) {
    G.build_consensus();
    o->put(">consensus sequence\n");
    o->put(G.cons);
    o->put('\n');
  } else if (P.result == 1) {
    output_msa(G, P, n_reads, name, comment, *o, &rid_to_ord);
  } else if (P.result == 3) {
    output_msa_cons(G, P, n_reads, name, comment, *o, &rid_to_ord);
  } else if (P.result == 2) {
    std::vector<int> len(n_reads);
    for (int i = 0; i < n_reads; i++) len[i] = (int)(off[i + 1] - off[i]);
    output_gfa(G, n_reads, len, name, *o);
  }
  return ORC_OK;
}

static std::vector<std::string> split(const char* buf, const int64_t* off, int n) {
  std::vector<std::string> v;
  if (!buf || !off) return v;
  for (int i = 0; i < n; i++) v.emplace_back(buf + off[i], buf + off[i + 1]);
  return v;
}

static void fill_stats(const WindowRun& W, orc_call_stat* st, int64_t cap, int64_t base) {
  if (!st) return;
  for (size_t c = 0; c < W.calls.size(); c++) {
    int64_t at = base + (int64_t)c;
    if (at >= cap) break;
    orc_call_stat& s = st[at];
    const CallOut& co = W.calls[c];
    s.read = W.call_read[c];
    s.rows = co.rows;
    s.qlen = co.qlen;
    s.score = co.score;
    s.ab_band = co.ab_band;
    s.max_width = co.max_width;
    s.cells = co.cells;
    s.edge_cells = co.edge_cells;
    s.logical_cells = co.logical;
    s.graph_nodes = W.call_nodes[c];
    s.graph_edges = W.call_edges[c];
  }
}

}  // namespace orc

using namespace orc;

extern "C" int orc_window(const orc_params* p, int32_t n_reads, const char* seq,
                          const int64_t* seq_off, const char* name, const int64_t* name_off,
                          const char* comment, const int64_t* comment_off, char* out,
                          int64_t out_cap, int64_t* out_len, orc_call_stat* stats,
                          int32_t stats_cap, int32_t* n_stats) {
  if (!p || n_reads < 0 || (n_reads > 0 && (!seq || !seq_off))) return ORC_E_ARG;
  Params P = normalise(*p);
  std::vector<std::string> nm = split(name, name_off, n_reads);
  std::vector<std::string> cm = split(comment, comment_off, n_reads);
  Text o;
  WindowRun W;
  int rc = run_window(P, n_reads, seq, seq_off, nm, cm, &o, W);
  fill_stats(W, stats, stats_cap, 0);
  if (n_stats) *n_stats = (int32_t)W.calls.size();
  if (rc != ORC_OK) return rc;
  if (out_len) *out_len = (int64_t)o.s.size();
  if (!out || (int64_t)o.s.size() > out_cap) return out ? ORC_E_OUT_CAP : ORC_OK;
  std::memcpy(out, o.s.data(), o.s.size());
  return ORC_OK;
}

extern "C" int orc_batch(const orc_params* p, int32_t n_windows, const int32_t* win_off,
                         const char* seq, const int64_t* seq_off, int32_t threads, char* out,
                         int64_t out_cap, int64_t* out_off, orc_call_stat* stats,
                         int64_t stats_cap, int64_t* calls_off) {
  if (!p || n_windows < 0 || !win_off || !seq || !seq_off) return ORC_E_ARG;
  Params P = normalise(*p);
  std::vector<std::string> texts(n_windows);
  std::vector<WindowRun> runs(stats ? n_windows : 0);
  std::vector<int64_t> n_calls(n_windows, 0);
  std::vector<int> rcs(n_windows, ORC_OK);
  int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int w = 0; w < n_windows; w++) {
    WindowRun local;
    WindowRun& W = stats ? runs[w] : local;
    Text o;
    int r0 = win_off[w];
    int nr = win_off[w + 1] - r0;
    std::vector<int64_t> off(nr + 1);
    for (int i = 0; i <= nr; i++) off[i] = seq_off[r0 + i] - seq_off[r0];
    rcs[w] = run_window(P, nr, seq + seq_off[r0], off.data(), {}, {}, &o, W);
    texts[w] = std::move(o.s);
    n_calls[w] = (int64_t)W.calls.size();
  }
  int64_t total = 0;
  for (int w = 0; w < n_windows; w++) total += (int64_t)texts[w].size();
  if (out_off) {
    out_off[0] = 0;
    for (int w = 0; w < n_windows; w++) out_off[w + 1] = out_off[w] + (int64_t)texts[w].size();
  }
  if (calls_off) {
    calls_off[0] = 0;
    for (int w = 0; w < n_windows; w++) calls_off[w + 1] = calls_off[w] + n_calls[w];
  }
  if (stats) {
    int64_t base = 0;
    for (int w = 0; w < n_windows; w++) {
      fill_stats(runs[w], stats, stats_cap, base);
      base += n_calls[w];
    }
  }
  for (int w = 0; w < n_windows; w++)
    if (rcs[w] != ORC_OK) return rcs[w];
  if (out) {
    if (total > out_cap) return ORC_E_OUT_CAP;
    int64_t at = 0;
    for (int w = 0; w < n_windows; w++) {
      std::memcpy(out + at, texts[w].data(), texts[w].size());
      at += (int64_t)texts[w].size();
    }
  }
  return ORC_OK;
}

namespace orc {
int export_graph(const Graph& G, int32_t* n_nodes, int32_t* n_edges, uint8_t* base,
                 int32_t* in_off, int32_t* in_src, int32_t* in_w, int32_t* out_off,
                 int32_t* out_dst, int32_t* out_w, int32_t* rank_to_id, int32_t* hlen,
                 int32_t* tlen, int32_t* par_id, int32_t* aligned);
}  // namespace orc

extern "C" int orc_graph(const orc_params* p, int32_t n_reads, const char* seq,
                         const int64_t* seq_off, int32_t* n_nodes, int32_t* n_edges,
                         uint8_t* base, int32_t* in_off, int32_t* in_src, int32_t* in_w,
                         int32_t* out_off, int32_t* out_dst, int32_t* out_w,
                         int32_t* rank_to_id, int32_t* hlen, int32_t* tlen, int32_t* par_id,
                         int32_t* aligned) {
  if (!p || n_reads < 0) return ORC_E_ARG;
  Params P = normalise(*p);
  WindowRun W;
  int rc = run_window(P, n_reads, seq, seq_off, {}, {}, nullptr, W, n_reads);
  if (rc != ORC_OK) return rc;
  return orc::export_graph(W.G, n_nodes, n_edges, base, in_off, in_src, in_w, out_off, out_dst,
                           out_w, rank_to_id, hlen, tlen, par_id, aligned);
}

namespace orc {
int export_graph(const Graph& G, int32_t* n_nodes, int32_t* n_edges, uint8_t* base,
                 int32_t* in_off, int32_t* in_src, int32_t* in_w, int32_t* out_off,
                 int32_t* out_dst, int32_t* out_w, int32_t* rank_to_id, int32_t* hlen,
                 int32_t* tlen, int32_t* par_id, int32_t* aligned) {
  const int N = (int)G.node.size();
  if (n_nodes) *n_nodes = N;
  if (n_edges) *n_edges = G.n_edges();
  if (!base) return ORC_OK;
  int ai = 0, ao = 0;
  for (int v = 0; v < N; v++) {
    const Node& x = G.node[v];
    base[v] = x.base;
    if (par_id) par_id[v] = x.par_id;
    if (aligned)
      for (int c = 0; c < NB; c++) aligned[v * NB + c] = x.aligned[c];
    if (in_off) in_off[v] = ai;
    for (size_t k = 0; k < x.in.size(); k++, ai++) {
      if (in_src) in_src[ai] = x.in[k];
      if (in_w) in_w[ai] = x.in_w[k];
    }
    if (out_off) out_off[v] = ao;
    for (size_t k = 0; k < x.out.size(); k++, ao++) {
      if (out_dst) out_dst[ao] = x.out[k];
      if (out_w) out_w[ao] = x.out_w[k];
    }
  }
  if (in_off) in_off[N] = ai;
  if (out_off) out_off[N] = ao;
  for (size_t r = 0; r < G.rank.size(); r++) {
    if (rank_to_id) rank_to_id[r] = G.rank[r];
    if (hlen) hlen[r] = r < G.hlen.size() ? G.hlen[r] : 0;
    if (tlen) tlen[r] = r < G.tlen.size() ? G.tlen[r] : 0;
  }
  return ORC_OK;
}
}  // namespace orc

extern "C" int orc_dp_matrix(const orc_params* p, int32_t n_reads, const char* seq,
                             const int64_t* seq_off, int32_t* n_rows, int64_t* total,
                             int32_t* Bs, int32_t* Be, int64_t* row_off, int32_t* H, int32_t* D,
                             int32_t* I, int32_t* score) {
  if (!p || n_reads < 2 || p->seeding) return ORC_E_ARG;
  Params P = normalise(*p);
  WindowRun W;
  int rc = run_window(P, n_reads - 1, seq, seq_off, {}, {}, nullptr, W, n_reads - 1);
  if (rc != ORC_OK) return rc;
  // The last read against the graph of the first n_reads-1 reads; every earlier topsort was
  // op=0 because the loop stopped before the last read.
  Graph& G = W.G;
  const int qlen = (int)(seq_off[n_reads] - seq_off[n_reads - 1]);
  std::vector<uint8_t> t(qlen);
  for (int j = 0; j < qlen; j++) t[j] = code_of((uint8_t)seq[seq_off[n_reads - 1] + j]);
  std::vector<Res> res;
  DpDump dump;
  CallOut co;
  if (G.node.size() <= 2) return ORC_E_ARG;
  rc = poa(P, G, 0, 1, t.data(), qlen, P.ab_band, res, &co, &dump);
  if (rc != ORC_OK && rc != ORC_E_BACKTRACK && rc != ORC_E_BAND) return rc;
  const int n = (int)dump.Bs.size();
  if (n_rows) *n_rows = n;
  if (total) *total = dump.off[n];
  if (score) *score = co.score;
  if (Bs) std::copy(dump.Bs.begin(), dump.Bs.end(), Bs);
  if (Be) std::copy(dump.Be.begin(), dump.Be.end(), Be);
  if (row_off) std::copy(dump.off.begin(), dump.off.end(), row_off);
  if (H) std::copy(dump.H.begin(), dump.H.end(), H);
  if (D) std::copy(dump.D.begin(), dump.D.end(), D);
  if (I) std::copy(dump.I.begin(), dump.I.end(), I);
  return rc;
}

// The global chain's anchors of read `rid` against a given consensus (codes or ASCII), as the
// -S dispatcher consumes them: out[2*i] = x, out[2*i+1] = y.
extern "C" int orc_anchors(const orc_params* p, int32_t n_reads, const char* seq,
                           const int64_t* seq_off, int32_t rid, const char* cons,
                           int32_t cons_len, uint64_t* out, int64_t cap, int64_t* n) {
  if (!p || rid < 0 || rid >= n_reads) return ORC_E_ARG;
  Params P = normalise(*p);
  Seeder S;
  S.init(P, n_reads, seq, seq_off);
  std::string c(cons, cons + cons_len);
  const int qlen = (int)(seq_off[rid + 1] - seq_off[rid]);
  std::vector<mm128> a = S.anchors_bycons(P, rid, qlen, c);
  if (n) *n = (int64_t)a.size();
  for (size_t i = 0; i < a.size() && (int64_t)i < cap; i++) {
    out[2 * i] = a[i].x;
    out[2 * i + 1] = a[i].y;
  }
  return ORC_OK;
}

// ---------------------------------------------------------------------------------------
// Call-level mirror of minipoa's functions on one graph (graph::init, poa(), alignment(),
// graph::add_path, graph::topsort, the outputs), the checker of the library's call-level API.
// ---------------------------------------------------------------------------------------
struct orc_dag {
  orc::Params P;
  orc::Graph G;
  orc::Seeder S;
  bool seeded = false;
  std::vector<int> read_len;  // reads added so far (add_path), in order
  std::vector<int> ord, rid_to_ord;  // the read set's order (-p: the guide tree), and inverse
  std::vector<int> scores;  // the score of every DP call completed on this graph (-V 1)
};

static void put_res(const std::vector<orc::Res>& res, int32_t* out, int64_t cap, int64_t* n) {
  if (n) *n = (int64_t)res.size();
  if (!out) return;
  for (int64_t i = 0; i < (int64_t)res.size() && i < cap; i++) {
    out[3 * i] = res[i].from;
    out[3 * i + 1] = res[i].base;
    out[3 * i + 2] = res[i].aligned_id;
  }
}

extern "C" orc_dag* orc_dag_new(const orc_params* p, int32_t n_reads, const char* seq,
                                    const int64_t* seq_off) {
  if (!p) return nullptr;
  orc_dag* g = new orc_dag();
  g->P = orc::normalise(*p);
  g->G.init();
  if (g->P.seeding && n_reads > 0 && seq && seq_off) {
    g->S.init(g->P, n_reads, seq, seq_off);
    g->seeded = true;
  }
  if (n_reads > 0 && seq && seq_off) {
    if (g->P.progressive) {
      g->ord = orc::guide_tree(g->P, n_reads, seq, seq_off);
    } else {
      g->ord.resize(n_reads);
      for (int i = 0; i < n_reads; i++) g->ord[i] = i;
    }
    g->rid_to_ord.resize(n_reads);
    for (int i = 0; i < n_reads; i++) g->rid_to_ord[g->ord[i]] = i;
  }
  return g;
}

extern "C" int orc_dag_order(orc_dag* g, int32_t* ord, int32_t cap) {
  if (!g) return ORC_E_ARG;
  for (int i = 0; i < (int)g->ord.size() && i < cap; i++) ord[i] = g->ord[i];
  return (int)g->ord.size();
}

extern "C" void orc_dag_free(orc_dag* g) { delete g; }

extern "C" int orc_dag_poa(orc_dag* g, int32_t beg_id, int32_t end_id, const char* q,
                             int32_t qlen, int32_t ab_band, int32_t* res, int64_t cap,
                             int64_t* n_res, int32_t* score) {
  if (!g || beg_id < 0 || end_id < 0 || beg_id >= (int)g->G.node.size() ||
      end_id >= (int)g->G.node.size())
    return ORC_E_ARG;
  if (qlen < 0) return ORC_E_EMPTY_WINDOW;
  std::vector<uint8_t> t(qlen);
  for (int j = 0; j < qlen; j++) t[j] = orc::code_of((uint8_t)q[j]);
  std::vector<orc::Res> out;
  orc::CallOut co;
  const int rc = orc::poa(g->P, g->G, beg_id, end_id, t.data(), qlen, ab_band != 0, out, &co,
                          nullptr);
  if (score) *score = co.score;
  if (rc == ORC_OK) g->scores.push_back(co.score);
  put_res(out, res, cap, n_res);
  return rc;
}

extern "C" int orc_dag_alignment(orc_dag* g, int32_t rid, const char* q, int32_t qlen,
                                   int32_t* res, int64_t cap, int64_t* n_res) {
  if (!g || qlen < 0 || (g->P.seeding && !g->seeded)) return ORC_E_ARG;
  std::vector<orc::Res> out;
  std::vector<orc::CallOut> calls;
  const int rc = orc::alignment(g->P, g->G, &g->S, rid, q, qlen, out, &calls);
  // the last call is the one that failed, if one did
  for (size_t c = 0; c + (rc != ORC_OK ? 1 : 0) < calls.size(); c++)
    g->scores.push_back(calls[c].score);
  put_res(out, res, cap, n_res);
  return rc;
}

extern "C" int orc_dag_scores(orc_dag* g, int32_t* out, int32_t cap) {
  if (!g) return ORC_E_ARG;
  for (int i = 0; i < (int)g->scores.size() && i < cap; i++) out[i] = g->scores[i];
  return (int)g->scores.size();
}

extern "C" int orc_dag_add_path(orc_dag* g, int32_t seq_id, const int32_t* res,
                                  int64_t n_res) {
  if (!g || !res || n_res < 1) return ORC_E_ARG;
  std::vector<orc::Res> r((size_t)n_res);
  for (int64_t i = 0; i < n_res; i++)
    r[i] = {res[3 * i], (uint8_t)res[3 * i + 1], res[3 * i + 2]};
  g->G.add_path(seq_id, r, 1);
  if (seq_id == (int)g->read_len.size()) g->read_len.push_back((int)n_res - 1);
  return ORC_OK;
}

// read_gfa (handle_input.cpp:444-495): S lines add their bases as nodes one by one
// (gfa_parse_S, :78-127, graph::add_node at graph.cpp:7-12), P lines add their '+' paths edge
// by edge with the read id of the path (gfa_parse_P, :128-261), then the op 0 sort (:482). The
// inputs the library refuses ('-' segments, unknown or duplicated segments, segments without
// bases, paths without segments, GFA output) are ORC_E_ARG here too.
extern "C" int orc_dag_read_gfa(orc_dag* g, const char* text, int64_t len, char* names,
                                  int64_t names_cap, int64_t* names_len, int32_t* n_paths) {
  if (!g || len < 0 || g->G.node.size() != 2 || !g->read_len.empty() || g->P.result == 2)
    return ORC_E_ARG;
  std::map<std::string, std::pair<int, int>> seg;  // name -> (first node, last node)
  std::string all;
  int paths = 0;
  int64_t at = 0;
  while (at < len) {
    int64_t e = at;
    while (e < len && text[e] != '\n') e++;
    std::string line(text + at, (size_t)(e - at));
    at = e + 1;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.size() < 3 || line[1] != '\t' || (line[0] != 'S' && line[0] != 'P')) continue;
    const size_t t1 = line.find('\t', 2);
    if (t1 == std::string::npos) return ORC_E_ARG;  // no second field
    const std::string name = line.substr(2, t1 - 2);
    const size_t t2 = line.find('\t', t1 + 1);
    const std::string body = line.substr(t1 + 1, t2 == std::string::npos ? std::string::npos
                                                                         : t2 - t1 - 1);
    if (line[0] == 'S') {
      if (body.empty() || seg.count(name)) return ORC_E_ARG;
      const int first = (int)g->G.node.size();
      for (char c : body) {
        const int id = (int)g->G.node.size();
        g->G.node.emplace_back(id, orc::code_of((uint8_t)c));
      }
      seg[name] = {first, (int)g->G.node.size() - 1};
    } else if (line[0] == 'P') {
      const int sid = (int)g->read_len.size();
      int last = 0, pos = 0, visited = 0;
      size_t from = 0;
      for (size_t k = 0; k < body.size(); k++) {
        if (body[k] == '-') return ORC_E_ARG;
        if (body[k] != '+') continue;
        auto it = seg.find(body.substr(from, k - from));
        if (it == seg.end()) return ORC_E_ARG;
        const int in = it->second.first, out = it->second.second;
        g->G.add_adj(sid, last, in, pos++);
        for (int v = in; v < out; v++) g->G.add_adj(sid, v, v + 1, pos++);
        visited += out - in + 1;
        last = out;
        from = k + 2;
      }
      if (visited == 0) return ORC_E_ARG;
      g->G.add_adj(sid, last, 1, pos);
      g->G.is_topsorted = false;
      g->read_len.push_back(visited);
      all += name + "\n";
      paths++;
    }
  }
  g->G.topsort(g->P, 0);
  if (n_paths) *n_paths = paths;
  if (names_len) *names_len = (int64_t)all.size();
  if (names) {
    if ((int64_t)all.size() > names_cap) return ORC_E_OUT_CAP;
    std::memcpy(names, all.data(), all.size());
  }
  return ORC_OK;
}

extern "C" int orc_dag_topsort(orc_dag* g, int32_t op) {
  if (!g || (op != 0 && op != 1)) return ORC_E_ARG;
  g->G.topsort(g->P, op);
  return ORC_OK;
}

extern "C" int orc_dag_output(orc_dag* g, int32_t result, const char* name,
                                const int64_t* name_off, const char* comment,
                                const int64_t* comment_off, char* out, int64_t out_cap,
                                int64_t* out_len) {
  if (!g || result < 0 || result > 3) return ORC_E_ARG;
  const int n = (int)g->read_len.size();
  const std::vector<std::string> nm = orc::split(name, name_off, n);
  const std::vector<std::string> cm = orc::split(comment, comment_off, n);
  orc::Text o;
  if (result == 0) {
    g->G.build_consensus();
    o.put(">consensus sequence\n");
    o.put(g->G.cons);
    o.put('\n');
  } else if (result == 1 || result == 3) {
    // rows through the read set's order when every read of it has been added
    const std::vector<int>* map = (int)g->rid_to_ord.size() == n ? &g->rid_to_ord : nullptr;
    if (result == 1) orc::output_msa(g->G, g->P, n, nm, cm, o, map);
    else orc::output_msa_cons(g->G, g->P, n, nm, cm, o, map);
  } else {
    orc::output_gfa(g->G, n, g->read_len, nm, o);
  }
  if (out_len) *out_len = (int64_t)o.s.size();
  if (!out || (int64_t)o.s.size() > out_cap) return out ? ORC_E_OUT_CAP : ORC_OK;
  std::memcpy(out, o.s.data(), o.s.size());
  return ORC_OK;
}

extern "C" int orc_dag_export(orc_dag* g, int32_t* n_nodes, int32_t* n_edges, uint8_t* base,
                                int32_t* in_off, int32_t* in_src, int32_t* in_w,
                                int32_t* out_off, int32_t* out_dst, int32_t* out_w,
                                int32_t* rank_to_id, int32_t* hlen, int32_t* tlen,
                                int32_t* par_id, int32_t* aligned) {
  if (!g) return ORC_E_ARG;
  return orc::export_graph(g->G, n_nodes, n_edges, base, in_off, in_src, in_w, out_off, out_dst,
                           out_w, rank_to_id, hlen, tlen, par_id, aligned);
}
