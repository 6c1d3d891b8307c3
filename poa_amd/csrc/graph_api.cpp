// Call-level C ABI (include/mpoa.h, "Call-level API"): minipoa's poa(), alignment(),
// graph::add_path, graph::topsort and the outputs over one graph that stays in device memory
// between calls. Each call is one launch of k_graph_op (pipeline.hip) on the graph's arena
// (slot 0 of its own pools); the host keeps the sizes, grows the arena before an add_path could
// outgrow it, and re-runs a poa() whose DP storage overflowed with exact capacities.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "host.h"
#include "mpoa.h"
#include "pipeline.h"
#include "seed.h"

using mpoa::Arena;
using mpoa::DpScores;
using mpoa::GraphOpArgs;
using mpoa::GraphOpOut;
using mpoa::KernelArgs;
using mpoa::SlotCaps;
using mpoa::WinDesc;
using mpoa::WinResult;
using mpoa::WinState;
using namespace mpoa::host;

struct mpoa_seeder {
  mpoa::seed::Seeder sd;
  int n_reads = 0;
  std::vector<int> ord;  // the order reads are added in (-p: the guide tree)
};

struct mpoa_graph {
  mpoa_ctx* ctx = nullptr;
  mpoa_params params{};
  DpScores sc{};
  SlotCaps caps{};
  DevBuf<int> d_int, d_dp;
  DevBuf<uint8_t> d_byte, d_q, d_text;
  DevBuf<WinState> d_state;
  DevBuf<GraphOpOut> d_out;
  DevBuf<WinResult> d_res;
  DevBuf<WinDesc> d_win;
  DevBuf<int64_t> d_read_off, d_gfa;
  DevBuf<int> d_cons;
  WinState st{};               // host copy after the last operation
  std::vector<int> read_len;   // reads added so far, in order
  int64_t path0 = 0;           // their total length
  int sorted = -1;             // op of the last sort since the last add_path, -1: none
  std::vector<int> cons;       // -S: (node id, rank, base) per consensus position
};

namespace {

// Arena capacities: nodes, in/out edge records, longest query, read-path positions, DP cells.
SlotCaps caps_of(int nodes, int edges, int qmax, int path, int64_t dp) {
  SlotCaps c{};
  c.nodes = (std::max(nodes, 64) + 63) / 64 * 64;
  c.edges = (std::max(edges, 64) + 63) / 64 * 64;
  c.qmax = std::max(qmax, 1);
  c.path = path;
  c.dp = (std::max<int64_t>(dp, 64) + 63) / 64 * 64;
  mpoa::finish_caps(c);
  return c;
}

KernelArgs args_of(mpoa_graph* g) {
  KernelArgs a{};
  a.scores = g->sc;
  a.caps = g->caps;
  a.read_off = g->d_read_off.p;
  a.win = g->d_win.p;
  a.result = g->d_res.p;
  a.n_jobs = 1;
  a.int_pool = g->d_int.p;
  a.byte_pool = g->d_byte.p;
  a.dp_pool = g->d_dp.p;
  a.out = g->d_text.p;
  a.gfa_lines = g->d_gfa.p;
  a.forms = g->ctx->forms;
  return a;
}

Arena arena_of(mpoa_graph* g) {
  return mpoa::carve_pools(g->caps, g->d_int.p, g->d_byte.p, g->d_dp.p, 0);
}

// New pools of capacities `nc` (each at least the current one); the graph, its last order and
// plan, the read paths and the consensus are copied over. The per-call scratch (DP matrices,
// band state, res arrays, query) is not.
int regrow(mpoa_graph* g, SlotCaps nc) {
  nc = caps_of(std::max(nc.nodes, g->caps.nodes), std::max(nc.edges, g->caps.edges),
               std::max(nc.qmax, g->caps.qmax), std::max(nc.path, g->caps.path),
               std::max(nc.dp, g->caps.dp));
  const SlotCaps oc = g->caps;
  DevBuf<int> ni, nd;
  DevBuf<uint8_t> nb;
  int rc;
  if ((rc = ni.alloc((size_t)nc.ints_per_slot)) != MPOA_OK ||
      (rc = nb.alloc((size_t)nc.bytes_per_slot)) != MPOA_OK ||
      (rc = nd.alloc((size_t)3 * nc.dp)) != MPOA_OK)
    return rc;
  if (g->d_int.p) {
    const Arena o = arena_of(g);
    const Arena n = mpoa::carve_pools(nc, ni.p, nb.p, nd.p, 0);
    const hipStream_t s = g->ctx->stream;
    const size_t N = (size_t)oc.nodes, E = (size_t)oc.edges;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
    };
    const std::pair<int*, int*> by_node[] = {
        {n.par, o.par}, {n.ind, o.ind}, {n.rank, o.rank}, {n.in_head, o.in_head},
        {n.in_tail, o.in_tail}, {n.out_head, o.out_head}, {n.out_tail, o.out_tail},
        {n.in_cnt, o.in_cnt}, {n.last_read, o.last_read}, {n.out_cnt, o.out_cnt},
        {n.rank_id, o.rank_id}, {n.hlen, o.hlen}, {n.tlen, o.tlen}, {n.cons_id, o.cons_id},
        {n.cons_rank, o.cons_rank}};
    const std::pair<int*, int*> by_edge[] = {
        {n.ie_src, o.ie_src}, {n.ie_w, o.ie_w}, {n.ie_next, o.ie_next}, {n.oe_dst, o.oe_dst},
        {n.oe_w, o.oe_w}, {n.oe_next, o.oe_next}, {n.pred_rank, o.pred_rank},
        {n.pred_w, o.pred_w}};
    HIPCHK(cp(n.rowrec, o.rowrec, 12 * N * sizeof(int)));
    HIPCHK(cp(n.aligned, o.aligned, mpoa::NB * N * sizeof(int)));
    HIPCHK(cp(n.pred_off, o.pred_off, (N + 1) * sizeof(int)));
    for (const auto& d : by_node) HIPCHK(cp(d.first, d.second, N * sizeof(int)));
    for (const auto& d : by_edge) HIPCHK(cp(d.first, d.second, E * sizeof(int)));
    if (oc.path > 0) HIPCHK(cp(n.path, o.path, (size_t)oc.path * sizeof(int)));
    HIPCHK(cp(n.base, o.base, N));
    HIPCHK(cp(n.cons_ch, o.cons_ch, N));
    HIPCHK(cp(n.pred_base, o.pred_base, E));
    HIPCHK(hipStreamSynchronize(s));
  }
  std::swap(g->d_int.p, ni.p);
  std::swap(g->d_int.n, ni.n);
  std::swap(g->d_byte.p, nb.p);
  std::swap(g->d_byte.n, nb.n);
  std::swap(g->d_dp.p, nd.p);
  std::swap(g->d_dp.n, nd.n);
  g->caps = nc;
  if (!g->d_cons.p || g->d_cons.n < 3 * (size_t)nc.nodes)
    if ((rc = g->d_cons.alloc(3 * (size_t)nc.nodes)) != MPOA_OK) return rc;
  return MPOA_OK;
}

// One graph operation, synchronously; the host copy of the state and the operation's record
// are refreshed.
int run_op(mpoa_graph* g, GraphOpArgs& op, GraphOpOut* out) {
  op.a = args_of(g);
  if (op.op == mpoa::GOP_OUTPUT) op.a.scores.result = op.result;  // the output asked for
  op.state = g->d_state.p;
  op.out = g->d_out.p;
  op.cons_out = g->d_cons.p;
  op.cons_cap = g->caps.nodes;
  if (mpoa_launch_graph_op(op, g->ctx->stream) != 0) {
    g_err = "graph operation launch failed";
    return MPOA_E_DEVICE;
  }
  HIPCHK(hipStreamSynchronize(g->ctx->stream));
  HIPCHK(hipMemcpy(out, g->d_out.p, sizeof(GraphOpOut), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&g->st, g->d_state.p, sizeof(WinState), hipMemcpyDeviceToHost));
  return MPOA_OK;
}

int check_status(int status) {
  if (status == MPOA_E_OVERFLOW) {
    g_err = "graph arena overflow";
    return MPOA_E_DEVICE;  // the host sizes every arena first: this is a library fault
  }
  return status;
}

int do_topsort(mpoa_graph* g, int op_code) {
  GraphOpArgs op{};
  op.op = mpoa::GOP_TOPSORT;
  op.sort_op = op_code;
  op.seeding = g->params.seeding != 0 && op_code == 0;
  GraphOpOut out{};
  int rc = run_op(g, op, &out);
  if (rc != MPOA_OK) return rc;
  if ((rc = check_status(out.status)) != MPOA_OK) return rc;
  g->sorted = op_code;
  if (op.seeding) {
    const int L = g->st.cons_len;
    g->cons.resize(3 * (size_t)std::max(L, 1));
    if (L > 0)
      HIPCHK(hipMemcpy(g->cons.data(), g->d_cons.p, 3 * (size_t)L * sizeof(int),
                       hipMemcpyDeviceToHost));
  }
  return MPOA_OK;
}

// poa() on [rank(beg_id), rank(end_id)]: appends its res entries to `res`.
int do_poa(mpoa_graph* g, int beg_id, int end_id, const uint8_t* q, int qlen, int ab_band,
           std::vector<mpoa_res>* res, int* score, bool sub_window = false) {
  if (beg_id < 0 || end_id < 0 || beg_id >= g->st.n_nodes || end_id >= g->st.n_nodes) {
    g_err = "node id out of range";
    return MPOA_E_ARG;
  }
  if (qlen < (sub_window ? -1 : 0)) {
    // a -S sub-window that ends more than one base before it begins in the read (overlapping
    // anchors), where minipoa's DP reads past its query; length -1 runs (no res entry)
    g_err = "query sub-window of negative length";
    return MPOA_E_EMPTY_WINDOW;
  }
  if (g->sorted != 0) {
    g_err = "the graph must be sorted (topsort op 0) after its last add_path";
    return MPOA_E_ARG;
  }
  int rc;
  const int m = qlen + 1;
  const int64_t full = (int64_t)(m / 8 + 2) * 8;
  if (qlen > g->caps.qmax) {
    SlotCaps nc = g->caps;
    nc.qmax = std::max(qlen, g->caps.qmax + g->caps.qmax / 2);
    if ((rc = regrow(g, nc)) != MPOA_OK) return rc;
  }
  if (!g->d_q.p || g->d_q.n < (size_t)qlen + 1)
    if ((rc = g->d_q.alloc(std::max<size_t>((size_t)qlen + 1, 4096))) != MPOA_OK) return rc;
  if (qlen > 0) HIPCHK(hipMemcpy(g->d_q.p, q, (size_t)qlen, hipMemcpyHostToDevice));
  // DP storage: the static band's estimate first, the exact need after an overflow
  const bool adaptive = ab_band != 0 && g->sc.f > 0;
  int64_t want = (int64_t)g->st.n_nodes *
                 (adaptive || g->sc.f <= 0 ? full
                                           : std::min<int64_t>(full, 2LL * (g->sc.b + m / g->sc.f) + 64));
  for (int attempt = 0; attempt < 2; attempt++) {
    if (want > g->caps.dp) {
      SlotCaps nc = g->caps;
      nc.dp = std::max<int64_t>(want, g->caps.dp + g->caps.dp / 2);
      if ((rc = regrow(g, nc)) != MPOA_OK) return rc;
    }
    GraphOpArgs op{};
    op.op = mpoa::GOP_POA;
    op.beg_id = beg_id;
    op.end_id = end_id;
    op.adaptive = adaptive;
    op.q = g->d_q.p;
    op.qlen = qlen;
    GraphOpOut out{};
    if ((rc = run_op(g, op, &out)) != MPOA_OK) return rc;
    if (out.status == MPOA_E_OVERFLOW && attempt == 0) {
      want = (int64_t)std::max(out.rows, 1) * full + 64;
      continue;
    }
    if ((rc = check_status(out.status)) != MPOA_OK) return rc;
    if (score) *score = out.score;
    const Arena A = arena_of(g);
    std::vector<int> from((size_t)m), al((size_t)m);
    std::vector<uint8_t> base((size_t)m);
    HIPCHK(hipMemcpy(from.data(), A.res_from, (size_t)m * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(al.data(), A.res_aligned, (size_t)m * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(base.data(), A.res_base, (size_t)m, hipMemcpyDeviceToHost));
    for (int j = 0; j < m; j++) res->push_back(mpoa_res{from[j], base[j], al[j]});
    return MPOA_OK;
  }
  g_err = "DP storage overflow after resizing";
  return MPOA_E_DEVICE;
}

int put_res(const std::vector<mpoa_res>& r, mpoa_res* out, int64_t cap, int64_t* n) {
  if (n) *n = (int64_t)r.size();
  if (!out) return MPOA_OK;
  if ((int64_t)r.size() > cap) return MPOA_E_OUT_CAP;
  std::copy(r.begin(), r.end(), out);
  return MPOA_OK;
}

// Device output buffer bound for the reads added so far (the batch form's bounds).
int64_t text_cap(int result, const std::vector<int>& len) {
  int64_t qtot = 0, qmax = 0;
  for (int l : len) {
    qtot += l;
    qmax = std::max<int64_t>(qmax, l);
  }
  const int64_t n = (int64_t)len.size();
  if (result == 1 || result == 3) return (n + 1) * (2 + qtot) + 64;
  if (result == 2) {
    int d = 1;
    for (int64_t v = 2 + qtot; v >= 10; v /= 10) d++;
    return (2 + qtot) * (6 + d) + (qtot + n) * (13 + 2 * d) + qtot * (d + 2) + n * (16 + d) + 64;
  }
  return 2 * std::max<int64_t>(qmax, (int64_t)1 << 16) + 64;
}

}  // namespace

extern "C" {

int mpoa_graph_create(mpoa_ctx* ctx, const mpoa_params* params, mpoa_graph** out) {
  if (!ctx || !params || !out) return MPOA_E_ARG;
  *out = nullptr;
  if (params->result < 0 || params->result > 3) {
    g_err = "result must be 0 (consensus), 1 (MSA), 2 (GFA) or 3 (MSA with a consensus row)";
    return MPOA_E_ARG;
  }
  HIPCHK(hipSetDevice(ctx->device));
  mpoa_graph* g = new mpoa_graph();
  g->ctx = ctx;
  g->params = *params;
  g->sc = scores_of(*params);
  const bool keep_path = params->result != 0;
  int rc;
  if ((rc = g->d_state.alloc(1)) != MPOA_OK || (rc = g->d_out.alloc(1)) != MPOA_OK ||
      (rc = g->d_res.alloc(1)) != MPOA_OK || (rc = g->d_win.alloc(1)) != MPOA_OK ||
      (rc = regrow(g, caps_of(1024, 4096, 1024, keep_path ? 4096 : 0, 1 << 20))) != MPOA_OK) {
    delete g;
    return rc;
  }
  GraphOpArgs op{};
  op.op = mpoa::GOP_INIT;
  GraphOpOut o{};
  if ((rc = run_op(g, op, &o)) != MPOA_OK || (rc = check_status(o.status)) != MPOA_OK) {
    delete g;
    return rc;
  }
  *out = g;
  return MPOA_OK;
}

void mpoa_graph_destroy(mpoa_graph* g) { delete g; }

int mpoa_poa(mpoa_graph* g, int32_t beg_id, int32_t end_id, const char* qseq, int32_t qlen,
             int32_t ab_band, mpoa_res* res, int64_t cap, int64_t* n_res, int32_t* score) {
  if (!g || (qlen > 0 && !qseq)) return MPOA_E_ARG;
  HIPCHK(hipSetDevice(g->ctx->device));
  std::vector<mpoa_res> r;
  const int rc = do_poa(g, beg_id, end_id, reinterpret_cast<const uint8_t*>(qseq), qlen, ab_band,
                        &r, score);
  if (rc != MPOA_OK) return rc;
  return put_res(r, res, cap, n_res);
}

int mpoa_seeder_create(const mpoa_params* params, int32_t n_reads, const char* seq,
                       const int64_t* seq_off, mpoa_seeder** out) {
  if (!params || !out || n_reads < 0 || (n_reads > 0 && (!seq || !seq_off))) return MPOA_E_ARG;
  mpoa_seeder* s = new mpoa_seeder();
  const int64_t b0 = n_reads > 0 ? seq_off[0] : 0;
  std::vector<int64_t> off((size_t)n_reads + 1, 0);
  std::vector<uint8_t> codes((size_t)(n_reads > 0 ? seq_off[n_reads] - b0 : 0) + 1);
  for (int i = 0; i <= n_reads && n_reads > 0; i++) off[i] = seq_off[i] - b0;
  for (int64_t j = 0; n_reads > 0 && j < off[n_reads]; j++)
    codes[j] = code_of_char((unsigned char)seq[b0 + j]);
  mpoa::seed::Config cfg;
  cfg.k = params->k;
  cfg.w = params->mm_w;
  cfg.bw = 1000;  // the chaining band minipoa's main sets (src/main.cpp:65)
  cfg.poa_w = params->poa_w;
  s->sd.init(cfg, n_reads, codes.data(), off.data());
  s->n_reads = n_reads;
  if (params->progressive) {
    s->ord = mpoa::seed::guide_order(cfg, n_reads, codes.data(), off.data());
  } else {
    s->ord.resize((size_t)n_reads);
    for (int i = 0; i < n_reads; i++) s->ord[i] = i;
  }
  *out = s;
  return MPOA_OK;
}

void mpoa_seeder_destroy(mpoa_seeder* s) { delete s; }

int32_t mpoa_seeder_order(mpoa_seeder* s, int32_t* ord, int32_t cap) {
  if (!s) return MPOA_E_ARG;
  for (int i = 0; i < s->n_reads && i < cap && ord; i++) ord[i] = s->ord[i];
  return s->n_reads;
}

int mpoa_alignment(mpoa_graph* g, mpoa_seeder* mm, int32_t rid, const char* qseq, int32_t qlen,
                   mpoa_res* res, int64_t cap, int64_t* n_res) {
  if (!g || qlen < 0 || (qlen > 0 && !qseq)) return MPOA_E_ARG;
  HIPCHK(hipSetDevice(g->ctx->device));
  const uint8_t* q = reinterpret_cast<const uint8_t*>(qseq);
  std::vector<mpoa_res> r;
  int rc;
  // a graph without bases: the source, then one new node per base (align.cpp:1747-1753)
  if (g->st.n_nodes <= 2) {
    r.push_back(mpoa_res{0, 4, -1});
    for (int j = 0; j < qlen; j++) r.push_back(mpoa_res{-1, code_of_char(q[j]), -1});
    return put_res(r, res, cap, n_res);
  }
  if (!g->params.seeding) {
    if ((rc = do_poa(g, 0, 1, q, qlen, g->params.ab_band, &r, nullptr)) != MPOA_OK) return rc;
    return put_res(r, res, cap, n_res);
  }
  if (!mm || rid < 0 || rid >= mm->n_reads) {
    g_err = "seeding needs the read set's seeder and a read index in it";
    return MPOA_E_ARG;
  }
  if (g->sorted < 0 && (rc = do_topsort(g, 0)) != MPOA_OK) return rc;
  const int L = (int)g->cons.size() / 3 >= g->st.cons_len ? g->st.cons_len : 0;
  std::vector<uint8_t> cons_code((size_t)std::max(L, 1));
  for (int p = 0; p < L; p++) cons_code[p] = (uint8_t)g->cons[3 * p + 2];
  std::vector<uint8_t> t((size_t)qlen + 1);
  for (int j = 0; j < qlen; j++) t[j] = code_of_char(q[j]);
  const std::vector<mpoa::seed::Pair> an = mm->sd.anchors(rid, qlen, cons_code.data(), L);
  // -S always takes the adaptive band (align.cpp:1780, 1808, 1826)
  if (an.empty()) {
    if ((rc = do_poa(g, 0, 1, t.data(), qlen, 1, &r, nullptr)) != MPOA_OK) return rc;
    return put_res(r, res, cap, n_res);
  }
  int beg_id = 0, beg_q = 0, j = 0;
  for (const mpoa::seed::Pair& a : an) {
    const int q_span = (int)(a.y >> 32 & 0xff);
    const int end_tpos = (int)(uint32_t)a.x - q_span + 1;
    const int end_qpos = (int)(uint32_t)a.y - q_span + 1;
    if (end_tpos < 0 || end_tpos + q_span > L) {
      g_err = "anchor outside the consensus";
      return MPOA_E_ARG;
    }
    const int sub_len = end_qpos - beg_q;
    int end_id = g->cons[3 * end_tpos];
    if ((rc = do_poa(g, beg_id, end_id, t.data() + j, sub_len, 1, &r, nullptr, true)) !=
        MPOA_OK)
      return rc;
    j += sub_len;
    for (int k = 0; k < q_span; k++, j++) {
      end_id = g->cons[3 * (end_tpos + k)];
      if (k != q_span - 1) r.push_back(mpoa_res{end_id, g->cons[3 * (end_tpos + k) + 2], -1});
    }
    beg_id = end_id;
    beg_q = j;
  }
  if ((rc = do_poa(g, beg_id, 1, t.data() + j, qlen - beg_q, 1, &r, nullptr)) != MPOA_OK)
    return rc;
  return put_res(r, res, cap, n_res);
}

int mpoa_graph_add_path(mpoa_graph* g, int32_t seq_id, const mpoa_res* res, int64_t n_res) {
  if (!g || !res || n_res < 1 || n_res > INT32_MAX / 4) return MPOA_E_ARG;
  if (seq_id != (int)g->read_len.size()) {
    g_err = "reads are added in order: seq_id must be the number of reads added so far";
    return MPOA_E_ARG;
  }
  const int N = g->st.n_nodes;
  for (int64_t i = 0; i < n_res; i++) {
    const mpoa_res& e = res[i];
    if (e.from < -1 || e.from >= N || e.base < 0 || e.base > 4 || e.aligned_id < -1 ||
        e.aligned_id >= N) {
      g_err = "res entry " + std::to_string(i) + " names no node of the graph";
      return MPOA_E_ARG;
    }
  }
  HIPCHK(hipSetDevice(g->ctx->device));
  int rc;
  const bool keep_path = g->params.result != 0;
  const int64_t len = n_res - 1;
  SlotCaps nc = g->caps;
  bool grow = false;
  const int64_t nodes = (int64_t)N + n_res;
  const int64_t edges = (int64_t)std::max(g->st.n_in_e, g->st.n_out_e) + n_res + 1;
  if (nodes > g->caps.nodes) nc.nodes = (int)std::max<int64_t>(nodes, g->caps.nodes * 3LL / 2), grow = true;
  if (edges > g->caps.edges) nc.edges = (int)std::max<int64_t>(edges, g->caps.edges * 3LL / 2), grow = true;
  if (len > g->caps.qmax) nc.qmax = (int)std::max<int64_t>(len, g->caps.qmax * 3LL / 2), grow = true;
  if (keep_path && g->path0 + len > g->caps.path)
    nc.path = (int)std::max<int64_t>(g->path0 + len, g->caps.path * 3LL / 2), grow = true;
  if (grow && (rc = regrow(g, nc)) != MPOA_OK) return rc;
  const Arena A = arena_of(g);
  std::vector<int> from((size_t)n_res), al((size_t)n_res);
  std::vector<uint8_t> base((size_t)n_res);
  for (int64_t i = 0; i < n_res; i++) {
    from[i] = res[i].from;
    al[i] = res[i].aligned_id;
    base[i] = (uint8_t)res[i].base;
  }
  HIPCHK(hipMemcpy(A.res_from, from.data(), (size_t)n_res * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(A.res_aligned, al.data(), (size_t)n_res * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(A.res_base, base.data(), (size_t)n_res, hipMemcpyHostToDevice));
  GraphOpArgs op{};
  op.op = mpoa::GOP_ADD_PATH;
  op.seq_id = seq_id;
  op.n_res = (int)n_res;
  op.path0 = (int)g->path0;
  GraphOpOut out{};
  if ((rc = run_op(g, op, &out)) != MPOA_OK) return rc;
  if ((rc = check_status(out.status)) != MPOA_OK) return rc;
  g->read_len.push_back((int)len);
  g->path0 += len;
  g->sorted = -1;
  return MPOA_OK;
}

int mpoa_graph_read_gfa(mpoa_graph* g, const char* text, int64_t len, char* names,
                        int64_t names_cap, int64_t* names_len, int32_t* n_paths) {
  if (!g || len < 0 || (len > 0 && !text)) return MPOA_E_ARG;
  if (g->st.n_nodes != 2 || !g->read_len.empty()) {
    g_err = "a GFA graph is read into a graph that holds only the source and the sink";
    return MPOA_E_ARG;
  }
  if (g->params.result == 2) {
    // minipoa's GFA output walks read_paths, which read_gfa leaves empty for the GFA's paths
    g_err = "GFA input with GFA output (-i with -r 2) is not supported";
    return MPOA_E_ARG;
  }
  HIPCHK(hipSetDevice(g->ctx->device));
  // Lines in order (handle_input.cpp:455-479): S adds a segment's bases as new nodes
  // (gfa_parse_S, :78-127), P adds a path over existing segments (gfa_parse_P, :128-261);
  // other lines are skipped. Nodes come first on the device (a path only joins existing nodes).
  struct Seg { int in, out; };
  std::unordered_map<std::string, Seg> segs;
  std::vector<uint8_t> bases;  // node 2 + k has base bases[k]
  std::vector<std::string> path_names;
  std::vector<std::vector<int>> paths;
  int next = 2;
  int64_t at = 0;
  while (at < len) {
    const char* nl = static_cast<const char*>(std::memchr(text + at, '\n', (size_t)(len - at)));
    const int64_t end = nl ? nl - text : len;
    std::string line(text + at, (size_t)(end - at));
    at = end + 1;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.size() < 3 || line[1] != '\t') continue;
    std::vector<std::string> f;
    for (size_t a = 2;;) {
      const size_t t = line.find('\t', a);
      f.push_back(line.substr(a, t == std::string::npos ? std::string::npos : t - a));
      if (t == std::string::npos) break;
      a = t + 1;
    }
    if (line[0] == 'S') {
      if (f.size() < 2 || f[1].empty()) {
        g_err = "GFA segment line without a sequence: " + f[0];
        return MPOA_E_ARG;
      }
      if (segs.count(f[0])) {
        g_err = "duplicated GFA segment: " + f[0];
        return MPOA_E_ARG;
      }
      segs[f[0]] = Seg{next, next + (int)f[1].size() - 1};
      for (char c : f[1]) bases.push_back(code_of_char((unsigned char)c));
      next += (int)f[1].size();
    } else if (line[0] == 'P') {
      if (f.size() < 2) {
        g_err = "GFA path line without a path: " + f[0];
        return MPOA_E_ARG;
      }
      std::vector<int> nodes;
      const std::string& w = f[1];
      size_t a = 0;
      for (size_t k = 0; k < w.size(); k++) {
        if (w[k] == '-') {
          g_err = "reverse-strand GFA path (" + f[0] + "): only '+' segments are read";
          return MPOA_E_ARG;
        }
        if (w[k] != '+') continue;
        auto it = segs.find(w.substr(a, k - a));
        if (it == segs.end()) {
          g_err = "GFA path " + f[0] + " names an unknown segment: " + w.substr(a, k - a);
          return MPOA_E_ARG;
        }
        for (int v = it->second.in; v <= it->second.out; v++) nodes.push_back(v);
        a = k + 2;  // past "+,"
      }
      if (nodes.empty()) {
        g_err = "GFA path without segments: " + f[0];
        return MPOA_E_ARG;
      }
      path_names.push_back(f[0]);
      paths.push_back(std::move(nodes));
    }
  }
  int rc;
  if (!bases.empty()) {
    const int cnt = (int)bases.size();
    if (2 + cnt > g->caps.nodes) {
      SlotCaps nc = g->caps;
      nc.nodes = 2 + cnt + cnt / 2;
      nc.edges = std::max(g->caps.edges, 2 * (2 + cnt));
      if ((rc = regrow(g, nc)) != MPOA_OK) return rc;
    }
    if (!g->d_q.p || g->d_q.n < (size_t)cnt)
      if ((rc = g->d_q.alloc((size_t)cnt)) != MPOA_OK) return rc;
    HIPCHK(hipMemcpy(g->d_q.p, bases.data(), (size_t)cnt, hipMemcpyHostToDevice));
    GraphOpArgs op{};
    op.op = mpoa::GOP_ADD_NODES;
    op.q = g->d_q.p;
    op.qlen = cnt;
    GraphOpOut out{};
    if ((rc = run_op(g, op, &out)) != MPOA_OK) return rc;
    if ((rc = check_status(out.status)) != MPOA_OK) return rc;
  }
  // each path: the source, its nodes, the sink, as a res of existing nodes (add_adj per step,
  // positions counted from 0: handle_input.cpp:196-205, 248-250)
  for (size_t k = 0; k < paths.size(); k++) {
    std::vector<mpoa_res> res;
    res.push_back(mpoa_res{0, 4, -1});
    for (int v : paths[k]) res.push_back(mpoa_res{v, bases[(size_t)(v - 2)], -1});
    if ((rc = mpoa_graph_add_path(g, (int32_t)g->read_len.size(), res.data(),
                                  (int64_t)res.size())) != MPOA_OK)
      return rc;
  }
  if ((rc = do_topsort(g, 0)) != MPOA_OK) return rc;  // handle_input.cpp:482
  std::string all;
  for (const std::string& nm : path_names) all += nm + "\n";
  if (n_paths) *n_paths = (int32_t)path_names.size();
  if (names_len) *names_len = (int64_t)all.size();
  if (names) {
    if ((int64_t)all.size() > names_cap) return MPOA_E_OUT_CAP;
    std::memcpy(names, all.data(), all.size());
  }
  return MPOA_OK;
}

int mpoa_graph_topsort(mpoa_graph* g, int32_t op) {
  if (!g || (op != 0 && op != 1)) return MPOA_E_ARG;
  HIPCHK(hipSetDevice(g->ctx->device));
  return do_topsort(g, op);
}

int mpoa_graph_output(mpoa_graph* g, int32_t result, const char* name, const int64_t* name_off,
                      const char* comment, const int64_t* comment_off, const int32_t* rid_to_ord,
                      char* out, int64_t out_cap, int64_t* out_len) {
  if (!g || result < 0 || result > 3) return MPOA_E_ARG;
  if (result != 0 && g->params.result == 0) {
    g_err = "MSA and GFA output need a graph created with result 1, 2 or 3 (read paths kept)";
    return MPOA_E_ARG;
  }
  HIPCHK(hipSetDevice(g->ctx->device));
  int rc;
  // output_rc_msa sorts an unsorted graph by aligned groups first (graph.cpp:238)
  if ((result == 1 || result == 3) && g->sorted < 0 && (rc = do_topsort(g, 1)) != MPOA_OK)
    return rc;
  const int n = (int)g->read_len.size();
  const int64_t cap = std::min<int64_t>(text_cap(result, g->read_len), INT32_MAX / 2);
  if (!g->d_text.p || g->d_text.n < (size_t)cap)
    if ((rc = g->d_text.alloc((size_t)cap)) != MPOA_OK) return rc;
  if (!g->d_gfa.p || g->d_gfa.n < (size_t)n + 1)
    if ((rc = g->d_gfa.alloc((size_t)n + 1)) != MPOA_OK) return rc;
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int i = 0; i < n; i++) off[i + 1] = off[i] + g->read_len[i];
  if (!g->d_read_off.p || g->d_read_off.n < (size_t)n + 1)
    if ((rc = g->d_read_off.alloc((size_t)n + 1)) != MPOA_OK) return rc;
  HIPCHK(hipMemcpy(g->d_read_off.p, off.data(), off.size() * sizeof(int64_t),
                   hipMemcpyHostToDevice));
  WinDesc wd{};
  wd.read0 = 0;
  wd.n_reads = n;
  wd.out0 = 0;
  wd.out_cap = (int)cap;
  HIPCHK(hipMemcpy(g->d_win.p, &wd, sizeof wd, hipMemcpyHostToDevice));
  // the MSA's columns are the ranks of the last order: the aligned groups after an op 1 sort,
  // else the nodes (rank.size() - 2 columns, graph.cpp:239)
  GraphOpArgs op{};
  op.op = mpoa::GOP_OUTPUT;
  op.result = result;
  GraphOpOut o{};
  if ((rc = run_op(g, op, &o)) != MPOA_OK) return rc;
  if ((rc = check_status(o.status)) != MPOA_OK) return rc;
  // the consensus output re-sorts the graph (op 0 without hlen/tlen): a later poa() needs a sort
  if ((result == 0 || result == 3) && g->sorted != 0) g->sorted = -1;
  WinResult r{};
  HIPCHK(hipMemcpy(&r, g->d_res.p, sizeof r, hipMemcpyDeviceToHost));
  std::string text;
  if (rid_to_ord) {
    std::vector<char> seen((size_t)n, 0);
    for (int i = 0; i < n; i++) {
      if (rid_to_ord[i] < 0 || rid_to_ord[i] >= n || seen[rid_to_ord[i]]) {
        g_err = "rid_to_ord is not a permutation of the reads added";
        return MPOA_E_ARG;
      }
      seen[rid_to_ord[i]] = 1;
    }
    if (result == 2) {
      g_err = "a read order with GFA output: minipoa's GFA writer indexes its read paths past "
              "their ends then (graph.cpp:369,423)";
      return MPOA_E_ARG;
    }
  }
  if ((rc = window_text(result, r, g->d_text.p, g->d_gfa.p, n, name, name_off, comment,
                        comment_off, rid_to_ord, &text)) != MPOA_OK)
    return rc;
  if (out_len) *out_len = (int64_t)text.size();
  if (!out) return MPOA_OK;
  if ((int64_t)text.size() > out_cap) return MPOA_E_OUT_CAP;
  std::memcpy(out, text.data(), text.size());
  return MPOA_OK;
}

int mpoa_graph_export(mpoa_graph* g, int32_t* n_nodes, int32_t* n_edges, uint8_t* base,
                      int32_t* in_off, int32_t* in_src, int32_t* in_w, int32_t* out_off,
                      int32_t* out_dst, int32_t* out_w, int32_t* rank_to_id, int32_t* hlen,
                      int32_t* tlen, int32_t* par_id, int32_t* aligned) {
  if (!g) return MPOA_E_ARG;
  HIPCHK(hipSetDevice(g->ctx->device));
  const int N = g->st.n_nodes;
  const Arena A = arena_of(g);
  const size_t E = (size_t)std::max(g->st.n_in_e, g->st.n_out_e);
  auto get = [&](std::vector<int>& v, const int* src, size_t n) {
    v.resize(std::max<size_t>(n, 1));
    return n ? hipMemcpy(v.data(), src, n * sizeof(int), hipMemcpyDeviceToHost) : hipSuccess;
  };
  std::vector<int> in_head, ie_src, ie_w, ie_next, out_head, oe_dst, oe_w, oe_next;
  HIPCHK(get(in_head, A.in_head, N));
  HIPCHK(get(ie_src, A.ie_src, E));
  HIPCHK(get(ie_w, A.ie_w, E));
  HIPCHK(get(ie_next, A.ie_next, E));
  HIPCHK(get(out_head, A.out_head, N));
  HIPCHK(get(oe_dst, A.oe_dst, E));
  HIPCHK(get(oe_w, A.oe_w, E));
  HIPCHK(get(oe_next, A.oe_next, E));
  int ne = 0;
  for (int v = 0; v < N; v++)
    for (int e = in_head[v]; e != -1; e = ie_next[e]) ne++;
  if (n_nodes) *n_nodes = N;
  if (n_edges) *n_edges = ne;
  if (!base) return MPOA_OK;
  HIPCHK(hipMemcpy(base, A.base, (size_t)N, hipMemcpyDeviceToHost));
  if (par_id) HIPCHK(hipMemcpy(par_id, A.par, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
  if (aligned)
    HIPCHK(hipMemcpy(aligned, A.aligned, (size_t)N * mpoa::NB * sizeof(int),
                     hipMemcpyDeviceToHost));
  int ai = 0, ao = 0;
  for (int v = 0; v < N; v++) {
    if (in_off) in_off[v] = ai;
    for (int e = in_head[v]; e != -1; e = ie_next[e], ai++) {
      if (in_src) in_src[ai] = ie_src[e];
      if (in_w) in_w[ai] = ie_w[e];
    }
    if (out_off) out_off[v] = ao;
    for (int e = out_head[v]; e != -1; e = oe_next[e], ao++) {
      if (out_dst) out_dst[ao] = oe_dst[e];
      if (out_w) out_w[ao] = oe_w[e];
    }
  }
  if (in_off) in_off[N] = ai;
  if (out_off) out_off[N] = ao;
  if (rank_to_id) HIPCHK(hipMemcpy(rank_to_id, A.rank_id, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
  if (hlen) HIPCHK(hipMemcpy(hlen, A.hlen, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
  if (tlen) HIPCHK(hipMemcpy(tlen, A.tlen, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
  return MPOA_OK;
}

}  // extern "C"
