// Host side of the C ABI (include/mpoa.h): device context, batch sizing, launches of the
// device pipeline (pipeline.hip), overflow re-runs and output formatting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <iterator>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpoa.h"
#include "pipeline.h"
#include "host.h"
#include "seed.h"

using mpoa::CallStat;
using mpoa::DpScores;
using mpoa::KernelArgs;
using mpoa::SlotCaps;
using mpoa::WinDesc;
using mpoa::WinResult;
using mpoa::SeedArgs;
using mpoa::SegDesc;
using mpoa::SegResult;
using mpoa::WinState;
namespace seed = mpoa::seed;

namespace mpoa {
namespace host {
thread_local std::string g_err;
}  // namespace host
}  // namespace mpoa

namespace {

using namespace mpoa::host;


mpoa_call_stat call_stat_of(const CallStat& c, int read, int window) {
  mpoa_call_stat o{};
  o.read = read;
  o.rows = c.rows;
  o.qlen = c.qlen;
  o.score = c.score;
  o.cells = c.cells;
  o.edge_cells = c.edge_cells;
  o.logical_cells = c.logical;
  o.width = c.width;
  o.nodes = c.nodes;
  o.edges = c.edges;
  o.window = window;
  return o;
}

}  // namespace

// One persistent launch configuration: a set of windows, a tile shape and slot capacities.
struct Bin {
  int nt = 64, cpl = 4;
  std::vector<int> wins;  // processing order
  SlotCaps caps{};
  int n_slots = 0;        // allocated slots
  bool mem_limited = false;
  DevBuf<int> d_int;
  DevBuf<uint8_t> d_byte;
  DevBuf<int> d_dp;
  DevBuf<int> d_order;
  DevBuf<int> d_queue;
  int used_slots = 0;     // slots of the last launch
};

struct mpoa_batch {
  mpoa_ctx* ctx = nullptr;
  mpoa_params params{};
  // -p: the reads permuted into each window's guide-tree order (host copy, what was uploaded)
  // and, per read in input order, the MSA row it prints (minipoa's rid_to_ord)
  std::string perm_seq;
  std::vector<int64_t> perm_off;
  std::vector<int> row_of;
  DpScores sc{};
  int n_windows = 0;
  std::vector<WinInfo> info;
  std::vector<WinDesc> desc;
  std::vector<int> order;     // processing order (largest predicted first)
  std::vector<double> cost;   // predicted cost per window
  int64_t out_total = 0;
  DevBuf<uint8_t> d_seq;
  DevBuf<int64_t> d_off;
  DevBuf<WinDesc> d_win;
  DevBuf<WinResult> d_res;
  DevBuf<uint8_t> d_out;
  Bin small, large, wide, rerun;  // large: the medium tile
  std::vector<int> bin_of;    // launch of each window (0 small, 1 medium, 2 re-run, 3 wide)
  bool split = true;          // large windows in their own concurrent launch
  // schedule knobs (mpoa_batch_set_schedule): windows the small tile takes, and how many of
  // the costliest others join the large launch (-1: half a window per CU)
  int small_qmax = 1500, small_width = 344, move_top = -1, small_nodes = 1 << 30;
  // the costliest medium-tile windows that run on the wide tile instead (faster per window)
  int wide_top = 8;
  std::vector<WinResult> res;
  int64_t reruns = 0;
  double device_ms = 0;
  double kernel_ms = 0;
  int launches = 0;
  int first_slots = 0;
  std::vector<mpoa_launch_stat> lstats;  // launches of the last execute
  DevBuf<int64_t> d_gfa;                  // one-window runs: offsets of the GFA path lines
  int mode = MPOA_MODE_BATCHED;
  int forms = 0;  // MPOA_FORM_* test hook
  // Learned predictions per window (mpoa_batch_set_predictions): DP cells of the costliest
  // call, widest band row and final node count. Empty when none were given.
  std::vector<float> pred_cells, pred_width, pred_nodes;
  // per-call statistics (single-window runs)
  DevBuf<CallStat> d_calls;
  DevBuf<int64_t> d_call_off;
  std::vector<int64_t> call_off;  // per window: first call record
  // Graph form of the batched pass (mpoa_batch_set_graph): per launch an event record, a
  // queue reset, the persistent kernel and an event record, the three launches as parallel
  // branches. Rebuilt when a launch's arguments change; replayed otherwise.
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> graph_key;
  hipEvent_t gev[6] = {};
  int graph_builds = 0;
  int graph_launches = 0;  // in the last execute
  ~mpoa_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : gev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

SlotCaps make_caps(const std::vector<Estimate>& est, const std::vector<WinInfo>& info,
                   const std::vector<int>& wins, bool keep_path) {
  SlotCaps c{};
  c.nodes = 2;
  c.edges = 4;
  c.qmax = 1;
  c.path = 0;
  c.dp = 64;
  for (int w : wins) {
    c.nodes = std::max(c.nodes, est[w].nodes);
    c.edges = std::max(c.edges, est[w].edges);
    c.qmax = std::max(c.qmax, info[w].qmax);
    if (keep_path) c.path = (int)std::max<int64_t>(c.path, info[w].qtot);
    c.dp = std::max(c.dp, est[w].dp);
  }
  c.nodes = (c.nodes + 63) / 64 * 64;
  c.edges = (c.edges + 63) / 64 * 64;
  c.dp = (c.dp + 63) / 64 * 64;
  mpoa::finish_caps(c);
  return c;
}

size_t slot_bytes_total(const SlotCaps& c) {
  return (size_t)c.ints_per_slot * 4 + (size_t)c.bytes_per_slot + (size_t)c.dp * 12;
}

// Allocates (or reuses) a bin's slot pools for its windows and capacities, within budget
// bytes of device memory. Returns MPOA_OK or an error.
int provision(Bin& bin, const SlotCaps& caps, int max_slots, size_t budget) {
  const size_t per_slot = slot_bytes_total(caps);
  const int want = std::min<int>(max_slots, (int)bin.wins.size());
  const bool same = bin.d_dp.p && caps.nodes == bin.caps.nodes && caps.edges == bin.caps.edges &&
                    caps.qmax == bin.caps.qmax && caps.path == bin.caps.path && caps.dp == bin.caps.dp;
  if (same && (bin.n_slots >= want || bin.mem_limited)) {
    bin.used_slots = std::min(bin.n_slots, want);
    return MPOA_OK;
  }
  bin.d_int.release();
  bin.d_byte.release();
  bin.d_dp.release();
  const size_t fit = budget / per_slot;
  const int n_slots = (int)std::min<size_t>((size_t)want, fit);
  if (n_slots < 1) {
    g_err = "window too large for device memory";
    return MPOA_E_OOM;
  }
  bin.mem_limited = (size_t)n_slots == fit && n_slots < want;
  int rc;
  if ((rc = bin.d_int.alloc((size_t)n_slots * caps.ints_per_slot)) != MPOA_OK) return rc;
  if ((rc = bin.d_byte.alloc((size_t)n_slots * caps.bytes_per_slot)) != MPOA_OK) return rc;
  if ((rc = bin.d_dp.alloc((size_t)n_slots * 3 * caps.dp)) != MPOA_OK) return rc;
  if (!bin.d_order.p || bin.d_order.n < bin.wins.size())
    if ((rc = bin.d_order.alloc(std::max<size_t>(bin.wins.size(), 1))) != MPOA_OK) return rc;
  if (!bin.d_queue.p && (rc = bin.d_queue.alloc(1)) != MPOA_OK) return rc;
  bin.n_slots = n_slots;
  bin.used_slots = n_slots;
  bin.caps = caps;
  return MPOA_OK;
}

// Uploads a bin's processing order on `stream` (allocating its order and queue buffers).
int upload_order(Bin& bin, hipStream_t stream) {
  int rc;
  if (!bin.d_order.p || bin.d_order.n < bin.wins.size())
    if ((rc = bin.d_order.alloc(bin.wins.size())) != MPOA_OK) return rc;
  if (!bin.d_queue.p && (rc = bin.d_queue.alloc(1)) != MPOA_OK) return rc;
  HIPCHK(hipMemcpyAsync(bin.d_order.p, bin.wins.data(), bin.wins.size() * sizeof(int),
                        hipMemcpyHostToDevice, stream));
  return MPOA_OK;
}

// The persistent kernel's arguments for a provisioned bin. Re-runs store the DP matrices as
// int32: a window re-run because a call's values left int16 must not take that form again.
KernelArgs args_of(const mpoa_batch* b, const Bin& bin) {
  KernelArgs a{};
  a.forms = b->forms | (&bin == &b->rerun ? MPOA_FORM_WIDE : 0);
  a.scores = b->sc;
  a.caps = bin.caps;
  a.read_seq = b->d_seq.p;
  a.read_off = b->d_off.p;
  a.win = b->d_win.p;
  a.result = b->d_res.p;
  a.order = bin.d_order.p;
  a.n_jobs = (int)bin.wins.size();
  a.queue = bin.d_queue.p;
  a.int_pool = bin.d_int.p;
  a.byte_pool = bin.d_byte.p;
  a.dp_pool = bin.d_dp.p;
  a.out = b->d_out.p;
  a.call_stats = b->d_calls.p;
  a.call_off = b->d_call_off.p;
  a.gfa_lines = b->d_gfa.p;
  return a;
}

// Enqueues one persistent launch of a provisioned bin on `stream` (no wait).
int enqueue(mpoa_batch* b, Bin& bin, hipStream_t stream) {
  if (bin.wins.empty()) return MPOA_OK;
  int rc;
  if ((rc = upload_order(bin, stream)) != MPOA_OK) return rc;
  HIPCHK(hipMemsetAsync(bin.d_queue.p, 0, sizeof(int), stream));
  const KernelArgs a = args_of(b, bin);
  const int lrc = mpoa_launch_windows(a, bin.used_slots, bin.nt, bin.cpl, stream);
  if (lrc != 0) {
    g_err = lrc == -2 ? "unsupported tile shape" : "kernel launch failed";
    return lrc == -2 ? MPOA_E_ARG : MPOA_E_DEVICE;
  }
  b->launches++;
  return MPOA_OK;
}

// A launch in flight: its bin's windows and the HIP events around it on its stream.
struct Pending {
  int kind;
  const Bin* bin;
  std::vector<int> wins;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool owned = true;  // events created for this launch (the graph's are the batch's)
};

int timed_enqueue(mpoa_batch* b, Bin& bin, hipStream_t stream, int kind,
                  std::vector<Pending>& pend) {
  if (bin.wins.empty()) return MPOA_OK;
  Pending p;
  p.kind = kind;
  p.bin = &bin;
  p.wins = bin.wins;
  HIPCHK(hipEventCreate(&p.e0));
  HIPCHK(hipEventCreate(&p.e1));
  HIPCHK(hipEventRecord(p.e0, stream));
  const int rc = enqueue(b, bin, stream);
  HIPCHK(hipEventRecord(p.e1, stream));
  pend.push_back(std::move(p));
  return rc;
}

// Turns finished launches into records (events synchronised, results copied back).
int close_pending(mpoa_batch* b, std::vector<Pending>& pend, const std::vector<WinResult>& res) {
  for (Pending& p : pend) {
    HIPCHK(hipEventSynchronize(p.e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, p.e0, p.e1));
    mpoa_launch_stat ls{};
    ls.launch = p.kind;
    ls.threads = p.bin->nt;
    ls.cols = p.bin->cpl;
    ls.slots = p.bin->used_slots;
    ls.windows = (int64_t)p.wins.size();
    for (int w : p.wins) {
      ls.cells += res[w].cells;
      ls.edge_cells += res[w].edge_cells;
    }
    ls.ms = ms;
    b->lstats.push_back(ls);
    if (p.owned) {
      (void)hipEventDestroy(p.e0);
      (void)hipEventDestroy(p.e1);
    }
  }
  pend.clear();
  return MPOA_OK;
}

// The batched pass as one graph launch on the context's stream: the wide, medium and small
// launches as parallel branches (queue reset, persistent kernel, an event record on either
// side), the small one held 2 ms so that the larger tiles hold their CUs first.
int graph_pass(mpoa_batch* b, std::vector<Pending>& pend) {
  mpoa_ctx* ctx = b->ctx;
  Bin* bins[3] = {&b->wide, &b->large, &b->small};
  const int kinds[3] = {3, 1, 0};
  std::vector<unsigned char> key;
  int rc;
  for (int k = 0; k < 3; k++) {
    Bin& bin = *bins[k];
    if (bin.wins.empty()) continue;
    if ((rc = upload_order(bin, ctx->stream)) != MPOA_OK) return rc;
    const KernelArgs a = args_of(b, bin);
    const int geo[4] = {k, bin.used_slots, bin.nt, bin.cpl};
    const unsigned char* pa = reinterpret_cast<const unsigned char*>(&a);
    const unsigned char* pg = reinterpret_cast<const unsigned char*>(geo);
    key.insert(key.end(), pa, pa + sizeof a);
    key.insert(key.end(), pg, pg + sizeof geo);
  }
  if (key.empty()) return MPOA_OK;
  if (!b->gexec || key != b->graph_key) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_key.clear();
    for (hipEvent_t& e : b->gev)
      if (!e) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipGraphCreate(&b->graph, 0));
    const bool others = !b->wide.wins.empty() || !b->large.wins.empty();
    for (int k = 0; k < 3; k++) {
      Bin& bin = *bins[k];
      if (bin.wins.empty()) continue;
      hipGraphNode_t head = nullptr, e0, kw, e1;
      size_t n_head = 0;
      if (kinds[k] == 0 && others) {
        if (mpoa_graph_add_hold(b->graph, nullptr, 0, (int64_t)ctx->wall_khz * 2, &head) != 0) {
          g_err = "graph construction failed";
          return MPOA_E_DEVICE;
        }
        n_head = 1;
      }
      HIPCHK(hipGraphAddEventRecordNode(&e0, b->graph, n_head ? &head : nullptr, n_head,
                                        b->gev[2 * k]));
      const int grc = mpoa_graph_add_windows(b->graph, &e0, 1, args_of(b, bin), bin.used_slots,
                                             bin.nt, bin.cpl, &kw);
      if (grc != 0) {
        g_err = grc == -2 ? "unsupported tile shape" : "graph construction failed";
        return grc == -2 ? MPOA_E_ARG : MPOA_E_DEVICE;
      }
      HIPCHK(hipGraphAddEventRecordNode(&e1, b->graph, &kw, 1, b->gev[2 * k + 1]));
    }
    HIPCHK(hipGraphInstantiateWithFlags(&b->gexec, b->graph, 0));
    b->graph_key = key;
    b->graph_builds++;
  }
  HIPCHK(hipGraphLaunch(b->gexec, ctx->stream));
  b->graph_launches++;
  for (int k = 0; k < 3; k++) {
    Bin& bin = *bins[k];
    if (bin.wins.empty()) continue;
    Pending p;
    p.kind = kinds[k];
    p.bin = &bin;
    p.wins = bin.wins;
    p.e0 = b->gev[2 * k];
    p.e1 = b->gev[2 * k + 1];
    p.owned = false;
    pend.push_back(std::move(p));
    b->launches++;
  }
  return MPOA_OK;
}

int upload(mpoa_batch* b, int32_t n_windows, const int32_t* win_off, const char* seq,
           const int64_t* seq_off) {
  mpoa_ctx* ctx = b->ctx;
  const int32_t r0 = win_off[0], r1 = win_off[n_windows];
  const int64_t base0 = seq_off[r0], bases = seq_off[r1] - base0;
  std::vector<int64_t> off(r1 - r0 + 1);
  for (int32_t r = r0; r <= r1; r++) off[r - r0] = seq_off[r] - base0;
  int rc;
  if ((rc = b->d_seq.alloc((size_t)bases + 16)) != MPOA_OK) return rc;
  if ((rc = b->d_off.alloc(off.size())) != MPOA_OK) return rc;
  HIPCHK(hipMemcpyAsync(b->d_seq.p, seq + base0, (size_t)bases, hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(b->d_off.p, off.data(), off.size() * sizeof(int64_t),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MPOA_OK;
}

// Splits the windows (kept in processing order) between two concurrent launches. The small
// tile (64 threads, 376-column LDS ring, query up to 1,500 bp in LDS, 27 KB of LDS: six
// workgroups per CU) takes every window it fits. The other launch uses the medium tile (128
// threads, 512-column ring, 4 KB query, ~38 KB) when all its windows fit that, else the wide
// tile (256 threads, 1,280-column ring, 24 KB query, ~88 KB). A single window (minipoa's own
// run) always takes the wide tile: there the per-row latency of one window is the whole time.
void assign_bins(mpoa_batch* b) {
  b->small.wins.clear();
  b->large.wins.clear();
  b->wide.wins.clear();
  const bool full = b->params.ab_band != 0 && b->params.band_f > 0;
  auto width_of = [&](const WinInfo& wi) {
    const int m = wi.qmax + 1;
    if (b->sc.f > 0 && !full) return 2 * (b->sc.b + m / b->sc.f) + 24;
    return (m / 8 + 2) * 8;
  };
  for (int w : b->order) {
    const WinInfo& wi = b->info[w];
    const int width = b->pred_width.empty() ? width_of(wi) : (int)b->pred_width[w];
    const int nodes = b->pred_nodes.empty() ? estimate(wi, b->sc, full).nodes
                                            : (int)b->pred_nodes[w];
    const bool large = !b->split || wi.qmax > b->small_qmax || width > b->small_width ||
                       nodes > b->small_nodes;
    // The wide tile for bands beyond the medium tile's ring and graphs beyond what its LDS
    // sorts (about 4,300 nodes).
    const bool wide = !b->split || width > 496 || nodes > 4200;
    (!large ? b->small.wins : wide ? b->wide.wins : b->large.wins).push_back(w);
  }
  b->large.nt = 128;
  b->large.cpl = 4;
  // One window (minipoa's own run) takes the 512-thread tile: its rows are the whole time.
  b->wide.nt = b->split ? 256 : 512;
  b->wide.cpl = 2;
  // One window alone whose band passes the 512 threads' 1,024 columns: four columns a thread,
  // so that every row is one pass (two passes cost twice the row's scan and barriers).
  // A band of at most 448 columns with a query the 256-thread tile holds in LDS: that tile,
  // whose rows need half the waves at each barrier and scan (a 2 kb, 100-read window: DP
  // 1.13 s against 1.35 s on the 512-thread tile).
  if (!b->split) {
    bool narrow = !b->wide.wins.empty();
    for (int w : b->wide.wins) {
      const WinInfo& wi = b->info[w];
      const int width = b->pred_width.empty() ? width_of(wi) : (int)b->pred_width[w];
      if (width > 1024) b->wide.cpl = 4;
      narrow = narrow && width <= 448 && wi.qmax <= 4000;
    }
    if (narrow) b->wide.nt = 256;
  }
  b->bin_of.assign(b->n_windows, 0);
  // The costliest small-tile windows join the medium launch up to move_top windows (default
  // half a window per CU): they set the makespan, and the larger tile shortens their rows.
  const size_t target = (size_t)std::max(0, b->move_top >= 0 ? b->move_top : b->ctx->n_cu / 2);
  if (b->split && b->large.wins.size() < target && !b->small.wins.empty()) {
    const size_t take = std::min(target - b->large.wins.size(), b->small.wins.size() / 8);
    std::vector<int> moved(b->small.wins.begin(), b->small.wins.begin() + take);
    b->small.wins.erase(b->small.wins.begin(), b->small.wins.begin() + take);
    std::vector<int> merged;
    std::merge(b->large.wins.begin(), b->large.wins.end(), moved.begin(), moved.end(),
               std::back_inserter(merged), [&](int x, int y) { return b->cost[x] > b->cost[y]; });
    b->large.wins = merged;
  }
  // The costliest windows of the medium launch set its end; on the wide tile the same window
  // runs ~1.2x faster (256 threads, two columns each, against 128 with four), so the first
  // wide_top of them move there.
  if (b->split && b->wide_top > 0 && !b->large.wins.empty()) {
    const size_t k = std::min<size_t>((size_t)b->wide_top, b->large.wins.size());
    std::vector<int> top(b->large.wins.begin(), b->large.wins.begin() + k);
    b->large.wins.erase(b->large.wins.begin(), b->large.wins.begin() + k);
    std::vector<int> merged;
    std::merge(b->wide.wins.begin(), b->wide.wins.end(), top.begin(), top.end(),
               std::back_inserter(merged), [&](int x, int y) { return b->cost[x] > b->cost[y]; });
    b->wide.wins = merged;
  }
  for (int w : b->large.wins) b->bin_of[w] = 1;
  for (int w : b->wide.wins) b->bin_of[w] = 3;
}

// Slot capacities for window w: the learned predictions when given (the caller includes its
// safety margin; a window that outgrows them overflows and is re-run with worst-case
// capacities), else the read-length estimate.
Estimate capacity_of(const mpoa_batch* b, int w, bool full) {
  Estimate e = estimate(b->info[w], b->sc, full);
  if (b->pred_cells.empty()) return e;
  const WinInfo& wi = b->info[w];
  const int64_t nodes = std::min<int64_t>(
      std::max<int64_t>((int64_t)b->pred_nodes[w] + 16, wi.qmax + 2), 2 + wi.qtot);
  e.nodes = (int)nodes;
  e.edges = (int)std::min<int64_t>(3 * nodes + wi.n_reads, wi.qtot + wi.n_reads + 2);
  // storage of one call: its cells, one sentinel block per row and row 0
  e.dp = (int64_t)b->pred_cells[w] + 8 * nodes + (int64_t)(wi.qmax / 8 + 2) * 8 + 64;
  if (full) e.dp = std::max(e.dp, estimate(wi, b->sc, full).dp);
  return e;
}

int run_seeded_dev(mpoa_batch* b, const int64_t* seq_off, std::vector<mpoa_call_stat>* calls);
int run_seeded(mpoa_batch* b, const char* seq, const int64_t* seq_off,
               std::vector<mpoa_call_stat>* calls);

// A persistent launch of k_seeded in flight: its seeding buffers and events.
struct SeededLaunch {
  Bin* bin = nullptr;
  int nt = 0, n_slots = 0;
  DevBuf<uint64_t> u_buf;
  DevBuf<int> i_buf;
  DevBuf<SegDesc> segs;
  DevBuf<mpoa::SeedDev> d_seed;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~SeededLaunch() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// Enqueues one persistent launch of k_seeded (nt threads per window) over the seeded windows
// of `bin` (processing order) on `stream`: as many slots as the tile keeps resident and
// `budget` bytes allow, each with graph pools sized for the bin's largest window and its own
// seeding buffers (SeedDev). Window statuses land in d_res once it completes.
int seeded_enqueue(mpoa_batch* b, const std::vector<Estimate>& est, Bin& bin, int nt,
                   size_t budget, hipStream_t stream, SeededLaunch& L) {
  mpoa_ctx* ctx = b->ctx;
  const bool keep_path = b->params.result != 0;
  const std::vector<int>& wins = bin.wins;
  SlotCaps c = make_caps(est, b->info, wins, keep_path);
  int64_t qtot = 1;
  int nr = 1;
  for (int w : wins) {
    qtot = std::max(qtot, b->info[w].qtot);
    nr = std::max(nr, b->info[w].n_reads);
  }
  const int64_t rm = qtot + (int64_t)nr * (2 * b->params.mm_w + 8);
  const int cons_cap = c.nodes;
  const size_t u_words = 4 * (size_t)rm + (size_t)qtot + 5 * (size_t)cons_cap +
                         8 * (size_t)mpoa::SEED_ANCHORS;
  // ints: rm_cnt, n_seg, n_splice, read_graph (2), splices, flag
  const size_t i_words = (size_t)nr + 4 + 2 * (size_t)mpoa::SEED_SPLICE + 1;
  const size_t extra = u_words * 8 + i_words * 4 + mpoa::SEED_SEGS * sizeof(SegDesc) +
                       sizeof(mpoa::SeedDev);
  const size_t per_slot = slot_bytes_total(c) + extra;
  const int per_cu = std::max(1, mpoa_seeded_occupancy(nt));
  const int slots = (int)std::min<size_t>(
      std::min<size_t>(wins.size(), (size_t)std::max(1, ctx->n_cu) * per_cu), budget / per_slot);
  if (slots < 1) {
    g_err = "window too large for device memory";
    return MPOA_E_OOM;
  }
  int rc;
  const int want = (int)std::min<size_t>(wins.size(), (size_t)std::max(1, ctx->n_cu) * per_cu);
  const SlotCaps& h = bin.caps;
  if (bin.d_dp.p && bin.nt == nt && h.nodes >= c.nodes && h.edges >= c.edges &&
      h.qmax >= c.qmax && h.path >= c.path && h.dp >= c.dp &&
      (bin.n_slots >= want || bin.n_slots >= slots)) {
    // the pools of an earlier execute cover this one: keep them (re-allocating costs seconds)
    c = h;
    bin.used_slots = std::min(bin.n_slots, want);
  } else {
    // The DP pool takes what the budget leaves, up to full-width rows over the whole graph,
    // so that few sub-windows outgrow it (each would re-run its window).
    int64_t full = c.dp;
    for (int w : wins) full = std::max(full, estimate(b->info[w], b->sc, true).dp);
    const int64_t spare = (int64_t)((budget - (size_t)slots * per_slot) / (size_t)slots / 12);
    if (full > c.dp && spare > 0) {
      c.dp += std::min(full - c.dp, spare) / 64 * 64;
      mpoa::finish_caps(c);
    }
    bin.nt = nt;
    bin.cpl = nt == 64 ? 4 : 2;
    if ((rc = provision(bin, c, slots, budget - (size_t)slots * extra)) != MPOA_OK) return rc;
  }
  const int n_slots = bin.used_slots;
  L.bin = &bin;
  L.nt = nt;
  L.n_slots = n_slots;
  if ((rc = L.u_buf.alloc(n_slots * u_words)) != MPOA_OK ||
      (rc = L.i_buf.alloc(n_slots * i_words)) != MPOA_OK ||
      (rc = L.segs.alloc((size_t)n_slots * mpoa::SEED_SEGS)) != MPOA_OK ||
      (rc = L.d_seed.alloc(n_slots)) != MPOA_OK)
    return rc;
  std::vector<mpoa::SeedDev> sd(n_slots);
  for (int s = 0; s < n_slots; s++) {
    mpoa::SeedDev& d = sd[s];
    uint64_t* u = L.u_buf.p + (size_t)s * u_words;
    auto take_u = [&](size_t m) { uint64_t* p = u; u += m; return p; };
    d.rmx = take_u(rm); d.rmy = take_u(rm); d.rtx = take_u(rm); d.rty = take_u(rm);
    d.hx = take_u((size_t)qtot + cons_cap);
    d.cmx = take_u(cons_cap); d.cmy = take_u(cons_cap);
    d.ctx = take_u(cons_cap); d.cty = take_u(cons_cap);
    d.ax = take_u(mpoa::SEED_ANCHORS); d.ay = take_u(mpoa::SEED_ANCHORS);
    d.bx = take_u(mpoa::SEED_ANCHORS); d.by = take_u(mpoa::SEED_ANCHORS);
    d.zx = take_u(mpoa::SEED_ANCHORS); d.zy = take_u(mpoa::SEED_ANCHORS);
    d.zsx = take_u(mpoa::SEED_ANCHORS); d.zsy = take_u(mpoa::SEED_ANCHORS);
    int* iw = L.i_buf.p + (size_t)s * i_words;
    d.rm_cnt = iw; iw += nr;
    d.n_seg = iw; iw += 1;
    d.n_splice = iw; iw += 1;
    d.read_graph = iw; iw += 2;
    d.splice = iw; iw += 2 * mpoa::SEED_SPLICE;
    d.flag = iw;
    d.rm_off = nullptr;  // each read's minimizers at its offset in the window plus 2w + 8 each
    d.hx_cons = qtot;
    d.cons_cap = cons_cap;
    d.segs = L.segs.p + (size_t)s * mpoa::SEED_SEGS;
    // sub-windows run one at a time in the slot's own DP pool, each checked against it by
    // the DP itself (a window past it re-runs with worst-case capacities)
    d.dp_cap = INT64_MAX;
    d.k = b->params.k;
    d.w = b->params.mm_w;
    d.bw = 1000;  // main.cpp:65
    d.poa_w = b->params.poa_w;
    d.pen_gap = (float)(0.01 * d.k);  // minimizer.cpp:498
    d.pen_skip = (float)(0.0 * 0.01 * d.k);
  }
  HIPCHK(hipMemcpyAsync(L.d_seed.p, sd.data(), n_slots * sizeof(mpoa::SeedDev),
                        hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemsetAsync(L.i_buf.p, 0, n_slots * i_words * sizeof(int), stream));
  if ((rc = upload_order(bin, stream)) != MPOA_OK) return rc;
  HIPCHK(hipMemsetAsync(bin.d_queue.p, 0, sizeof(int), stream));
  KernelArgs a = args_of(b, bin);
  a.forms = b->forms;
  a.call_stats = nullptr;
  a.call_off = nullptr;
  a.seed = L.d_seed.p;
  HIPCHK(hipEventCreate(&L.e0));
  HIPCHK(hipEventCreate(&L.e1));
  HIPCHK(hipEventRecord(L.e0, stream));
  if (mpoa_launch_seeded(a, n_slots, nt, stream) != 0) {
    g_err = "seeded launch failed";
    return MPOA_E_DEVICE;
  }
  HIPCHK(hipEventRecord(L.e1, stream));
  b->launches++;
  b->first_slots += n_slots;
  return MPOA_OK;
}

// Waits for an enqueued seeded launch; its HIP-event milliseconds.
int seeded_wait(SeededLaunch& L, float* ms) {
  *ms = 0;
  if (!L.e1) return MPOA_OK;
  HIPCHK(hipEventSynchronize(L.e1));
  HIPCHK(hipEventElapsedTime(ms, L.e0, L.e1));
  return MPOA_OK;
}

// Window w of a seeded batch through the single-window form (its own sizing; device seeding,
// else host seeding): for windows past a capacity of the batched form, and for every window
// under MPOA_FORM_HOST_SEED. Its status, output and result record replace the batch's.
int seeded_single(mpoa_batch* b, int w) {
  mpoa_ctx* ctx = b->ctx;
  const WinInfo& wi = b->info[w];
  const WinDesc& wd = b->desc[w];
  std::vector<int64_t> off(wi.n_reads + 1);
  HIPCHK(hipMemcpy(off.data(), b->d_off.p + wd.read0, off.size() * sizeof(int64_t),
                   hipMemcpyDeviceToHost));
  std::string seq((size_t)(off.back() - off[0]), '\0');
  if (!seq.empty())
    HIPCHK(hipMemcpy(seq.data(), b->d_seq.p + off[0], seq.size(), hipMemcpyDeviceToHost));
  for (int i = wi.n_reads; i >= 0; i--) off[i] -= off[0];
  mpoa_params pp = b->params;
  pp.seeding = 0;
  pp.progressive = 0;  // the batch holds the reads in guide-tree order already
  int32_t win_off[2] = {0, wi.n_reads};
  mpoa_batch* nb = nullptr;
  int rc = mpoa_batch_prepare(ctx, &pp, 1, win_off, seq.data(), off.data(), nullptr, &nb);
  if (rc != MPOA_OK) return rc;
  nb->params.seeding = 1;
  nb->forms = b->forms;
  std::vector<mpoa_call_stat> calls;
  rc = (b->forms & MPOA_FORM_HOST_SEED) ? 1 : run_seeded_dev(nb, off.data(), &calls);
  if (rc == 1) rc = run_seeded(nb, seq.data(), off.data(), &calls);
  if (!nb->res.empty()) {
    b->res[w] = nb->res[0];
    if (nb->desc[0].out_cap > 0 &&
        hipMemcpy(b->d_out.p + wd.out0, nb->d_out.p, (size_t)nb->desc[0].out_cap,
                  hipMemcpyDeviceToDevice) != hipSuccess)
      rc = MPOA_E_DEVICE;
    if (hipMemcpy(b->d_res.p + w, &b->res[w], sizeof(WinResult), hipMemcpyHostToDevice) !=
        hipSuccess)
      rc = MPOA_E_DEVICE;
  }
  b->launches += nb->launches;
  mpoa_batch_free(nb);
  return rc == MPOA_E_DEVICE || rc == MPOA_E_OOM ? rc : MPOA_OK;
}

// Capacities of a seeded window: graph as for any window; DP for the largest sub-window,
// which the seeding keeps to about min_w = consensus / 24 of the read between anchors, taken
// four times over, on rows of up to the whole graph (a window past it re-runs).
Estimate seeded_estimate(const mpoa_batch* b, int w) {
  const WinInfo& wi = b->info[w];
  Estimate e = estimate(wi, b->sc, true);
  const int64_t width = std::min<int64_t>((int64_t)(wi.qmax + 1) / 8 * 8 + 16,
                                          (int64_t)(wi.qmax / 6 + 2) / 8 * 8 + 64);
  e.dp = (int64_t)e.nodes * width;
  return e;
}

// Seeded (-S) windows of a batch, split like the non-seeded launches: windows with reads up
// to 1,500 bases on the 64-thread seeded tile (30 KB of LDS, several per CU), the rest on the
// 256-thread tile (one per CU), both launches at once on two streams. Windows that outgrow
// their estimate (graph, or a seeding past the small tile's LDS) run again on the 256-thread
// tile with worst-case graph capacities; those still past a capacity (seeding's own limits)
// run through the single-window form.
int execute_seeded(mpoa_batch* b, std::vector<Estimate>& est) {
  mpoa_ctx* ctx = b->ctx;
  const int n = b->n_windows;
  b->first_slots = 0;
  std::vector<WinResult> res(n);
  std::vector<int> again;
  int rc;
  if (!(b->forms & MPOA_FORM_HOST_SEED)) {
    for (int w = 0; w < n; w++) est[w] = seeded_estimate(b, w);
    b->small.wins.clear();
    b->wide.wins.clear();
    for (int w : b->order) (b->info[w].qmax <= 1500 ? b->small : b->wide).wins.push_back(w);
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    // what the bins hold already counts as free for them (provision reuses or replaces it)
    auto held = [](const Bin& x) {
      return x.d_int.n * sizeof(int) + x.d_byte.n + x.d_dp.n * sizeof(int);
    };
    const size_t avail = free_b + held(b->small) + held(b->wide);
    SeededLaunch ls, lw;
    hipEvent_t ev0;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventRecord(ev0, ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->stream3, ev0, 0));
    (void)hipEventDestroy(ev0);
    if (!b->wide.wins.empty() &&
        (rc = seeded_enqueue(b, est, b->wide, 256, avail / 4, ctx->stream3, lw)) != MPOA_OK)
      return rc;
    if (!b->small.wins.empty() &&
        (rc = seeded_enqueue(b, est, b->small, 64, avail / 2, ctx->stream, ls)) != MPOA_OK)
      return rc;
    float ms_s = 0, ms_w = 0;
    if ((rc = seeded_wait(ls, &ms_s)) != MPOA_OK || (rc = seeded_wait(lw, &ms_w)) != MPOA_OK)
      return rc;
    b->kernel_ms = std::max(ms_s, ms_w);
    HIPCHK(hipMemcpy(res.data(), b->d_res.p, n * sizeof(WinResult), hipMemcpyDeviceToHost));
    for (int w = 0; w < n; w++)
      if (res[w].status == MPOA_E_OVERFLOW) again.push_back(w);
    if (!again.empty()) {
      b->reruns += (int64_t)again.size();
      for (int w : again) est[w] = worst_case(b->info[w]);
      b->rerun.wins = again;
      HIPCHK(hipMemGetInfo(&free_b, &total_b));
      SeededLaunch lr;
      float ms = 0;
      if ((rc = seeded_enqueue(b, est, b->rerun, 256, (free_b + held(b->rerun)) / 4 * 3,
                               ctx->stream, lr)) != MPOA_OK ||
          (rc = seeded_wait(lr, &ms)) != MPOA_OK)
        return rc;
      b->kernel_ms += ms;
      HIPCHK(hipMemcpy(res.data(), b->d_res.p, n * sizeof(WinResult), hipMemcpyDeviceToHost));
      again.clear();
      for (int w = 0; w < n; w++)
        if (res[w].status == MPOA_E_OVERFLOW) again.push_back(w);
    }
  } else {
    for (int w = 0; w < n; w++) again.push_back(w);
  }
  b->res = res;
  for (int w : again)
    if ((rc = seeded_single(b, w)) != MPOA_OK) return rc;
  b->device_ms = b->kernel_ms;
  for (int w = 0; w < n; w++)
    if (b->res[w].status != MPOA_OK) return b->res[w].status;
  return MPOA_OK;
}

// The plain form: one launch per window, one workgroup, one window after another on one
// stream, in input order. Same kernel and capacities as the batched form.
int plain_pass(mpoa_batch* b, const std::vector<Estimate>& est, const std::vector<int>& wins,
               bool worst, float* ms_out);

int execute_plain(mpoa_batch* b, std::vector<Estimate>& est) {
  const int n = b->n_windows;
  std::vector<int> all(n);
  for (int w = 0; w < n; w++) all[w] = w;
  float ms = 0, rms = 0;
  int rc = plain_pass(b, est, all, false, &ms);
  if (rc != MPOA_OK) return rc;
  b->kernel_ms = ms;
  std::vector<WinResult> res(n);
  HIPCHK(hipMemcpy(res.data(), b->d_res.p, n * sizeof(WinResult), hipMemcpyDeviceToHost));
  std::vector<int> again;
  for (int w = 0; w < n; w++)
    if (res[w].status == MPOA_E_OVERFLOW) again.push_back(w);
  if (!again.empty()) {
    b->reruns += (int64_t)again.size();
    for (int w : again) est[w] = worst_case(b->info[w]);
    if ((rc = plain_pass(b, est, again, true, &rms)) != MPOA_OK) return rc;
    b->kernel_ms += rms;
    HIPCHK(hipMemcpy(res.data(), b->d_res.p, n * sizeof(WinResult), hipMemcpyDeviceToHost));
  }
  b->device_ms = b->kernel_ms;
  b->res = res;
  b->first_slots = 1;
  for (int w = 0; w < n; w++)
    if (res[w].status != MPOA_OK) return res[w].status;
  return MPOA_OK;
}

int plain_pass(mpoa_batch* b, const std::vector<Estimate>& est, const std::vector<int>& wins,
               bool worst, float* ms_out) {
  mpoa_ctx* ctx = b->ctx;
  const int n = b->n_windows;
  const bool keep_path = b->params.result != 0;
  int rc;
  size_t free_b = 0, total_b = 0;
  // Worst-case re-runs take the wide tile (the re-run bin of the batched form).
  if (worst) {
    b->rerun.nt = 256;
    b->rerun.cpl = 2;
    b->rerun.wins = wins;
  }
  for (Bin* bin : {&b->small, &b->large, &b->wide, &b->rerun}) {
    if (bin->wins.empty() || (worst != (bin == &b->rerun))) continue;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    if ((rc = provision(*bin, make_caps(est, b->info, bin->wins, keep_path), 1, free_b / 3)) !=
        MPOA_OK)
      return rc;
  }
  std::vector<int> ident(n);
  for (int w = 0; w < n; w++) ident[w] = w;
  DevBuf<int> d_ident, d_q;
  if ((rc = d_ident.alloc(n)) != MPOA_OK || (rc = d_q.alloc(n)) != MPOA_OK) return rc;
  HIPCHK(hipMemcpy(d_ident.p, ident.data(), n * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(d_q.p, 0, n * sizeof(int)));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, ctx->stream));
  for (int w : wins) {
    const Bin& bin = worst              ? b->rerun
                     : b->bin_of[w] == 1 ? b->large
                     : b->bin_of[w] == 3 ? b->wide
                                         : b->small;
    KernelArgs a{};
    a.forms = b->forms | (worst ? MPOA_FORM_WIDE : 0);
    a.scores = b->sc;
    a.caps = bin.caps;
    a.read_seq = b->d_seq.p;
    a.read_off = b->d_off.p;
    a.win = b->d_win.p;
    a.result = b->d_res.p;
    a.order = d_ident.p + w;
    a.n_jobs = 1;
    a.queue = d_q.p + w;
    a.int_pool = bin.d_int.p;
    a.byte_pool = bin.d_byte.p;
    a.dp_pool = bin.d_dp.p;
    a.out = b->d_out.p;
    a.call_stats = b->d_calls.p;
    a.call_off = b->d_call_off.p;
    if (mpoa_launch_windows(a, 1, bin.nt, bin.cpl, ctx->stream) != 0) {
      g_err = "kernel launch failed";
      return MPOA_E_DEVICE;
    }
    b->launches++;
  }
  HIPCHK(hipEventRecord(e1, ctx->stream));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(ms_out, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return MPOA_OK;
}

// A seeded (-S) window seeded on the device (SeedDev): one launch sketches every read, then per
// read the launch of its sub-windows and its commit, whose workgroup seeds the next read
// against the new consensus. Everything is enqueued at once and synchronised once. Returns 1
// when the window outgrew a capacity of this form (seeding buffers, sub-window count or pool,
// graph estimate; run_seeded below then runs it), else the window's status.
int run_seeded_dev(mpoa_batch* b, const int64_t* seq_off, std::vector<mpoa_call_stat>* calls) {
  using mpoa::SeedDev;
  mpoa_ctx* ctx = b->ctx;
  const WinInfo& wi = b->info[0];
  const int n = wi.n_reads;
  const int k = b->params.k, w = b->params.mm_w;
  if (n < 2 || k < 1 || k > 28 || w < 1 || w > 255) {
    ctx->last_seeding = 0;
    return 1;
  }
  const bool keep_path = b->params.result != 0;
  const Estimate e = estimate(wi, b->sc, true);
  SlotCaps gc{};
  gc.nodes = (std::max(e.nodes, 2) + 63) / 64 * 64;
  gc.edges = (std::max(e.edges, 4) + 63) / 64 * 64;
  gc.qmax = std::max(1, wi.qmax);
  gc.path = keep_path ? (int)wi.qtot : 0;
  gc.dp = 64;
  mpoa::finish_caps(gc);
  SlotCaps sc_caps = gc;
  sc_caps.edges = 64;
  sc_caps.path = 0;
  sc_caps.dp = 0;  // every sub-window takes its share of one pool (SegDesc::dp0)
  mpoa::finish_caps(sc_caps);
  // One read's sub-windows: a band row per rank between two anchors, about a min_w-th of
  // the read wide; dp_cap allows every rank an eighth of the longest read.
  const int64_t dp_cap =
      std::max<int64_t>(3LL * (16 << 20), 3LL * (gc.nodes + 64) * (gc.qmax / 8 + 64));
  const int cons_cap = gc.nodes;
  std::vector<int64_t> rm_off(n + 1, 0);
  for (int r = 0; r < n; r++)
    rm_off[r + 1] = rm_off[r] + (seq_off[r + 1] - seq_off[r]) + 2 * w + 8;
  DevBuf<WinState> d_state;
  DevBuf<int> g_int, g_dp, s_int, s_dp, i_buf;
  DevBuf<uint8_t> g_byte, s_byte;
  DevBuf<uint64_t> u_buf;
  DevBuf<int64_t> d_rm_off;
  DevBuf<SegDesc> d_segs;
  DevBuf<SegResult> d_segres;
  int rc;
  const size_t rm = (size_t)rm_off[n], hx = (size_t)wi.qtot + cons_cap;
  const size_t u_words = 4 * rm + hx + 4 * (size_t)cons_cap + 8 * (size_t)mpoa::SEED_ANCHORS;
  // ints: rm_cnt, n_seg, n_splice, read_graph (2 per read), splices, flag
  const size_t i_words = 5 * (size_t)n + 2 * (size_t)n * mpoa::SEED_SPLICE + 1;
  if ((rc = d_state.alloc(1)) != MPOA_OK || (rc = g_int.alloc(gc.ints_per_slot)) != MPOA_OK ||
      (rc = g_byte.alloc(gc.bytes_per_slot)) != MPOA_OK || (rc = g_dp.alloc(3 * 64)) != MPOA_OK ||
      (rc = s_int.alloc(mpoa::SEED_SEGS * sc_caps.ints_per_slot)) != MPOA_OK ||
      (rc = s_byte.alloc(mpoa::SEED_SEGS * sc_caps.bytes_per_slot)) != MPOA_OK ||
      (rc = s_dp.alloc((size_t)dp_cap)) != MPOA_OK || (rc = u_buf.alloc(u_words)) != MPOA_OK ||
      (rc = i_buf.alloc(i_words)) != MPOA_OK || (rc = d_rm_off.alloc(n + 1)) != MPOA_OK ||
      (rc = d_segs.alloc((size_t)n * mpoa::SEED_SEGS)) != MPOA_OK ||
      (rc = d_segres.alloc((size_t)n * mpoa::SEED_SEGS)) != MPOA_OK)
    return rc;
  HIPCHK(hipMemcpyAsync(d_rm_off.p, rm_off.data(), (n + 1) * sizeof(int64_t),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(i_buf.p, 0, i_words * sizeof(int), ctx->stream));
  SeedDev d{};
  uint64_t* u = u_buf.p;
  auto take_u = [&](size_t m) { uint64_t* p = u; u += m; return p; };
  d.rmx = take_u(rm); d.rmy = take_u(rm); d.rtx = take_u(rm); d.rty = take_u(rm);
  d.hx = take_u(hx);
  d.cmx = take_u(cons_cap); d.cmy = take_u(cons_cap);
  d.ctx = take_u(cons_cap); d.cty = take_u(cons_cap);
  d.ax = take_u(mpoa::SEED_ANCHORS); d.ay = take_u(mpoa::SEED_ANCHORS);
  d.bx = take_u(mpoa::SEED_ANCHORS); d.by = take_u(mpoa::SEED_ANCHORS);
  d.zx = take_u(mpoa::SEED_ANCHORS); d.zy = take_u(mpoa::SEED_ANCHORS);
  d.zsx = take_u(mpoa::SEED_ANCHORS); d.zsy = take_u(mpoa::SEED_ANCHORS);
  int* iw = i_buf.p;
  d.rm_cnt = iw; iw += n;
  d.n_seg = iw; iw += n;
  d.n_splice = iw; iw += n;
  d.read_graph = iw; iw += 2 * n;
  d.splice = iw; iw += 2 * (size_t)n * mpoa::SEED_SPLICE;
  d.flag = iw;
  d.rm_off = d_rm_off.p;
  d.hx_cons = wi.qtot;
  d.cons_cap = cons_cap;
  d.segs = d_segs.p;
  d.dp_cap = dp_cap;
  d.k = k;
  d.w = w;
  d.bw = 1000;  // main.cpp:65
  d.poa_w = b->params.poa_w;
  d.pen_gap = (float)(0.01 * k);  // minimizer.cpp:498
  d.pen_skip = (float)(0.0 * 0.01 * k);
  KernelArgs ga{};
  ga.forms = b->forms;
  ga.scores = b->sc;
  ga.caps = gc;
  ga.read_seq = b->d_seq.p;
  ga.read_off = b->d_off.p;
  ga.win = b->d_win.p;
  ga.result = b->d_res.p;
  ga.n_jobs = 1;
  ga.int_pool = g_int.p;
  ga.byte_pool = g_byte.p;
  ga.dp_pool = g_dp.p;
  ga.out = b->d_out.p;
  ga.gfa_lines = b->d_gfa.p;
  SeedArgs sa{};
  sa.graph = ga;
  sa.seg = ga;
  sa.seg.caps = sc_caps;
  sa.seg.int_pool = s_int.p;
  sa.seg.byte_pool = s_byte.p;
  sa.seg.dp_pool = s_dp.p;
  sa.n_reads = n;
  sa.state = d_state.p;
  sa.use_dev = 1;
  sa.dev = d;
  // The whole launch sequence (the copies and clears above stay outside) is captured as one
  // hipGraph and launched once, unless graph replay is off for the batch.
  const bool capture = b->use_graph;
  if (capture) HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const char* lerr = nullptr;
  if (mpoa_launch_seed_reads(sa, ctx->stream) != 0) lerr = "read sketch launch failed";
  int path0 = 0;
  for (int r = 0; r < n && !lerr; r++) {
    sa.read = r;
    sa.path0 = path0;
    sa.segs = d.segs + (size_t)r * mpoa::SEED_SEGS;
    sa.seg_res = d_segres.p + (size_t)r * mpoa::SEED_SEGS;
    sa.splice = d.splice + (size_t)r * 2 * mpoa::SEED_SPLICE;
    if (r > 0 && mpoa_launch_segments(sa, 128, 4, ctx->stream) != 0)
      lerr = "sub-window launch failed";
    else if (mpoa_launch_commit(sa, ctx->stream) != 0)
      lerr = "commit launch failed";
    b->launches += r > 0 ? 2 : 1;
    path0 += (int)(seq_off[r + 1] - seq_off[r]);
  }
  if (capture) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (!lerr && e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (!lerr && e == hipSuccess) e = hipGraphLaunch(exec, ctx->stream);
    if (!lerr && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (!lerr && e != hipSuccess) lerr = hipGetErrorString(e);
  }
  if (lerr) {
    g_err = lerr;
    return MPOA_E_DEVICE;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<int> iv(i_words);
  HIPCHK(hipMemcpy(iv.data(), i_buf.p, i_words * sizeof(int), hipMemcpyDeviceToHost));
  b->res.assign(1, WinResult{});
  HIPCHK(hipMemcpy(b->res.data(), b->d_res.p, sizeof(WinResult), hipMemcpyDeviceToHost));
  const int* h_nseg = iv.data() + n;
  const int* h_graph = iv.data() + 3 * n;
  if (iv[i_words - 1] != 0 || b->res[0].status == MPOA_E_OVERFLOW) {
    ctx->last_seeding = iv[i_words - 1] != 0 ? iv[i_words - 1] : mpoa::SEED_WHY_GRAPH;
    return 1;
  }
  ctx->last_seeding = 1;
  calls->clear();
  std::vector<SegResult> sr((size_t)n * mpoa::SEED_SEGS);
  HIPCHK(hipMemcpy(sr.data(), d_segres.p, sr.size() * sizeof(SegResult), hipMemcpyDeviceToHost));
  for (int r = 1; r < n; r++) {
    for (int s = 0; s < h_nseg[r]; s++) {
      const SegResult& g = sr[(size_t)r * mpoa::SEED_SEGS + s];
      mpoa_call_stat c{};
      c.read = r;
      c.rows = g.rows;
      c.qlen = g.qlen;
      c.score = g.score;
      c.cells = g.cells;
      c.edge_cells = g.edge_cells;
      c.logical_cells = g.logical;
      c.width = g.width;
      c.nodes = h_graph[2 * r];
      c.edges = h_graph[2 * r + 1];
      calls->push_back(c);
    }
  }
  return b->res[0].status;
}

// A seeded (-S) window, read by read (main.cpp:110-121 with alignment()'s -S branch,
// align.cpp:1763-1827). The host seeds each read against the consensus the device built
// after the previous read, splits the read at the chain's anchors, and launches one
// workgroup per sub-window, then the commit of the read. The graph stays on the device.
int run_seeded(mpoa_batch* b, const char* seq, const int64_t* seq_off,
               std::vector<mpoa_call_stat>* calls) {
  mpoa_ctx* ctx = b->ctx;
  const WinInfo& wi = b->info[0];
  const int n = wi.n_reads;
  const bool keep_path = b->params.result != 0;
  const int64_t base0 = seq_off[0];
  std::vector<uint8_t> codes((size_t)(seq_off[n] - base0) + 1);
  std::vector<int64_t> off(n + 1);
  for (int i = 0; i <= n; i++) off[i] = seq_off[i] - base0;
  for (int64_t j = 0; j < off[n]; j++) codes[j] = code_of_char((unsigned char)seq[base0 + j]);
  seed::Config cfg;
  cfg.k = b->params.k;
  cfg.w = b->params.mm_w;
  cfg.bw = 1000;  // main.cpp:65
  cfg.poa_w = b->params.poa_w;
  seed::Seeder sd;
  sd.init(cfg, n, codes.data(), off.data());

  DevBuf<WinState> d_state;
  DevBuf<SegDesc> d_segs;
  DevBuf<SegResult> d_segres;
  DevBuf<int> d_splice, d_cons;
  DevBuf<int> g_int, s_int, s_dp, g_dp;
  DevBuf<uint8_t> g_byte, s_byte;
  int rc;
  if ((rc = d_state.alloc(1)) != MPOA_OK) return rc;
  for (int attempt = 0; attempt < 2; attempt++) {
    calls->clear();
    const Estimate e = attempt == 0 ? estimate(wi, b->sc, true) : worst_case(wi);
    SlotCaps gc{};
    gc.nodes = (std::max(e.nodes, 2) + 63) / 64 * 64;
    gc.edges = (std::max(e.edges, 4) + 63) / 64 * 64;
    gc.qmax = std::max(1, wi.qmax);
    gc.path = keep_path ? (int)wi.qtot : 0;
    gc.dp = 64;
    mpoa::finish_caps(gc);
    if ((rc = g_int.alloc((size_t)gc.ints_per_slot)) != MPOA_OK ||
        (rc = g_byte.alloc((size_t)gc.bytes_per_slot)) != MPOA_OK ||
        (rc = g_dp.alloc(3 * 64)) != MPOA_OK || (rc = d_cons.alloc(3 * (size_t)gc.nodes)) != MPOA_OK)
      return rc;
    KernelArgs ga{};
    ga.forms = b->forms;
    ga.scores = b->sc;
    ga.caps = gc;
    ga.read_seq = b->d_seq.p;
    ga.read_off = b->d_off.p;
    ga.win = b->d_win.p;
    ga.result = b->d_res.p;
    ga.n_jobs = 1;
    ga.int_pool = g_int.p;
    ga.byte_pool = g_byte.p;
    ga.dp_pool = g_dp.p;
    ga.out = b->d_out.p;
    ga.gfa_lines = b->d_gfa.p;
    SlotCaps sc_caps{};
    size_t seg_slots = 0;
    int64_t seg_dp = 0;
    int path0 = 0, prev_seg = 0;
    WinState st{};
    bool failed = false;
    std::vector<int> cons;
    std::vector<uint8_t> cons_code;
    auto collect = [&](int read, int nseg) -> int {
      if (nseg == 0) return MPOA_OK;
      std::vector<SegResult> sr(nseg);
      HIPCHK(hipMemcpy(sr.data(), d_segres.p, nseg * sizeof(SegResult), hipMemcpyDeviceToHost));
      for (const SegResult& g : sr) {
        mpoa_call_stat c{};
        c.read = read;
        c.rows = g.rows;
        c.qlen = g.qlen;
        c.score = g.score;
        c.cells = g.cells;
        c.edge_cells = g.edge_cells;
        c.logical_cells = g.logical;
        c.width = g.width;
        c.nodes = st.n_nodes;
        c.edges = st.n_in_e;
        calls->push_back(c);
      }
      return MPOA_OK;
    };
    for (int r = 0; r < n && !failed; r++) {
      const int qlen = (int)(off[r + 1] - off[r]);
      std::vector<SegDesc> segs;
      std::vector<int> splice;
      std::vector<int64_t> need;
      if (r > 0) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(&st, d_state.p, sizeof(WinState), hipMemcpyDeviceToHost));
        if ((rc = collect(r - 1, prev_seg)) != MPOA_OK) return rc;
        if (st.status != MPOA_OK) {
          failed = true;
          break;
        }
        if (st.n_nodes > 2) {
          const int L = st.cons_len;
          cons.resize(3 * (size_t)std::max(L, 1));
          if (L > 0)
            HIPCHK(hipMemcpy(cons.data(), d_cons.p, 3 * (size_t)L * sizeof(int),
                             hipMemcpyDeviceToHost));
          cons_code.resize(std::max(L, 1));
          for (int p = 0; p < L; p++) cons_code[p] = (uint8_t)cons[3 * p + 2];
          const std::vector<seed::Pair> an = sd.anchors(r, qlen, cons_code.data(), L);
          const int sink_rank = st.n_nodes - 1;
          int beg_id = 0, beg_rank = 0, beg_q = 0, j = 0, res_pos = 0;
          // <= 0: a sub-window that ends before it begins, in ranks (minipoa writes to an
          // empty vector) or in the read by more than one base (anchors overlapping there,
          // where minipoa's DP reads past its query). Length -1 is a DP call that adds no res
          // entry (the device's null_call_score).
          auto add_seg = [&](int end_id, int end_rank, int sub_len) {
            if (sub_len < -1) return 0;
            SegDesc d{};
            d.beg_id = beg_id;
            d.end_id = end_id;
            d.q0 = j;
            d.qlen = sub_len;
            d.res0 = res_pos;
            segs.push_back(d);
            const int rows = end_rank - beg_rank + 1;
            need.push_back((int64_t)std::max(rows, 1) * ((sub_len + 1) / 8 + 2) * 8);
            return rows;
          };
          for (const seed::Pair& a : an) {
            const int q_span = (int)(a.y >> 32 & 0xff);
            const int end_tpos = (int)(uint32_t)a.x - q_span + 1;
            const int end_qpos = (int)(uint32_t)a.y - q_span + 1;
            if (end_tpos < 0 || end_tpos + q_span > L) {
              g_err = "anchor outside the consensus";
              return MPOA_E_ARG;
            }
            const int sub_len = end_qpos - beg_q;
            const int end_id = cons[3 * end_tpos], end_rank = cons[3 * end_tpos + 1];
            if (add_seg(end_id, end_rank, sub_len) <= 0) return MPOA_E_EMPTY_WINDOW;
            res_pos += sub_len + 1;
            j += sub_len;
            int last_id = end_id, last_rank = end_rank;
            for (int k = 0; k < q_span; k++, j++) {
              last_id = cons[3 * (end_tpos + k)];
              last_rank = cons[3 * (end_tpos + k) + 1];
              if (k != q_span - 1) {
                splice.push_back(res_pos++);
                splice.push_back(last_id);
              }
            }
            beg_id = last_id;
            beg_rank = last_rank;
            beg_q = j;
          }
          if (add_seg(1, sink_rank, qlen - beg_q) <= 0) return MPOA_E_EMPTY_WINDOW;
        }
      }
      const int nseg = (int)segs.size();
      if (nseg > 0) {
        const int64_t dp_need = (*std::max_element(need.begin(), need.end()) + 63) / 64 * 64;
        if ((size_t)nseg > seg_slots || dp_need > seg_dp) {
          sc_caps = gc;
          sc_caps.edges = 64;
          sc_caps.path = 0;
          sc_caps.dp = std::max<int64_t>(dp_need, seg_dp);
          mpoa::finish_caps(sc_caps);
          const size_t slots = std::max<size_t>((size_t)nseg, seg_slots);
          HIPCHK(hipStreamSynchronize(ctx->stream));
          if ((rc = s_int.alloc(slots * sc_caps.ints_per_slot)) != MPOA_OK ||
              (rc = s_byte.alloc(slots * sc_caps.bytes_per_slot)) != MPOA_OK ||
              (rc = s_dp.alloc(slots * 3 * (size_t)sc_caps.dp)) != MPOA_OK ||
              (rc = d_segres.alloc(slots)) != MPOA_OK)
            return rc;
          seg_slots = slots;
          seg_dp = sc_caps.dp;
        }
        if (!d_segs.p || d_segs.n < (size_t)nseg)
          if ((rc = d_segs.alloc(std::max<size_t>(nseg, 16))) != MPOA_OK) return rc;
        HIPCHK(hipMemcpyAsync(d_segs.p, segs.data(), nseg * sizeof(SegDesc),
                              hipMemcpyHostToDevice, ctx->stream));
        if (!splice.empty()) {
          if (!d_splice.p || d_splice.n < splice.size())
            if ((rc = d_splice.alloc(std::max<size_t>(splice.size(), 64))) != MPOA_OK) return rc;
          HIPCHK(hipMemcpyAsync(d_splice.p, splice.data(), splice.size() * sizeof(int),
                                hipMemcpyHostToDevice, ctx->stream));
        }
      }
      SeedArgs sa{};
      sa.graph = ga;
      sa.seg = ga;
      sa.seg.caps = sc_caps;
      sa.seg.int_pool = s_int.p;
      sa.seg.byte_pool = s_byte.p;
      sa.seg.dp_pool = s_dp.p;
      sa.segs = d_segs.p;
      sa.seg_res = d_segres.p;
      sa.n_seg = nseg;
      sa.splice = d_splice.p;
      sa.n_splice = (int)splice.size() / 2;
      sa.read = r;
      sa.n_reads = n;
      sa.path0 = path0;
      sa.state = d_state.p;
      sa.cons_out = d_cons.p;
      sa.cons_cap = gc.nodes;
      if (nseg > 0 && mpoa_launch_segments(sa, 128, 4, ctx->stream) != 0) {
        g_err = "sub-window launch failed";
        return MPOA_E_DEVICE;
      }
      if (mpoa_launch_commit(sa, ctx->stream) != 0) {
        g_err = "commit launch failed";
        return MPOA_E_DEVICE;
      }
      b->launches += nseg > 0 ? 2 : 1;
      prev_seg = nseg;
      path0 += qlen;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&st, d_state.p, sizeof(WinState), hipMemcpyDeviceToHost));
    if (!failed && (rc = collect(n - 1, prev_seg)) != MPOA_OK) return rc;
    b->res.assign(1, WinResult{});
    HIPCHK(hipMemcpy(b->res.data(), b->d_res.p, sizeof(WinResult), hipMemcpyDeviceToHost));
    if (b->res[0].status == MPOA_E_OVERFLOW && attempt == 0) {
      b->reruns++;
      continue;
    }
    return b->res[0].status;
  }
  return b->res.empty() ? MPOA_E_DEVICE : b->res[0].status;
}

}  // namespace

namespace mpoa {
namespace host {

int window_text(int result, const WinResult& r, const uint8_t* d_out, const int64_t* d_gfa,
                int n_reads, const char* name, const int64_t* name_off, const char* comment,
                const int64_t* comment_off, const int* row_of, std::string* out) {
  std::string& text = *out;
  text.clear();
  if (result == 0) {
    std::vector<char> cons((size_t)std::max(r.out_len, 1));
    if (r.out_len > 0 &&
        hipMemcpy(cons.data(), d_out, (size_t)r.out_len, hipMemcpyDeviceToHost) != hipSuccess) {
      return MPOA_E_DEVICE;
    }
    text = ">consensus sequence\n";
    text.append(cons.data(), (size_t)r.out_len);
    text.push_back('\n');
  } else if (result == 2) {
    // output_gfa (src/graph.cpp:358-433): the device names path lines by read index; the
    // read names replace those where given.
    std::vector<char> g((size_t)std::max(r.out_len, 1));
    std::vector<int64_t> at((size_t)n_reads + 1);
    if (r.out_len > 0 &&
        (hipMemcpy(g.data(), d_out, (size_t)r.out_len, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(at.data(), d_gfa, at.size() * sizeof(int64_t), hipMemcpyDeviceToHost) !=
             hipSuccess)) {
      return MPOA_E_DEVICE;
    }
    if (r.out_len > 0) {
      text.assign(g.data(), (size_t)at[0]);
      for (int i = 0; i < n_reads; i++) {
        const char* line = g.data() + at[i];
        const int64_t len = at[i + 1] - at[i];
        const char* tab = (const char*)std::memchr(line + 2, '\t', (size_t)(len - 2));
        std::string nm;
        if (name && name_off) nm.assign(name + name_off[i], name + name_off[i + 1]);
        text += "P\t";
        text += nm.empty() ? std::to_string(i + 1) : nm;
        if (tab) text.append(tab, (size_t)(line + len - tab));
      }
    }
  } else {
    // output_rc_msa (src/graph.cpp:237-251); result 3 adds the consensus row
    const int L = r.out_len;
    const int n_rows = n_reads + (result == 3 ? 1 : 0);
    std::vector<char> rows((size_t)std::max<int64_t>((int64_t)n_rows * L, 1));
    if ((int64_t)n_rows * L > 0 &&
        hipMemcpy(rows.data(), d_out, (size_t)n_rows * L, hipMemcpyDeviceToHost) !=
            hipSuccess) {
      return MPOA_E_DEVICE;
    }
    for (int i = 0; i < n_reads; i++) {
      std::string nm, cm;
      if (name && name_off) nm.assign(name + name_off[i], name + name_off[i + 1]);
      if (comment && comment_off) cm.assign(comment + comment_off[i], comment + comment_off[i + 1]);
      if (nm.empty()) text += ">" + std::to_string(i + 1) + "\n";
      else text += ">" + nm + " " + cm + "\n";
      // read i prints the row of the read added at position row_of[i] (graph.cpp:249)
      text.append(rows.data() + (int64_t)(row_of ? row_of[i] : i) * L, (size_t)L);
      text.push_back('\n');
    }
    if (result == 3) {
      text += ">consensus sequence\n";
      text.append(rows.data() + (int64_t)n_reads * L, (size_t)L);
      text.push_back('\n');
    }
  }
  return MPOA_OK;
}

}  // namespace host
}  // namespace mpoa

extern "C" {

const char* mpoa_strerror(int rc) {
  switch (rc) {
    case MPOA_OK: return "ok";
    case MPOA_E_ARG: return g_err.empty() ? "invalid argument" : g_err.c_str();
    case MPOA_E_BAND: return "backtrack left the band";
    case MPOA_E_BACKTRACK: return "backtrack error";
    case MPOA_E_OUT_CAP: return "output buffer too small";
    case MPOA_E_EMPTY_WINDOW: return "empty seeding sub-window";
    case MPOA_E_OOM: return g_err.empty() ? "device out of memory" : g_err.c_str();
    case MPOA_E_DEVICE: return g_err.empty() ? "HIP device error" : g_err.c_str();
    case MPOA_E_OVERFLOW: return "arena overflow";
    case MPOA_E_SCRATCH:
      return g_err.empty() ? "window exceeds the allocation-free scratch" : g_err.c_str();
    default: return "unknown error";
  }
}

int mpoa_create(int device, mpoa_ctx** out) {
  if (!out) return MPOA_E_ARG;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= device || device < 0) {
    g_err = "no HIP device available (the library has no CPU path)";
    return MPOA_E_DEVICE;
  }
  HIPCHK(hipSetDevice(device));
  mpoa_ctx* c = new mpoa_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreate(&c->stream) != hipSuccess || hipStreamCreate(&c->stream2) != hipSuccess ||
      hipStreamCreate(&c->stream3) != hipSuccess) {
    delete c;
    g_err = "device initialisation failed";
    return MPOA_E_DEVICE;
  }
  void* scr = nullptr;
  if (mpoa_scratch_address(device, &scr) != 0) {
    (void)hipStreamDestroy(c->stream);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream3);
    delete c;
    g_err = "device scratch lookup failed";
    return MPOA_E_DEVICE;
  }
  c->scratch = static_cast<uint8_t*>(scr);
  c->n_cu = prop.multiProcessorCount;
  c->mem_total = prop.totalGlobalMem;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
      khz > 0)
    c->wall_khz = khz;
  *out = c;
  return MPOA_OK;
}

void mpoa_destroy(mpoa_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->stream3) (void)hipStreamDestroy(ctx->stream3);
  delete ctx;
}

int mpoa_batch_prepare(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_windows,
                       const int32_t* win_off, const char* seq, const int64_t* seq_off,
                       const double* size_hint, mpoa_batch** out) {
  if (!ctx || !params || !out || n_windows < 0 || !win_off || !seq_off || (!seq && n_windows))
    return MPOA_E_ARG;
  *out = nullptr;
  if (params->result < 0 || params->result > 3) {
    g_err = "result must be 0 (consensus), 1 (MSA), 2 (GFA) or 3 (MSA with a consensus row)";
    return MPOA_E_ARG;
  }
  if (params->progressive && params->result == 2) {
    g_err = "progressive order (-p) with GFA output: minipoa's GFA writer indexes its read "
            "paths past their ends then (graph.cpp:369,423)";
    return MPOA_E_ARG;
  }
  HIPCHK(hipSetDevice(ctx->device));
  mpoa_batch* b = new mpoa_batch();
  b->ctx = ctx;
  b->forms = ctx->forms;
  b->params = *params;
  if (params->progressive && n_windows > 0) {
    // each window's reads in minipoa's guide-tree order (minimizer.cpp:195-386); the device
    // sees them in that order, the outputs map rows back
    const int r0 = win_off[0], r1 = win_off[n_windows];
    b->perm_off.assign(1, 0);
    b->row_of.assign((size_t)(r1 - r0), 0);
    seed::Config cfg;
    cfg.k = params->k;
    cfg.w = params->mm_w;
    for (int w = 0; w < n_windows; w++) {
      const int a = win_off[w], n = win_off[w + 1] - win_off[w];
      std::vector<int64_t> off((size_t)n + 1, 0);
      for (int i = 0; i < n; i++) off[i + 1] = off[i] + (seq_off[a + i + 1] - seq_off[a + i]);
      std::vector<uint8_t> codes((size_t)off[n] + 1);
      for (int64_t j = 0; j < off[n]; j++)
        codes[j] = code_of_char((unsigned char)seq[seq_off[a] + j]);
      const std::vector<int> ord = seed::guide_order(cfg, n, codes.data(), off.data());
      for (int i = 0; i < n; i++) {
        const int rid = a + ord[i];
        b->perm_seq.append(seq + seq_off[rid], (size_t)(seq_off[rid + 1] - seq_off[rid]));
        b->perm_off.push_back((int64_t)b->perm_seq.size());
        b->row_of[(size_t)(a - r0 + ord[i])] = i;
      }
    }
    seq = b->perm_seq.data();
    std::vector<int64_t> shifted((size_t)r1 + 1, 0);
    for (int r = r0; r <= r1; r++) shifted[r] = b->perm_off[r - r0];
    b->perm_off.swap(shifted);
    seq_off = b->perm_off.data();
  }
  b->sc = scores_of(*params);
  b->n_windows = n_windows;
  b->info.resize(n_windows);
  b->desc.resize(n_windows);
  b->cost.resize(n_windows);
  const int r_base = win_off[0];
  int64_t out0 = 0;
  for (int w = 0; w < n_windows; w++) {
    WinInfo& wi = b->info[w];
    wi.n_reads = win_off[w + 1] - win_off[w];
    wi.qmax = 0;
    wi.qtot = 0;
    for (int r = win_off[w]; r < win_off[w + 1]; r++) {
      const int64_t l = seq_off[r + 1] - seq_off[r];
      wi.qmax = (int)std::max<int64_t>(wi.qmax, l);
      wi.qtot += l;
    }
    WinDesc& d = b->desc[w];
    d.read0 = win_off[w] - r_base;
    d.n_reads = wi.n_reads;
    d.out0 = out0;
    int64_t cap;
    if (params->result == 1 || params->result == 3) {
      cap = (int64_t)(wi.n_reads + 1) * (2 + wi.qtot);
    } else if (params->result == 2) {
      // output_gfa bound: nodes <= 2 + sum q, in-edges <= sum q + reads, d-digit ids
      int d = 1;
      for (int64_t v = 2 + wi.qtot; v >= 10; v /= 10) d++;
      cap = (2 + wi.qtot) * (6 + d) + (wi.qtot + wi.n_reads) * (13 + 2 * d) +
            wi.qtot * (d + 2) + (int64_t)wi.n_reads * (16 + d) + 64;
    } else {
      cap = 2 * (int64_t)wi.qmax + 64;
    }
    cap = std::min<int64_t>(cap, INT32_MAX / 2);
    d.out_cap = (int)cap;
    out0 += (cap + 15) / 16 * 16;
    const Estimate e = estimate(wi, b->sc, params->ab_band != 0 && params->band_f > 0);
    b->cost[w] = size_hint ? size_hint[w] : (double)wi.n_reads * (double)e.dp;
  }
  b->out_total = out0;
  b->order.resize(n_windows);
  for (int w = 0; w < n_windows; w++) b->order[w] = w;
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return b->cost[x] > b->cost[y]; });
  assign_bins(b);
  int rc;
  if (n_windows > 0 && (rc = upload(b, n_windows, win_off, seq, seq_off)) != MPOA_OK) {
    delete b;
    return rc;
  }
  if ((rc = b->d_win.alloc(std::max(n_windows, 1))) != MPOA_OK ||
      (rc = b->d_res.alloc(std::max(n_windows, 1))) != MPOA_OK ||
      (rc = b->d_out.alloc((size_t)std::max<int64_t>(out0, 16))) != MPOA_OK) {
    delete b;
    return rc;
  }
  if (n_windows > 0 &&
      hipMemcpy(b->d_win.p, b->desc.data(), n_windows * sizeof(WinDesc), hipMemcpyHostToDevice) !=
          hipSuccess) {
    delete b;
    return MPOA_E_DEVICE;
  }
  *out = b;
  return MPOA_OK;
}

int mpoa_batch_execute(mpoa_batch* b) {
  if (!b) return MPOA_E_ARG;
  mpoa_ctx* ctx = b->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  const int n = b->n_windows;
  b->res.assign(n, WinResult{});
  b->reruns = 0;
  b->kernel_ms = 0;
  b->launches = 0;
  b->graph_launches = 0;
  b->lstats.clear();
  if (n == 0) return MPOA_OK;
  std::vector<Pending> pend;
  const bool keep_path = b->params.result != 0;
  const bool full = b->params.ab_band != 0 && b->params.band_f > 0;
  std::vector<Estimate> est(n);
  if (b->params.seeding) return execute_seeded(b, est);
  for (int w = 0; w < n; w++) est[w] = capacity_of(b, w, full);
  if (b->mode == MPOA_MODE_PLAIN) return execute_plain(b, est);

  // Device memory: the large-window bin first (few slots), the small-window bin takes the rest.
  size_t free_b = 0, total_b = 0;
  int rc;
  hipEvent_t ev0, ev1, ev_large;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  HIPCHK(hipEventCreate(&ev_large));
  if (!b->wide.wins.empty()) {
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const SlotCaps c = make_caps(est, b->info, b->wide.wins, keep_path);
    if ((rc = provision(b->wide, c, std::max(1, ctx->n_cu), free_b / 4)) != MPOA_OK) return rc;
  }
  if (!b->large.wins.empty()) {
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const SlotCaps c = make_caps(est, b->info, b->large.wins, keep_path);
    if ((rc = provision(b->large, c, std::max(1, ctx->n_cu), free_b / 3)) != MPOA_OK) return rc;
  }
  if (!b->small.wins.empty()) {
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const SlotCaps c = make_caps(est, b->info, b->small.wins, keep_path);
    const int per_cu = mpoa_tile_occupancy(b->small.nt, b->small.cpl);
    const int max_slots = std::max(1, ctx->n_cu * std::max(1, per_cu));
    if ((rc = provision(b->small, c, max_slots, free_b / 4 * 3)) != MPOA_OK) return rc;
  }
  b->first_slots = b->small.used_slots + b->large.used_slots + b->wide.used_slots;
  if (b->use_graph) {
    HIPCHK(hipEventRecord(ev0, ctx->stream));
    if ((rc = graph_pass(b, pend)) != MPOA_OK) return rc;
  } else {
  HIPCHK(hipEventRecord(ev0, ctx->stream));
  HIPCHK(hipStreamWaitEvent(ctx->stream2, ev0, 0));
  HIPCHK(hipStreamWaitEvent(ctx->stream3, ev0, 0));
  // Large windows first so that they hold their CUs from the start: the small-window grid
  // (persistent, more workgroups than fit) would otherwise take every CU's LDS and keep the
  // large windows waiting until its queue drained. The host gives the large launch a short
  // head start before enqueueing the small one.
  if ((rc = timed_enqueue(b, b->wide, ctx->stream3, 3, pend)) != MPOA_OK) return rc;
  if ((rc = timed_enqueue(b, b->large, ctx->stream2, 1, pend)) != MPOA_OK) return rc;
  if ((!b->large.wins.empty() || !b->wide.wins.empty()) && !b->small.wins.empty()) {
    HIPCHK(hipStreamQuery(ctx->stream2) == hipErrorNotReady ? hipSuccess
                                                              : hipStreamQuery(ctx->stream2));
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
  }
  if ((rc = timed_enqueue(b, b->small, ctx->stream, 0, pend)) != MPOA_OK) return rc;
  HIPCHK(hipEventRecord(ev_large, ctx->stream2));
  HIPCHK(hipStreamWaitEvent(ctx->stream, ev_large, 0));
  HIPCHK(hipEventRecord(ev_large, ctx->stream3));
  HIPCHK(hipStreamWaitEvent(ctx->stream, ev_large, 0));
  }
  HIPCHK(hipEventRecord(ev1, ctx->stream));
  HIPCHK(hipEventSynchronize(ev1));
  float kms = 0;
  HIPCHK(hipEventElapsedTime(&kms, ev0, ev1));
  b->kernel_ms = kms;
  std::vector<WinResult> res(n);
  HIPCHK(hipMemcpy(res.data(), b->d_res.p, n * sizeof(WinResult), hipMemcpyDeviceToHost));
  if ((rc = close_pending(b, pend, res)) != MPOA_OK) return rc;
  // Windows that overflowed an estimated capacity: re-run with worst-case capacities.
  std::vector<int> wins;
  for (int w = 0; w < n; w++)
    if (res[w].status == MPOA_E_OVERFLOW) wins.push_back(w);
  for (int pass = 0; pass < 2 && !wins.empty(); pass++) {
    b->reruns += (int64_t)wins.size();
    for (int w : wins) est[w] = worst_case(b->info[w]);
    Bin& rb = b->rerun;
    rb.nt = 256;
    rb.cpl = 2;
    rb.wins = wins;
    for (int w : wins) b->bin_of[w] = 2;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    if ((rc = provision(rb, make_caps(est, b->info, wins, keep_path), ctx->n_cu,
                        free_b / 4 * 3)) != MPOA_OK)
      return rc;
    hipEvent_t r0, r1;
    HIPCHK(hipEventCreate(&r0));
    HIPCHK(hipEventCreate(&r1));
    HIPCHK(hipEventRecord(r0, ctx->stream));
    if ((rc = timed_enqueue(b, rb, ctx->stream, 2, pend)) != MPOA_OK) return rc;
    HIPCHK(hipEventRecord(r1, ctx->stream));
    HIPCHK(hipEventSynchronize(r1));
    float rms = 0;
    HIPCHK(hipEventElapsedTime(&rms, r0, r1));
    b->kernel_ms += rms;
    (void)hipEventDestroy(r0);
    (void)hipEventDestroy(r1);
    HIPCHK(hipMemcpy(res.data(), b->d_res.p, n * sizeof(WinResult), hipMemcpyDeviceToHost));
    if ((rc = close_pending(b, pend, res)) != MPOA_OK) return rc;
    std::vector<int> again;
    for (int w : wins)
      if (res[w].status == MPOA_E_OVERFLOW) again.push_back(w);
    wins = again;
  }
  float ms = 0;
  HIPCHK(hipEventRecord(ev1, ctx->stream));
  HIPCHK(hipEventSynchronize(ev1));
  HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  (void)hipEventDestroy(ev_large);
  b->device_ms = ms;
  b->res = res;
  for (int w = 0; w < n; w++)
    if (res[w].status != MPOA_OK) return res[w].status;
  return MPOA_OK;
}

int mpoa_batch_results(mpoa_batch* b, char* out, int64_t out_cap, int64_t* out_off,
                       mpoa_batch_stats* stats) {
  if (!b) return MPOA_E_ARG;
  const int n = b->n_windows;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->windows = n;
    for (int w = 0; w < n && w < (int)b->res.size(); w++) {
      stats->cells += b->res[w].cells;
      stats->edge_cells += b->res[w].edge_cells;
      stats->logical_cells += b->res[w].logical;
      stats->narrow_cells += b->res[w].narrow_cells;
      stats->dp_calls += std::max(0, b->info[w].n_reads - 1);
    }
    stats->reruns = b->reruns;
    stats->device_ms = b->device_ms;
    stats->kernel_ms = b->kernel_ms;
    stats->launches = b->launches;
    stats->slots = b->first_slots;
    stats->graph_launches = b->graph_launches;
    stats->graph_builds = b->graph_builds;
    stats->pool_bytes = 0;
    for (const Bin* bin : {&b->small, &b->large, &b->wide, &b->rerun})
      stats->pool_bytes += (int64_t)(bin->d_int.n * sizeof(int) + bin->d_byte.n +
                                     bin->d_dp.n * sizeof(int));
  }
  if (!out_off) return MPOA_OK;
  std::vector<uint8_t> raw((size_t)std::max<int64_t>(b->out_total, 1));
  if (b->out_total > 0)
    HIPCHK(hipMemcpy(raw.data(), b->d_out.p, (size_t)b->out_total, hipMemcpyDeviceToHost));
  static const char kHead[] = ">consensus sequence\n";
  int64_t at = 0;
  out_off[0] = 0;
  for (int w = 0; w < n; w++) {
    const WinResult& r = b->res[w];
    const uint8_t* src = raw.data() + b->desc[w].out0;
    std::string text;
    if (b->params.result == 0) {
      text.assign(kHead);
      text.append((const char*)src, (size_t)std::max(0, r.out_len));
      text.push_back('\n');
    } else if (b->params.result == 2) {
      text.assign((const char*)src, (size_t)std::max(0, r.out_len));
    } else {
      const int L = std::max(0, r.out_len);
      const int* row_of = b->row_of.empty() ? nullptr : b->row_of.data() + b->desc[w].read0;
      for (int i = 0; i < b->info[w].n_reads; i++) {
        text += ">" + std::to_string(i + 1) + "\n";
        text.append((const char*)src + (int64_t)(row_of ? row_of[i] : i) * L, (size_t)L);
        text.push_back('\n');
      }
      if (b->params.result == 3) {
        text += kHead;
        text.append((const char*)src + (int64_t)b->info[w].n_reads * L, (size_t)L);
        text.push_back('\n');
      }
    }
    if (out && at + (int64_t)text.size() <= out_cap) std::memcpy(out + at, text.data(), text.size());
    at += (int64_t)text.size();
    out_off[w + 1] = at;
  }
  return out && at > out_cap ? MPOA_E_OUT_CAP : MPOA_OK;
}

void mpoa_batch_free(mpoa_batch* b) { delete b; }

int mpoa_batch_window_stats(mpoa_batch* b, mpoa_window_stat* out, int32_t cap, double* tick_hz) {
  if (!b) return MPOA_E_ARG;
  if (tick_hz) {
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, b->ctx->device) !=
        hipSuccess)
      rate_khz = 100000;
    *tick_hz = rate_khz * 1000.0;
  }
  if (!out) return MPOA_OK;
  for (int w = 0; w < b->n_windows && w < cap && w < (int)b->res.size(); w++) {
    const WinResult& r = b->res[w];
    mpoa_window_stat& o = out[w];
    o.status = r.status;
    o.nodes = r.nodes;
    o.edges = r.edges;
    o.launch = b->bin_of.empty() ? 0 : b->bin_of[w];
    o.t_begin = r.t_begin;
    o.t_end = r.t_end;
    o.cells = r.cells;
    o.logical_cells = r.logical;
    o.narrow_cells = r.narrow_cells;
    o.edge_cells = r.edge_cells;
    for (int k = 0; k < 8; k++) o.ticks[k] = r.ticks[k];
  }
  return MPOA_OK;
}

int mpoa_batch_set_graph(mpoa_batch* b, int32_t enable) {
  if (!b) return MPOA_E_ARG;
  b->use_graph = enable != 0;
  return MPOA_OK;
}

int mpoa_batch_set_mode(mpoa_batch* b, int32_t mode) {
  if (!b || (mode != MPOA_MODE_BATCHED && mode != MPOA_MODE_PLAIN)) return MPOA_E_ARG;
  b->mode = mode;
  return MPOA_OK;
}

constexpr int kAllForms = MPOA_FORM_SORT_HBM | MPOA_FORM_SORT_LANE | MPOA_FORM_BT_LANE |
                          MPOA_FORM_ADD_LANE | MPOA_FORM_CONS_ROUNDS | MPOA_FORM_WIDE |
                          MPOA_FORM_HOST_SEED | MPOA_FORM_BT_RANKS;

int mpoa_batch_set_forms(mpoa_batch* b, int32_t forms) {
  if (!b || (forms & ~kAllForms) != 0) return MPOA_E_ARG;
  b->forms = forms;
  return MPOA_OK;
}

int mpoa_last_seeding(const mpoa_ctx* ctx) { return ctx ? ctx->last_seeding : -1; }

int mpoa_set_forms(mpoa_ctx* ctx, int32_t forms) {
  if (!ctx || (forms & ~kAllForms) != 0) return MPOA_E_ARG;
  ctx->forms = forms;
  ctx->held_key.clear();
  ctx->held_text.clear();
  ctx->held_calls.clear();
  return MPOA_OK;
}

int mpoa_batch_set_predictions(mpoa_batch* b, const float* pred, int32_t stride) {
  if (!b || stride < 4 || (b->n_windows > 0 && !pred)) return MPOA_E_ARG;
  const int n = b->n_windows;
  b->pred_cells.assign(n, 0.f);
  b->pred_width.assign(n, 0.f);
  b->pred_nodes.assign(n, 0.f);
  for (int w = 0; w < n; w++) {
    const float* p = pred + (int64_t)w * stride;
    if (!(p[0] >= 0.f) || !(p[1] >= 0.f) || !(p[2] >= 0.f) || !(p[3] >= 0.f)) {
      g_err = "predictions must be finite and non-negative";
      return MPOA_E_ARG;
    }
    b->cost[w] = p[0];
    b->pred_cells[w] = p[1];
    b->pred_width[w] = p[2];
    b->pred_nodes[w] = p[3];
  }
  for (int w = 0; w < n; w++) b->order[w] = w;
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return b->cost[x] > b->cost[y]; });
  assign_bins(b);
  return MPOA_OK;
}

int mpoa_batch_set_traces(mpoa_batch* b, int32_t enable) {
  if (!b) return MPOA_E_ARG;
  if (!enable) {
    b->d_calls.release();
    b->d_call_off.release();
    b->call_off.clear();
    return MPOA_OK;
  }
  const int n = b->n_windows;
  b->call_off.assign(n + 1, 0);
  for (int w = 0; w < n; w++) b->call_off[w + 1] = b->call_off[w] + std::max(0, b->info[w].n_reads - 1);
  int rc;
  if ((rc = b->d_calls.alloc((size_t)std::max<int64_t>(b->call_off[n], 1))) != MPOA_OK) return rc;
  if ((rc = b->d_call_off.alloc((size_t)n + 1)) != MPOA_OK) return rc;
  HIPCHK(hipMemcpy(b->d_call_off.p, b->call_off.data(), (n + 1) * sizeof(int64_t),
                   hipMemcpyHostToDevice));
  return MPOA_OK;
}

int mpoa_batch_traces(mpoa_batch* b, mpoa_call_stat* out, int64_t cap, int64_t* n) {
  if (!b || (cap > 0 && !out)) return MPOA_E_ARG;
  const int64_t total = b->call_off.empty() ? 0 : b->call_off.back();
  if (n) *n = total;
  if (!out || total == 0) return MPOA_OK;
  std::vector<CallStat> cs((size_t)total);
  HIPCHK(hipMemcpy(cs.data(), b->d_calls.p, cs.size() * sizeof(CallStat), hipMemcpyDeviceToHost));
  for (int w = 0; w < b->n_windows; w++)
    for (int64_t k = b->call_off[w]; k < b->call_off[w + 1] && k < cap; k++)
      out[k] = call_stat_of(cs[k], (int)(k - b->call_off[w]) + 1, w);
  return MPOA_OK;
}

int mpoa_seed_anchors(const mpoa_params* params, int32_t n_reads, const char* seq,
                      const int64_t* seq_off, int32_t rid, const char* cons, int32_t cons_len,
                      uint64_t* out, int64_t cap, int64_t* n) {
  if (!params || !seq || !seq_off || rid < 0 || rid >= n_reads || cons_len < 0) return MPOA_E_ARG;
  const int64_t base0 = seq_off[0];
  std::vector<uint8_t> codes((size_t)(seq_off[n_reads] - base0) + 1);
  std::vector<int64_t> off(n_reads + 1);
  for (int i = 0; i <= n_reads; i++) off[i] = seq_off[i] - base0;
  for (int64_t j = 0; j < off[n_reads]; j++)
    codes[j] = code_of_char((unsigned char)seq[base0 + j]);
  std::vector<uint8_t> cc((size_t)cons_len + 1);
  for (int j = 0; j < cons_len; j++) cc[j] = code_of_char((unsigned char)cons[j]);
  seed::Config cfg;
  cfg.k = params->k;
  cfg.w = params->mm_w;
  cfg.bw = 1000;
  cfg.poa_w = params->poa_w;
  seed::Seeder sd;
  sd.init(cfg, n_reads, codes.data(), off.data());
  const int qlen = (int)(off[rid + 1] - off[rid]);
  const std::vector<seed::Pair> a = sd.anchors(rid, qlen, cc.data(), cons_len);
  if (n) *n = (int64_t)a.size();
  for (size_t i = 0; i < a.size() && (int64_t)i < cap; i++) {
    out[2 * i] = a[i].x;
    out[2 * i + 1] = a[i].y;
  }
  return MPOA_OK;
}

int mpoa_seed_anchors_device(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_reads,
                             const char* seq, const int64_t* seq_off, int32_t rid,
                             const char* cons, int32_t cons_len, uint64_t* out, int64_t cap,
                             int64_t* n) {
  if (!ctx || !params || !seq || !seq_off || rid < 0 || rid >= n_reads || cons_len < 0 ||
      (cap > 0 && !out) || params->k < 1 || params->k > 28 || params->mm_w < 1 ||
      params->mm_w > 255)
    return MPOA_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  // the reads go up as one window; the consensus as codes
  mpoa_params pp = *params;
  pp.seeding = 0;
  pp.progressive = 0;
  int32_t win_off[2] = {0, n_reads};
  mpoa_batch* b = nullptr;
  int rc = mpoa_batch_prepare(ctx, &pp, 1, win_off, seq, seq_off, nullptr, &b);
  if (rc != MPOA_OK) return rc;
  const int w = params->mm_w, k = params->k;
  std::vector<int64_t> rm_off(n_reads + 1, 0);
  for (int r = 0; r < n_reads; r++)
    rm_off[r + 1] = rm_off[r] + (seq_off[r + 1] - seq_off[r]) + 2 * w + 8;
  const int64_t qtot = seq_off[n_reads] - seq_off[0];
  const int cons_cap = std::max(cons_len, 1);
  const size_t rm = (size_t)rm_off[n_reads];
  DevBuf<uint64_t> u_buf;
  DevBuf<int> i_buf;
  DevBuf<int64_t> d_rm_off;
  DevBuf<uint8_t> d_cons;
  const size_t u_words = 4 * rm + (size_t)qtot + 5 * (size_t)cons_cap +
                         10 * (size_t)mpoa::SEED_ANCHORS;
  const size_t i_words = (size_t)n_reads + 2;
  if ((rc = u_buf.alloc(u_words)) != MPOA_OK || (rc = i_buf.alloc(i_words)) != MPOA_OK ||
      (rc = d_rm_off.alloc(n_reads + 1)) != MPOA_OK || (rc = d_cons.alloc(cons_cap)) != MPOA_OK) {
    mpoa_batch_free(b);
    return rc;
  }
  std::vector<uint8_t> cc((size_t)cons_cap, 4);
  for (int j = 0; j < cons_len; j++) cc[j] = code_of_char((unsigned char)cons[j]);
  HIPCHK(hipMemcpy(d_cons.p, cc.data(), cc.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_rm_off.p, rm_off.data(), rm_off.size() * sizeof(int64_t),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemset(i_buf.p, 0, i_words * sizeof(int)));
  mpoa::SeedDev d{};
  uint64_t* u = u_buf.p;
  auto take_u = [&](size_t m) { uint64_t* p = u; u += m; return p; };
  d.rmx = take_u(rm); d.rmy = take_u(rm); d.rtx = take_u(rm); d.rty = take_u(rm);
  d.hx = take_u((size_t)qtot + cons_cap);
  d.cmx = take_u(cons_cap); d.cmy = take_u(cons_cap);
  d.ctx = take_u(cons_cap); d.cty = take_u(cons_cap);
  d.ax = take_u(mpoa::SEED_ANCHORS); d.ay = take_u(mpoa::SEED_ANCHORS);
  d.bx = take_u(mpoa::SEED_ANCHORS); d.by = take_u(mpoa::SEED_ANCHORS);
  d.zx = take_u(mpoa::SEED_ANCHORS); d.zy = take_u(mpoa::SEED_ANCHORS);
  d.zsx = take_u(mpoa::SEED_ANCHORS); d.zsy = take_u(mpoa::SEED_ANCHORS);
  d.pick_out = take_u(2 * (size_t)mpoa::SEED_ANCHORS);
  d.rm_cnt = i_buf.p;
  d.n_pick_out = i_buf.p + n_reads;
  d.flag = i_buf.p + n_reads + 1;
  d.rm_off = d_rm_off.p;
  d.hx_cons = qtot;
  d.cons_cap = cons_cap;
  d.cons_codes = d_cons.p;
  d.dp_cap = INT64_MAX;
  d.k = k;
  d.w = w;
  d.bw = 1000;  // main.cpp:65
  d.poa_w = params->poa_w;
  d.pen_gap = (float)(0.01 * k);  // minimizer.cpp:498
  d.pen_skip = (float)(0.0 * 0.01 * k);
  SeedArgs sa{};
  sa.graph.read_seq = b->d_seq.p;
  sa.graph.read_off = b->d_off.p;
  sa.graph.win = b->d_win.p;
  sa.n_reads = n_reads;
  sa.read = rid;
  sa.cons_cap = cons_len;
  sa.dev = d;
  if (mpoa_launch_seed_reads(sa, ctx->stream) != 0 ||
      mpoa_launch_seed_anchors(sa, ctx->stream) != 0) {
    mpoa_batch_free(b);
    g_err = "seeding launch failed";
    return MPOA_E_DEVICE;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int got[2];
  HIPCHK(hipMemcpy(got, i_buf.p + n_reads, sizeof(got), hipMemcpyDeviceToHost));
  mpoa_batch_free(b);
  if (got[0] < 0 || got[1] != 0 || got[0] > mpoa::SEED_ANCHORS) {
    g_err = "past the device seeding's capacities (mpoa_seed_anchors seeds on the host)";
    return MPOA_E_OVERFLOW;
  }
  std::vector<uint64_t> picks(2 * (size_t)std::max(got[0], 1));
  if (got[0] > 0)
    HIPCHK(hipMemcpy(picks.data(), d.pick_out, 2 * (size_t)got[0] * sizeof(uint64_t),
                     hipMemcpyDeviceToHost));
  if (n) *n = got[0];
  // picked last anchor first: the sub-windows take them in read order
  for (int i = 0; i < got[0] && i < cap; i++) {
    out[2 * i] = picks[2 * (size_t)(got[0] - 1 - i)];
    out[2 * i + 1] = picks[2 * (size_t)(got[0] - 1 - i) + 1];
  }
  return MPOA_OK;
}

int mpoa_batch_launch_stats(mpoa_batch* b, mpoa_launch_stat* out, int32_t cap, int32_t* n) {
  if (!b || (cap > 0 && !out)) return MPOA_E_ARG;
  const int k = (int)b->lstats.size();
  for (int i = 0; i < k && i < cap; i++) out[i] = b->lstats[i];
  if (n) *n = k;
  return MPOA_OK;
}

int mpoa_batch_set_schedule(mpoa_batch* b, int32_t small_qmax, int32_t small_width,
                            int32_t move_top, int32_t small_nodes) {
  if (!b || small_qmax < 0 || small_width < 0 || small_nodes < 0) return MPOA_E_ARG;
  b->small_qmax = small_qmax;
  b->small_width = small_width;
  b->move_top = move_top;
  b->small_nodes = small_nodes;
  assign_bins(b);
  return MPOA_OK;
}

int mpoa_batch_set_tile(mpoa_batch* b, int32_t threads, int32_t cols) {
  if (!b || !mpoa_tile_supported(threads, cols)) return MPOA_E_ARG;
  b->small.nt = threads;
  b->small.cpl = cols;
  return MPOA_OK;
}

int mpoa_align_reads(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_reads, const char* seq,
                     const int64_t* seq_off, const char* name, const int64_t* name_off,
                     const char* comment, const int64_t* comment_off, char* out,
                     int64_t out_cap, int64_t* out_len, mpoa_call_stat* stats,
                     int32_t stats_cap, int32_t* n_stats) {
  if (!ctx || !params || n_reads < 0 || (n_reads > 0 && (!seq || !seq_off))) return MPOA_E_ARG;
  // Every argument the text and the stats depend on, as bytes: a call with the same key as the
  // held one (a call whose text did not fit) takes the held text instead of running again.
  std::string key;
  {
    auto put = [&key](const void* p, size_t n) { key.append(static_cast<const char*>(p), n); };
    auto put_list = [&](const char* data, const int64_t* off) {
      const int flag = data && off ? 1 : 0;
      put(&flag, sizeof flag);
      if (!flag) return;
      put(off, sizeof(int64_t) * ((size_t)n_reads + 1));
      if (n_reads > 0 && off[n_reads] > off[0]) put(data + off[0], (size_t)(off[n_reads] - off[0]));
    };
    put(params, sizeof *params);
    put(&n_reads, sizeof n_reads);
    if (n_reads > 0) put_list(seq, seq_off);
    put_list(name, name_off);
    put_list(comment, comment_off);
  }
  auto deliver = [&](std::string& text, const std::vector<mpoa_call_stat>& calls) {
    if (n_stats) *n_stats = (int32_t)calls.size();
    if (stats)
      for (int i = 0; i < (int)calls.size() && i < stats_cap; i++) stats[i] = calls[i];
    if (out_len) *out_len = (int64_t)text.size();
    if (!out || (int64_t)text.size() > out_cap) {
      if (&text != &ctx->held_text) {
        ctx->held_key.swap(key);
        ctx->held_text.swap(text);
        ctx->held_calls = calls;
      }
      return out ? MPOA_E_OUT_CAP : MPOA_OK;
    }
    std::memcpy(out, text.data(), text.size());
    ctx->held_key.clear();
    ctx->held_text.clear();
    ctx->held_calls.clear();
    return MPOA_OK;
  };
  if (!ctx->held_key.empty() && key == ctx->held_key) {
    const std::vector<mpoa_call_stat> calls = ctx->held_calls;
    return deliver(ctx->held_text, calls);
  }
  int32_t win_off[2] = {0, n_reads};
  std::vector<int64_t> zero_off{0};
  const int64_t* so = n_reads > 0 ? seq_off : zero_off.data();
  // The batch carries the reads, the window descriptor and the output buffer; a seeded window
  // is then driven read by read (run_seeded), any other in one launch.
  mpoa_params pp = *params;
  const bool seeding = pp.seeding != 0 && n_reads > 0;
  pp.seeding = 0;
  mpoa_batch* b = nullptr;
  int rc = mpoa_batch_prepare(ctx, &pp, 1, win_off, seq ? seq : "", so, nullptr, &b);
  if (rc != MPOA_OK) return rc;
  b->params.seeding = params->seeding;
  if (params->result == 2 && (rc = b->d_gfa.alloc((size_t)n_reads + 1)) != MPOA_OK) {
    mpoa_batch_free(b);
    return rc;
  }
  std::vector<mpoa_call_stat> calls;
  if (seeding) {
    // with -p the batch holds the reads in guide-tree order; the seeding follows that order
    const bool perm = !b->row_of.empty();
    const int64_t* po = perm ? b->perm_off.data() : so;
    ctx->last_seeding = 0;
    rc = (ctx->forms & MPOA_FORM_HOST_SEED) ? 1 : run_seeded_dev(b, po, &calls);
    if (rc == 1) rc = run_seeded(b, perm ? b->perm_seq.data() : seq, po, &calls);
  } else {
    b->split = false;  // one window: the wide-tile launch
    assign_bins(b);
    if ((rc = mpoa_batch_set_traces(b, 1)) != MPOA_OK) {
      mpoa_batch_free(b);
      return rc;
    }
    rc = mpoa_batch_execute(b);
    if (n_reads > 1) {
      std::vector<CallStat> cs(n_reads - 1);
      if (hipMemcpy(cs.data(), b->d_calls.p, cs.size() * sizeof(CallStat),
                    hipMemcpyDeviceToHost) == hipSuccess)
        for (int i = 0; i < (int)cs.size(); i++) calls.push_back(call_stat_of(cs[i], i + 1, 0));
    }
  }
  if (n_stats) *n_stats = (int32_t)calls.size();
  if (stats)
    for (int i = 0; i < (int)calls.size() && i < stats_cap; i++) stats[i] = calls[i];
  if (rc != MPOA_OK) {
    mpoa_batch_free(b);
    return rc;
  }
  const WinResult r = b->res.empty() ? WinResult{} : b->res[0];
  std::string text;
  rc = mpoa::host::window_text(params->result, r, b->d_out.p, b->d_gfa.p, n_reads, name,
                               name_off, comment, comment_off,
                               b->row_of.empty() ? nullptr : b->row_of.data(), &text);
  if (rc != MPOA_OK) {
    mpoa_batch_free(b);
    return rc;
  }
  mpoa_batch_free(b);
  return deliver(text, calls);
}

int mpoa_align_reads_noalloc(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_reads,
                             const char* d_seq, const int64_t* d_seq_off, const int64_t* seq_off,
                             char* d_out, int64_t out_cap, int64_t* d_res, void* stream) {
  if (!ctx || !params || n_reads < 0 || !d_seq_off || !seq_off || !d_out || out_cap < 0 ||
      !d_res || (n_reads > 0 && !d_seq))
    return MPOA_E_ARG;
  if (params->seeding) {
    g_err = "seeding (-S) windows are not supported by the allocation-free form";
    return MPOA_E_ARG;
  }
  if (params->result != 0 && params->result != 1) {
    g_err = "the allocation-free form writes consensus (0) or MSA (1) output";
    return MPOA_E_ARG;
  }
  if (!ctx->scratch) return MPOA_E_DEVICE;
  HIPCHK(hipSetDevice(ctx->device));
  const DpScores sc = scores_of(*params);
  WinInfo wi{n_reads, 0, 0};
  for (int r = 0; r < n_reads; r++) {
    const int64_t l = seq_off[r + 1] - seq_off[r];
    if (l < 0) return MPOA_E_ARG;
    wi.qmax = (int)std::max<int64_t>(wi.qmax, l);
    wi.qtot += l;
  }
  const bool keep_path = params->result != 0;
  const bool full = params->ab_band != 0 && params->band_f > 0;
  // Scratch layout: window records in the first 4 KB, then the int, byte and DP pools of one
  // slot, each on a 256-byte boundary.
  constexpr size_t kHead = 4096;
  auto layout_bytes = [&](const SlotCaps& c) {
    return kHead + (size_t)c.ints_per_slot * 4 + (size_t)c.bytes_per_slot + (size_t)c.dp * 12;
  };
  auto caps_of = [&](const Estimate& e) {
    SlotCaps c{};
    c.nodes = (std::max(e.nodes, 2) + 63) / 64 * 64;
    c.edges = (std::max(e.edges, 4) + 63) / 64 * 64;
    c.qmax = std::max(1, wi.qmax);
    c.path = keep_path ? (int)wi.qtot : 0;
    c.dp = (std::max<int64_t>(e.dp, 64) + 63) / 64 * 64;
    mpoa::finish_caps(c);
    return c;
  };
  // The read-length estimate, widened towards the worst case as far as the array allows.
  const Estimate est = estimate(wi, sc, full), worst = worst_case(wi);
  SlotCaps caps = caps_of(est);
  if (layout_bytes(caps) > MPOA_NOALLOC_SCRATCH_BYTES) {
    g_err = "window needs " + std::to_string(layout_bytes(caps)) +
            " bytes of device scratch; the allocation-free form has " +
            std::to_string((size_t)MPOA_NOALLOC_SCRATCH_BYTES);
    return MPOA_E_SCRATCH;
  }
  for (double grow : {4.0, 2.0, 1.5}) {
    Estimate e;
    e.nodes = (int)std::min<double>(worst.nodes, grow * est.nodes);
    e.edges = (int)std::min<double>(worst.edges, grow * est.edges);
    e.dp = (int64_t)std::min<double>((double)worst.dp, grow * (double)est.dp);
    const SlotCaps c = caps_of(e);
    if (layout_bytes(c) <= MPOA_NOALLOC_SCRATCH_BYTES) {
      caps = c;
      break;
    }
  }
  uint8_t* base = ctx->scratch;
  KernelArgs a{};
  // int32 storage only: this entry point has no re-run for a call whose values leave int16
  a.forms = ctx->forms | MPOA_FORM_WIDE;
  a.scores = sc;
  a.caps = caps;
  a.read_seq = reinterpret_cast<const uint8_t*>(d_seq);
  a.read_off = d_seq_off;
  a.win = reinterpret_cast<const WinDesc*>(base);
  a.result = reinterpret_cast<WinResult*>(base + 256);
  a.order = reinterpret_cast<const int*>(base + 2048);
  a.n_jobs = 1;
  a.queue = reinterpret_cast<int*>(base + 2304);
  a.int_pool = reinterpret_cast<int*>(base + kHead);
  a.byte_pool = base + kHead + (size_t)caps.ints_per_slot * 4;
  a.dp_pool = reinterpret_cast<int*>(a.byte_pool + caps.bytes_per_slot);
  a.out = reinterpret_cast<uint8_t*>(d_out);
  static_assert(sizeof(WinDesc) <= 256 && sizeof(WinResult) <= 2048 - 256, "scratch head");
  WinDesc wd{};
  wd.read0 = 0;
  wd.n_reads = n_reads;
  wd.out0 = 0;
  wd.out_cap = (int)std::min<int64_t>(out_cap, INT32_MAX / 2);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  // one window alone: the tile mpoa_align_reads gives it
  const int rc = mpoa_launch_noalloc(a, wd, 512, 2, d_res, st);
  if (rc != 0) {
    g_err = "kernel launch failed";
    return MPOA_E_DEVICE;
  }
  return MPOA_OK;
}

}  // extern "C"
