// Kernel-argument structures shared by the device pipeline (pipeline.hip) and the host
// library (mpoa.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpoa.h"

namespace mpoa {

// Scores after minipoa's normalisation: o1 = gap_open + gap_ext (src/align.cpp:1369).
struct DpScores {
  int match, mismatch, o1, e1;
  int b, f;
  int ab_band;
  int result;
};

// Capacities of one slot (one resident workgroup). A window that outgrows any of them
// stops with MPOA_E_OVERFLOW and is re-run by the host with larger capacities.
struct SlotCaps {
  int nodes;       // graph nodes
  int edges;       // edge records per direction
  int qmax;        // longest read
  int path;        // sum of read lengths kept for MSA/GFA paths (0 in consensus mode)
  int64_t dp;      // int32 cells per DP matrix (H, D and I each)
  int64_t ints_per_slot;
  int64_t bytes_per_slot;
};

// Int32 and byte words a slot's arena takes: carve() below stays within them (checked on the
// host for many capacities by tests/carve_check.cpp).
inline __host__ __device__ int64_t slot_ints(int N, int E, int qmax, int path) {
  return 46LL * N + 8LL * E + 2LL * qmax + 8 + path;
}
inline __host__ __device__ int64_t slot_bytes(int N, int E, int qmax) {
  return 2LL * N + E + 3LL * qmax + 128;
}

// Slot sizes of caps whose nodes, edges, qmax, path and dp are set: whole 256-byte blocks.
inline void finish_caps(SlotCaps& c) {
  c.ints_per_slot = (slot_ints(c.nodes, c.edges, c.qmax, c.path) + 63) / 64 * 64;
  c.bytes_per_slot = (slot_bytes(c.nodes, c.edges, c.qmax) + 255) / 256 * 256;
}

constexpr int NB = 5;         // alphabet incl. N (minipoa's m = 5, src/handle_input.cpp:51)
constexpr int QPAD = 32;      // query bytes past qmax: the N padding to col_size plus 2*Q
constexpr int RING_ROWS = 8;  // rows kept in the LDS ring (tiles above 64 threads)

// Per-slot arena: carved from the int pool / byte pool / DP pool at fixed offsets. Every
// device function that takes the arena is force-inlined: one left as a call needs the arena's
// address, which moves the whole structure to scratch memory and turns each of its pointers
// into a scratch load inside the DP, sort and backtrack loops.
struct Arena {
  uint8_t* base;
  int *par, *aligned, *ind, *rank, *in_head, *in_tail, *out_head, *out_tail, *in_cnt, *last_read;
  int* out_cnt;
  int *ie_src, *ie_w, *ie_next, *oe_dst, *oe_w, *oe_next;
  int *rank_id, *t0, *t1, *t2, *hlen, *tlen;
  int *pred_off, *pred_rank, *pred_w, *Bs, *Be, *row_off;
  uint8_t* pred_base;
  int* rowrec;  // RowRec per rank (12 ints)
  int *Mp, *Sl, *Pl, *Pr, *Ol, *Or, *Ow;
  int *res_from, *res_aligned;
  int *cons_id, *cons_rank;  // consensus path for -S seeding
  uint8_t *res_base, *qseq, *qsh;
  uint8_t* cons_ch;
  int* path;
  int *H, *D, *I;
  int* int_end;       // one past the last int word taken
  uint8_t* byte_end;  // one past the last byte taken
};

// Slot `slot` of the three pools. Pointer arithmetic only (no access), so the host can check
// the layout against slot_ints / slot_bytes. The int arrays start with the row records (int4
// accesses: ints_per_slot is a multiple of 64); qseq and qsh are 16-byte aligned.
__host__ __device__ inline Arena carve_pools(const SlotCaps& c, int* int_pool, uint8_t* byte_pool,
                                             int* dp_pool, int slot) {
  Arena A;
  int* p = int_pool + (int64_t)slot * c.ints_per_slot;
  auto take = [&](int64_t n) { int* r = p; p += n; return r; };
  const int N = c.nodes, E = c.edges;
  A.rowrec = take(12LL * N);
  A.par = take(N); A.aligned = take((int64_t)NB * N); A.ind = take(N); A.rank = take(N);
  A.in_head = take(N); A.in_tail = take(N); A.out_head = take(N); A.out_tail = take(N);
  A.in_cnt = take(N); A.last_read = take(N); A.out_cnt = take(N);
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
  A.cons_id = take(N); A.cons_rank = take(N);
  A.int_end = p;
  uint8_t* b = byte_pool + (int64_t)slot * c.bytes_per_slot;
  A.base = b; b += N;
  A.cons_ch = b; b += N;
  A.pred_base = b; b += E;
  A.res_base = b; b += c.qmax + 2;
  auto align16 = [](uint8_t* x) {
    return reinterpret_cast<uint8_t*>((reinterpret_cast<uintptr_t>(x) + 15) & ~uintptr_t(15));
  };
  A.qseq = align16(b); b = A.qseq + c.qmax + QPAD;
  A.qsh = align16(b); b = A.qsh + c.qmax + QPAD;
  A.byte_end = b;
  int* dp = dp_pool + (int64_t)slot * 3 * c.dp;
  A.H = dp; A.D = dp + c.dp; A.I = dp + 2 * c.dp;
  return A;
}

struct WinDesc {
  int read0, n_reads;  // reads [read0, read0 + n_reads) of the read table
  int64_t out0;        // output bytes [out0, out0 + out_cap)
  int out_cap;
  int pad;
};

// Device timers (wall_clock64 ticks) per phase of the window pipeline.
// PH_KAHN and PH_DEPTH are parts of PH_SORT (the order itself; hlen/tlen).
enum Phase { PH_PLAN = 0, PH_DP, PH_BT, PH_ADD, PH_SORT, PH_OUT, PH_KAHN, PH_DEPTH, PH_N };

struct WinResult {
  int status;
  int out_len;     // consensus length, or MSA columns
  int64_t cells, edge_cells;
  int64_t logical;  // sum over DP rows 1..n-1 of Me-Ms+1 (the unquantised band)
  int nodes, edges;
  int64_t ticks[PH_N];
  int64_t t_begin, t_end;  // device wall clock at the window's start and end
  int64_t narrow_cells;    // band cells of the calls whose H, D and I were stored as int16
};

// One DP call (one read against the graph of the reads before it).
struct CallStat {
  int score, rows, qlen, width;  // width: widest row, 8*block_num
  int64_t cells, edge_cells, logical;
  int nodes, edges;              // graph size when the call ran
};

struct SeedDev;

struct KernelArgs {
  DpScores scores;
  SlotCaps caps;
  const uint8_t* read_seq;
  const int64_t* read_off;
  const WinDesc* win;
  WinResult* result;
  const int* order;
  int n_jobs;
  int* queue;
  int* int_pool;
  uint8_t* byte_pool;
  int* dp_pool;
  uint8_t* out;
  int64_t* gfa_lines;       // optional (one window): offset of each read's GFA path line
  CallStat* call_stats;     // optional: per DP call, window w's calls from call_off[w]
  const int64_t* call_off;  // n_windows + 1 offsets (reads - 1 per window)
  int forms;                // MPOA_FORM_* bits: fallback forms taken for every read (0 in use)
  const SeedDev* seed;      // seeded batches: each slot's seeding buffers (k_seeded)
};

// -S sub-window of the current read: the DP of read characters [q0, q0 + qlen) against the
// graph's rank window [rank(beg_id), rank(end_id)]; its alignment goes to res [res0, ...).
// With dp_len > 0 its H, D and I (dp_len int32 each) lie at dp_pool + dp0 instead of the slot's
// share of the pool (the sub-windows a read's device seeding carves out of one pool).
struct SegDesc {
  int beg_id, end_id, q0, qlen, res0, pad;
  int64_t dp0, dp_len;
};

struct SegResult {
  int status, score, rows, qlen, width, pad;
  int64_t cells, edge_cells, logical;
  int64_t ticks[PH_N];
};

// Window state kept in device memory between the launches of a seeded window.
struct WinState {
  int n_nodes, n_in_e, n_out_e, status, msa_cols, cons_len, pad0, pad1;
  int64_t cells, edge_cells, logical;
  int64_t ticks[PH_N];
};

// Capacities of device seeding (per read): sub-windows, anchors kept for chaining, entries one
// LDS sort takes, and anchor-base splices.
constexpr int SEED_SEGS = 32;
constexpr int SEED_ANCHORS = 4096;
constexpr int SEED_SORT = 8192;
constexpr int SEED_SPLICE = 1024;
// Why a window left device seeding (SeedDev::flag bits; bit 1 is the graph estimate's, set by
// the host): read minimizers past capacity or with ties, consensus past cons_cap, its
// minimizers past cons_cap, ties in their sort, anchors past SEED_ANCHORS, ties among them or
// among chain ends or chains, sub-windows past SEED_SEGS, their pool, splices.
enum {
  SEED_WHY_GRAPH = 2, SEED_WHY_READS = 4, SEED_WHY_CONS = 8, SEED_WHY_CONS_MM = 16,
  SEED_WHY_CONS_SORT = 32, SEED_WHY_ANCHORS = 64, SEED_WHY_ANCHOR_SORT = 128,
  SEED_WHY_END_SORT = 256, SEED_WHY_CHAIN_SORT = 512, SEED_WHY_SEGS = 1024,
  SEED_WHY_POOL = 2048, SEED_WHY_SPLICE = 4096
};

// Device seeding of one -S window (minipoa's minimizer.cpp / lchain.cpp steps on the GPU): the
// reads' minimizers sorted by hash, once; after each read's commit, the consensus's minimizers,
// the anchors of the next read, its chains and its sub-windows. Buffers in HBM, sized by the
// host; a step that outgrows them sets *flag and stops the window (the host then seeds it).
struct SeedDev {
  uint64_t *rmx, *rmy;      // read minimizers sorted by x: read r at [rm_off[r], + rm_cnt[r])
  uint64_t *rtx, *rty;      // the same before sorting
  const int64_t* rm_off;    // n_reads + 1 capacity offsets
  int* rm_cnt;              // n_reads
  uint64_t* hx;             // x per position: the window's read bases, then cons_cap more
  int64_t hx_cons;          // first consensus position in hx
  int cons_cap;             // consensus positions and minimizers kept
  uint64_t *cmx, *cmy, *ctx, *cty;  // consensus minimizers: sorted / as emitted (cons_cap)
  uint64_t *ax, *ay, *bx, *by;      // anchors: joined / sorted (SEED_ANCHORS)
  uint64_t *zx, *zy, *zsx, *zsy;    // chain ends, then chains: unsorted / sorted (SEED_ANCHORS)
  SegDesc* segs;            // SEED_SEGS per read
  int* splice;              // SEED_SPLICE (position, node id) pairs per read
  int* n_seg;               // per read (0: the read is added as a new path)
  int* n_splice;            // per read
  int* read_graph;          // per read: nodes and in-edges after its commit
  int* flag;                // nonzero: a capacity was exceeded; the host seeds this window
  int64_t dp_cap;           // int32 words of dp_pool for one read's sub-windows (H, D, I each)
  int k, w, bw, poa_w;
  float pen_gap, pen_skip;
  // anchors only (k_seed_anchors): the consensus codes, and the picked anchors out (x, y
  // pairs as collect_anchors_bycons returns them, last anchor first; SEED_ANCHORS at most)
  const uint8_t* cons_codes;
  uint64_t* pick_out;
  int* n_pick_out;
};

struct SeedArgs {
  KernelArgs graph;        // the window (win[0]) and the graph in slot 0 of its pools
  KernelArgs seg;          // private scratch per sub-window (slot = sub-window)
  const SegDesc* segs;
  SegResult* seg_res;
  int n_seg;               // 0: trivial read (graph without bases)
  const int* splice;       // (res position, node id) pairs: anchor bases
  int n_splice;
  int read, n_reads;
  int path0;               // bases of the reads before this one
  WinState* state;
  int* cons_out;           // after a commit: (node id, rank, base) per consensus position
  int cons_cap;
  int use_dev;             // 1: sub-window counts from dev, and the commit seeds the next read
  SeedDev dev;
};

// One step of minipoa's per-read loop on a graph kept in device memory between calls (the
// call-level API: poa(), graph::add_path, graph::topsort, the outputs). The graph is slot 0 of
// a's pools, its counters and status in `state`.
enum GraphOp {
  GOP_INIT = 0, GOP_POA = 1, GOP_ADD_PATH = 2, GOP_TOPSORT = 3, GOP_OUTPUT = 4,
  GOP_ADD_NODES = 5  // graph::add_node for each of q[0..qlen) (codes 0..4), without edges
};

struct GraphOpOut {
  int status, score, rows, qlen, width, pad;
  int64_t cells, edge_cells, logical;
};

struct GraphOpArgs {
  KernelArgs a;           // a.win[0] / a.read_off: the reads added so far (outputs)
  WinState* state;
  GraphOpOut* out;
  int op;
  int beg_id, end_id, adaptive;  // GOP_POA: rank window of node ids, band mode
  const uint8_t* q;              // GOP_POA: query (ASCII or codes 0..4), qlen bases
  int qlen;
  int seq_id, n_res, path0;      // GOP_ADD_PATH: read index, res entries (already in the
                                 // graph's res arrays), bases of the reads before it
  int sort_op, seeding;          // GOP_TOPSORT: 0 or 1; with seeding, the consensus too
  int* cons_out;                 // (node id, rank, base) per consensus position
  int cons_cap;
  int result;                    // GOP_OUTPUT: 0 consensus, 1 MSA, 2 GFA, 3 MSA + consensus
};

__host__ __device__ inline Arena carve(const KernelArgs& a, int slot) {
  return carve_pools(a.caps, a.int_pool, a.byte_pool, a.dp_pool, slot);
}

}  // namespace mpoa

// Launches the persistent pipeline with n_slots workgroups of nt threads, cpl DP columns per
// thread per pass. Returns 0, or -2 for an unsupported (nt, cpl).
int mpoa_launch_windows(const mpoa::KernelArgs& a, int n_slots, int nt, int cpl,
                        hipStream_t stream);
bool mpoa_tile_supported(int nt, int cpl);
// Workgroups of the (nt, cpl) window kernel one CU holds at once (0 if unsupported).
int mpoa_tile_occupancy(int nt, int cpl);
// Graph form of one persistent launch: a node resetting the work queue after `deps`, then the
// k_windows node (returned in *last). 0, -1 on a graph error, -2 for an unsupported tile.
int mpoa_graph_add_windows(hipGraph_t g, const hipGraphNode_t* deps, size_t n_deps,
                           const mpoa::KernelArgs& a, int n_slots, int nt, int cpl,
                           hipGraphNode_t* last);
// A one-wavefront node that waits `ticks` of the device wall clock after `deps`.
int mpoa_graph_add_hold(hipGraph_t g, const hipGraphNode_t* deps, size_t n_deps, int64_t ticks,
                        hipGraphNode_t* node);
// Seeded windows: every sub-window of one read (one workgroup each), then the commit of that
// read (one workgroup). 0, -1 on a launch error, -2 for an unsupported tile.
int mpoa_launch_segments(const mpoa::SeedArgs& s, int nt, int cpl, hipStream_t stream);
int mpoa_launch_commit(const mpoa::SeedArgs& s, hipStream_t stream);
// Device seeding: every read's minimizers (one workgroup per read), before the first commit.
int mpoa_launch_seed_reads(const mpoa::SeedArgs& s, hipStream_t stream);
// The anchors of read s.read on the consensus s.dev.cons_codes (s.cons_cap bases), after
// mpoa_launch_seed_reads: one workgroup, into s.dev.pick_out / n_pick_out.
int mpoa_launch_seed_anchors(const mpoa::SeedArgs& s, hipStream_t stream);
// Seeded (-S) windows of a batch: the persistent launch, one window per workgroup at a time,
// every read seeded and its sub-windows run in turn; nt 64 (a 30 KB tile, several per CU) or
// 256 (one per CU, room for any window's seeding). 0, -1 on a launch error, -2 for another nt.
int mpoa_launch_seeded(const mpoa::KernelArgs& a, int n_slots, int nt, hipStream_t stream);
// Workgroups of the nt-thread seeded tile one CU holds at once (0 if unsupported).
int mpoa_seeded_occupancy(int nt);
// Device address of this device's allocation-free scratch array (MPOA_NOALLOC_SCRATCH_BYTES),
// looked up once per device with that device current. 0, or -1 when the lookup fails.
int mpoa_scratch_address(int device, void** out);
// The allocation-free window: descriptor and queue set-up, the persistent kernel with one slot,
// and the copy of status and output length to d_res[0..1]; every pointer in `a` points into
// the scratch array or the caller's buffers. 0, -1 on a launch error, -2 for an unsupported
// tile.
int mpoa_launch_noalloc(const mpoa::KernelArgs& a, const mpoa::WinDesc& wd, int nt, int cpl,
                        int64_t* d_res, hipStream_t stream);
// One graph operation (GraphOpArgs::op) by one 512-thread workgroup. 0, or -1 on a launch error.
int mpoa_launch_graph_op(const mpoa::GraphOpArgs& g, hipStream_t stream);
