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

// Int32 and byte words a slot's arena takes; must match carve() in pipeline.hip.
inline __host__ __device__ int64_t slot_ints(int N, int E, int qmax, int path) {
  return 46LL * N + 8LL * E + 2LL * qmax + 8 + path;
}
inline __host__ __device__ int64_t slot_bytes(int N, int E, int qmax) {
  return 2LL * N + E + 3LL * qmax + 128;
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
  int nodes, edges;
  int64_t ticks[PH_N];
  int64_t t_begin, t_end;  // device wall clock at the window's start and end
};

// One DP call (one read against the graph of the reads before it).
struct CallStat {
  int score, rows, qlen, width;  // width: widest row, 8*block_num
  int64_t cells, edge_cells;
  int nodes, edges;              // graph size when the call ran
};

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
};

// -S sub-window of the current read: the DP of read characters [q0, q0 + qlen) against the
// graph's rank window [rank(beg_id), rank(end_id)]; its alignment goes to res [res0, ...).
struct SegDesc {
  int beg_id, end_id, q0, qlen, res0, pad;
};

struct SegResult {
  int status, score, rows, qlen, width, pad;
  int64_t cells, edge_cells;
  int64_t ticks[PH_N];
};

// Window state kept in device memory between the launches of a seeded window.
struct WinState {
  int n_nodes, n_in_e, n_out_e, status, msa_cols, cons_len, pad0, pad1;
  int64_t cells, edge_cells;
  int64_t ticks[PH_N];
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
};

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
