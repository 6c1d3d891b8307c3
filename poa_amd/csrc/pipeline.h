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
  return 44LL * N + 8LL * E + 2LL * qmax + 8 + path;
}
inline __host__ __device__ int64_t slot_bytes(int N, int E, int qmax) {
  return (int64_t)N + E + 3LL * qmax + 128;
}

struct WinDesc {
  int read0, n_reads;  // reads [read0, read0 + n_reads) of the read table
  int64_t out0;        // output bytes [out0, out0 + out_cap)
  int out_cap;
  int pad;
};

// Device timers (wall_clock64 ticks) per phase of the window pipeline.
enum Phase { PH_PLAN = 0, PH_DP, PH_BT, PH_ADD, PH_SORT, PH_OUT, PH_N };

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
  CallStat* call_stats;     // optional: per DP call, window w's calls from call_off[w]
  const int64_t* call_off;  // n_windows + 1 offsets (reads - 1 per window)
};

}  // namespace mpoa

// Launches the persistent pipeline with n_slots workgroups of nt threads, cpl DP columns per
// thread per pass. Returns 0, or -2 for an unsupported (nt, cpl).
int mpoa_launch_windows(const mpoa::KernelArgs& a, int n_slots, int nt, int cpl,
                        hipStream_t stream);
bool mpoa_tile_supported(int nt, int cpl);
