// Host-side internals shared by the library's C++ sources (mpoa.cpp, graph_api.cpp): the
// device context, error text, device buffers, score normalisation and capacity estimates.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "mpoa.h"
#include "pipeline.h"

struct mpoa_ctx {
  int device = 0;
  hipStream_t stream = nullptr;   // small-window launches, copies
  hipStream_t stream2 = nullptr;  // medium-tile launches, concurrent with the first
  hipStream_t stream3 = nullptr;  // wide-tile launches, concurrent with both
  int n_cu = 0;
  size_t mem_total = 0;
  int wall_khz = 100000;          // device wall clock (the k_hold timer)
  // mpoa_align_reads: the last call whose text did not fit its buffer, for the follow-up call
  // with the same arguments
  std::string held_key, held_text;
  std::vector<mpoa_call_stat> held_calls;
  int forms = 0;  // MPOA_FORM_* test hook: the forms of batches created afterwards
  int last_seeding = -1;  // the last seeded mpoa_align_reads: 1 device seeding, 0 host, -1 none
  uint8_t* scratch = nullptr;  // this device's allocation-free scratch array (device address)
};

namespace mpoa {
namespace host {

// Text of the last error on this thread (mpoa_strerror).
extern thread_local std::string g_err;

#define HIPCHK(expr)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      g_err = std::string(#expr) + ": " + hipGetErrorString(e_);                         \
      return e_ == hipErrorOutOfMemory ? MPOA_E_OOM : MPOA_E_DEVICE;                      \
    }                                                                                    \
  } while (0)

// nt4 encoding (src/sequence.cpp:6-23): A C G T/U -> 0..3, anything else 4.
inline uint8_t code_of_char(unsigned char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
  }
}

// initPara (src/handle_input.cpp:45-77) + o1 = O + E (src/align.cpp:1369).
inline mpoa::DpScores scores_of(const mpoa_params& p) {
  mpoa::DpScores s;
  s.match = p.match;
  s.mismatch = p.mismatch > 0 ? -p.mismatch : p.mismatch;
  const int O = p.gap_open > 0 ? -p.gap_open : p.gap_open;
  s.e1 = p.gap_ext > 0 ? -p.gap_ext : p.gap_ext;
  s.o1 = O + s.e1;
  s.b = p.band_b;
  s.f = p.band_f;
  s.ab_band = p.ab_band;
  s.result = p.result;
  return s;
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      g_err = std::string("hipMalloc: ") + hipGetErrorString(e);
      return MPOA_E_OOM;
    }
    n = count;
    return MPOA_OK;
  }
};

struct WinInfo {
  int n_reads;
  int qmax;
  int64_t qtot;
};

// Capacities estimated from the reads alone. Nodes: the longest read plus 0.16 node per other
// base (reads at up to 15% error add at most ~0.12 new nodes per base); DP cells: nodes times
// the static band width 2w + 64. The device checks every capacity exactly and a window that
// outgrows one is re-run with worst-case capacities.
struct Estimate {
  int nodes, edges;
  int64_t dp;
};

inline Estimate estimate(const WinInfo& w, const DpScores& s, bool full_width) {
  Estimate e;
  int64_t n = 2 + w.qmax + (int64_t)std::ceil(0.16 * (double)(w.qtot - w.qmax));
  n = std::min<int64_t>(n, 2 + w.qtot);
  e.nodes = (int)n;
  e.edges = (int)std::min<int64_t>(3 * n + w.n_reads, w.qtot + w.n_reads + 2);
  const int m = w.qmax + 1;
  int64_t width;
  if (s.f > 0 && !full_width) {
    const int wb = s.b + m / s.f;
    width = std::min<int64_t>(2LL * wb + 64, (int64_t)(m / 8 + 2) * 8);
  } else {
    width = (int64_t)(m / 8 + 2) * 8;
  }
  e.dp = n * width;
  return e;
}

// Exact worst case: every base a new node; every row full width.
inline Estimate worst_case(const WinInfo& w) {
  Estimate e;
  e.nodes = (int)std::min<int64_t>(2 + w.qtot, INT32_MAX / 64);
  e.edges = (int)std::min<int64_t>(w.qtot + w.n_reads + 2, INT32_MAX / 64);
  e.dp = (int64_t)e.nodes * ((w.qmax + 1) / 8 + 2) * 8;
  return e;
}

// minipoa's stdout for a finished window (result 0: ">consensus sequence" + bases; 1: the
// row-column MSA with its headers; 2: GFA with read names on the path lines) from the device
// output region d_out (and, for GFA, the path-line offsets d_gfa) of the result record r.
// row_of (optional): the MSA row read i prints, when reads were added in another order (-p).
int window_text(int result, const mpoa::WinResult& r, const uint8_t* d_out, const int64_t* d_gfa,
                int n_reads, const char* name, const int64_t* name_off, const char* comment,
                const int64_t* comment_off, const int* row_of, std::string* out);

}  // namespace host
}  // namespace mpoa
