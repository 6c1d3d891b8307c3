// Host check of the per-slot arena layout (carve_pools in poa_amd/csrc/pipeline.h): for many
// slot capacities, carves three consecutive slots out of host buffers sized exactly as the
// library sizes its pools, and checks that every slot stays inside its share of each pool, that
// the vector-accessed arrays are 16-byte aligned, and that the query arrays hold the longest
// query the capacities allow (col_size + 2*Q bytes, src/align.cpp:1362-1373). Prints "ok" or
// the first failing capacity set; exit status 1 on a failure.
#include <cstdio>
#include <cstdint>
#include <cstdlib>

#include "pipeline.h"

using mpoa::Arena;
using mpoa::SlotCaps;

static int fail(const SlotCaps& c, int slot, const char* what) {
  std::printf("FAIL nodes=%d edges=%d qmax=%d path=%d dp=%lld slot=%d: %s\n", c.nodes, c.edges,
              c.qmax, c.path, (long long)c.dp, slot, what);
  return 1;
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

static int check(SlotCaps c) {
  mpoa::finish_caps(c);
  const int slots = 3;
  // hipMalloc returns 256-byte aligned pools; so do these (never touched: carve_pools only
  // computes addresses)
  struct Buf {
    void* p;
    explicit Buf(size_t n) : p(std::malloc(n + 256)) {}
    ~Buf() { std::free(p); }
  };
  Buf ints((size_t)slots * c.ints_per_slot * 4), bytes((size_t)slots * c.bytes_per_slot),
      dp((size_t)slots * 3 * c.dp * 4);
  if (!ints.p || !bytes.p || !dp.p) return fail(c, 0, "host allocation failed");
  auto up256 = [](void* p) {
    return reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(p) + 255) & ~uintptr_t(255));
  };
  int* ip = static_cast<int*>(up256(ints.p));
  uint8_t* bp = static_cast<uint8_t*>(up256(bytes.p));
  int* dpp = static_cast<int*>(up256(dp.p));
  for (int s = 0; s < slots; s++) {
    const Arena A = mpoa::carve_pools(c, ip, bp, dpp, s);
    int* i0 = ip + (int64_t)s * c.ints_per_slot;
    uint8_t* b0 = bp + (int64_t)s * c.bytes_per_slot;
    if (A.rowrec != i0) return fail(c, s, "row records do not open the slot");
    if (A.int_end > i0 + c.ints_per_slot) return fail(c, s, "int words past the slot");
    if (A.int_end - i0 > mpoa::slot_ints(c.nodes, c.edges, c.qmax, c.path))
      return fail(c, s, "int words past slot_ints");
    if (A.byte_end > b0 + c.bytes_per_slot) return fail(c, s, "bytes past the slot");
    if (A.byte_end - b0 > mpoa::slot_bytes(c.nodes, c.edges, c.qmax))
      return fail(c, s, "bytes past slot_bytes");
    if (!aligned16(A.rowrec) || !aligned16(A.qseq) || !aligned16(A.qsh))
      return fail(c, s, "vector-accessed array not 16-byte aligned");
    if (!aligned16(A.H) || !aligned16(A.D) || !aligned16(A.I))
      return fail(c, s, "DP matrix not 16-byte aligned");
    if (A.I + c.dp > dpp + (int64_t)(s + 1) * 3 * c.dp) return fail(c, s, "DP past the slot");
    // longest query: m = qmax + 1, col_size = ceil((m + 1) / Q) * Q, then 2*Q more bytes
    const int m = c.qmax + 1;
    const int qbytes = (m + 1 + 7) / 8 * 8 + 16;
    if (A.qseq + qbytes > A.qsh) return fail(c, s, "query codes overlap the shifted query");
    if (A.qsh + qbytes > A.byte_end) return fail(c, s, "shifted query past the slot");
    if (A.res_base + c.qmax + 2 > A.qseq) return fail(c, s, "result bases overlap the query");
    if (A.cons_id + c.nodes > A.cons_rank) return fail(c, s, "consensus arrays overlap");
  }
  return 0;
}

int main() {
  const int nodes[] = {2, 64, 100, 4096, 9000, 70000, 131072};
  const int edges[] = {4, 64, 200, 8192, 300000};
  const int qmax[] = {1, 7, 8, 15, 150, 1500, 2000, 16569, 20001};
  const int path[] = {0, 1, 333, 200000};
  const int64_t dps[] = {64, 4096, 1 << 20};
  int bad = 0;
  long n = 0;
  for (int N : nodes)
    for (int E : edges)
      for (int q : qmax)
        for (int P : path)
          for (int64_t D : dps) {
            SlotCaps c{};
            // the library rounds nodes, edges and dp to multiples of 64 (make_caps)
            c.nodes = (N + 63) / 64 * 64;
            c.edges = (E + 63) / 64 * 64;
            c.qmax = q;
            c.path = P;
            c.dp = (D + 63) / 64 * 64;
            bad |= check(c);
            n++;
            if (bad) return 1;
          }
  std::printf("ok %ld capacity sets\n", n);
  return 0;
}
