// Host seeding for -S windows: minimizers of every read and of the current consensus,
// anchors between them, local chains and the global chain whose anchors split a read into
// independent sub-windows (minipoa's minimizer.cpp and lchain.cpp; the steps cited in
// seed.cpp). Results must match minipoa bit for bit: the sort is the same unstable MSD radix
// sort and the chain penalties are the same float expressions.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace mpoa {
namespace seed {

struct Pair {
  uint64_t x, y;
};

struct Config {
  int k = 19, w = 10;   // -k, -w
  int bw = 100;         // -b: chaining band and maximum gap
  int poa_w = 0;        // -W, 0 when not given
};

// Progressive order (-p, minipoa's guide tree, minimizer.cpp:195-386): read indices in the
// order they are added to the graph. Reads are sketched with the seeding's k and w; pairs are
// compared by their shared minimizers (length-weighted Jaccard similarity); the most similar
// pair opens the order, then each next read is the one most similar in sum to those already
// in it. Fewer than two reads, or none with a minimizer: input order.
std::vector<int> guide_order(const Config& c, int n_reads, const uint8_t* codes,
                             const int64_t* off);

// One window's reads, sketched once; the consensus is sketched per call.
class Seeder {
 public:
  // reads: n codes arrays (0..3 bases, 4 other) as concatenation + offsets.
  void init(const Config& c, int n_reads, const uint8_t* codes, const int64_t* off);
  // Anchors of read `rid` (length qlen) on the consensus `cons` (codes): the global chain's
  // picked anchors in ascending order; x = tpos (last base), y = qspan << 32 | qpos.
  std::vector<Pair> anchors(int rid, int qlen, const uint8_t* cons, int cons_len);

 private:
  Config cfg_;
  int n_reads_ = 0;
  std::vector<Pair> mins_;      // sorted per read, then the consensus
  std::vector<int64_t> start_;  // n_reads + 2
};

}  // namespace seed
}  // namespace mpoa
