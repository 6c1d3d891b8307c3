// minipoa_amd.hpp: header-only C++ facade with minipoa's own names (include/parameter.h,
// include/result.h, include/graph.h, include/minimizer.h, include/align.h of NCl3-lhd/POA) over
// the C ABI of include/mpoa.h. A program written against minipoa's alignment path keeps its
// loop (src/main.cpp:110-126):
//
//     minipoa_amd::graph* DAG = new minipoa_amd::graph();
//     DAG->init(para);
//     minipoa_amd::minimizer_t* mm = new minimizer_t(para, seqs);
//     for (int i = 0; i < seqs.size(); i++) {
//       std::vector<res_t> res = alignment(para, DAG, mm, i, seqs[i].seq.c_str(),
//                                          seqs[i].seq.size());
//       DAG->add_path(para->m, i, res, 1);
//       DAG->topsort(para, i + 1 == seqs.size());
//     }
//     DAG->output_consensus();   // or output_rc_msa / output_gfa
//
// and runs it on the GPU: the graph lives in device memory and every call is a device launch.
// Errors that make minipoa exit (band violation, backtrack error) or crash (an empty -S
// sub-window) throw minipoa_amd::error here. The device is chosen with set_device() before the
// first graph is created (default 0).
#pragma once
#include <cstdint>
#include <fstream>
#include <iostream>
#include <iterator>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpoa.h"

namespace minipoa_amd {

struct error : std::runtime_error {
  int code;
  error(int rc, const char* what) : std::runtime_error(what), code(rc) {}
};

inline void check(int rc) {
  if (rc != MPOA_OK) throw error(rc, mpoa_strerror(rc));
}

// The process's device context (one per process, as minipoa has one process per run).
inline int& device_index() {
  static int d = 0;
  return d;
}
inline void set_device(int d) { device_index() = d; }
inline mpoa_ctx* context() {
  static std::unique_ptr<mpoa_ctx, void (*)(mpoa_ctx*)> ctx(
      [] {
        mpoa_ctx* c = nullptr;
        check(mpoa_create(device_index(), &c));
        return c;
      }(),
      mpoa_destroy);
  return ctx.get();
}

// para_t (include/parameter.h:5-19): the fields the alignment path reads, with minipoa's
// defaults (src/main.cpp:18-33). Penalties may be given as typed (positive) or as initPara
// leaves them (negative).
struct para_t {
  int thread = 1;
  int m = 5;  // alphabet size incl. N
  int match = 2, mismatch = 4, gap_open1 = 4, gap_ext1 = 2;
  int b = 100, f = 40;
  bool ab_band = false;
  int k = 19, mm_w = 10;
  bool enable_seeding = false;
  int bw = 1000, poa_w = 0;
  bool progressive_poa = false;
  int result = 0;  // 0 consensus, 1 row-column MSA, 2 GFA, 3 MSA with a consensus row
  int verbose = 0;

  mpoa_params c_params() const {
    mpoa_params p{};
    p.match = match;
    p.mismatch = mismatch;
    p.gap_open = gap_open1;
    p.gap_ext = gap_ext1;
    p.band_b = b;
    p.band_f = f;
    p.ab_band = ab_band;
    p.seeding = enable_seeding;
    p.k = k;
    p.mm_w = mm_w;
    p.poa_w = poa_w;
    p.result = result;
    p.progressive = progressive_poa;
    return p;
  }
};

// res_t (include/result.h:3-13)
struct res_t {
  int from;
  unsigned char base;
  int aligned_id = -1;
  res_t(int _from, unsigned char _base, int _aligned_id = -1)
      : from(_from), base(_base), aligned_id(_aligned_id) {}
};

// seq_t (include/sequence.h)
struct seq_t {
  std::string name, comment, seq, qual;
};

namespace detail {
inline std::vector<res_t> to_res(const std::vector<mpoa_res>& v) {
  std::vector<res_t> r;
  r.reserve(v.size());
  for (const mpoa_res& e : v) r.emplace_back(e.from, (unsigned char)e.base, e.aligned_id);
  return r;
}
template <class F>
std::vector<res_t> res_call(F f, int64_t guess) {
  std::vector<mpoa_res> buf((size_t)(guess > 0 ? guess : 1));
  int64_t n = 0;
  int rc = f(buf.data(), (int64_t)buf.size(), &n);
  if (rc == MPOA_E_OUT_CAP) {
    buf.resize((size_t)n);
    rc = f(buf.data(), n, &n);
  }
  check(rc);
  buf.resize((size_t)n);
  return to_res(buf);
}
inline void lists(const std::vector<seq_t>& seqs, std::string* name, std::vector<int64_t>* name_off,
                  std::string* comment, std::vector<int64_t>* comment_off) {
  name_off->assign(1, 0);
  comment_off->assign(1, 0);
  for (const seq_t& s : seqs) {
    *name += s.name;
    *comment += s.comment;
    name_off->push_back((int64_t)name->size());
    comment_off->push_back((int64_t)comment->size());
  }
}
}  // namespace detail

// minimizer_t (include/minimizer.h:9-32): the read set's minimizers for -S seeding, and the
// read order (input order until get_guide_tree).
struct minimizer_t {
  mpoa_seeder* h = nullptr;
  std::vector<int> ord, rid_to_ord;
  minimizer_t(para_t* para, const std::vector<seq_t>& seqs) {
    std::string all;
    std::vector<int64_t> off(1, 0);
    for (const seq_t& s : seqs) {
      all += s.seq;
      off.push_back((int64_t)all.size());
    }
    const mpoa_params p = para->c_params();
    check(mpoa_seeder_create(&p, (int32_t)seqs.size(), all.data(), off.data(), &h));
    for (int i = 0; i < (int)seqs.size(); i++) {
      ord.push_back(i);
      rid_to_ord.push_back(i);
    }
  }
  ~minimizer_t() { mpoa_seeder_destroy(h); }
  // minimizer_t::get_guide_tree (minimizer.cpp:195-386): with -p, the order reads are added in
  void get_guide_tree(para_t* para) {
    if (!para->progressive_poa) return;
    const int32_t n = mpoa_seeder_order(h, nullptr, 0);
    std::vector<int32_t> o((size_t)(n > 0 ? n : 1));
    mpoa_seeder_order(h, o.data(), n);
    for (int i = 0; i < n; i++) {
      ord[i] = o[i];
      rid_to_ord[o[i]] = i;
    }
  }
  minimizer_t(const minimizer_t&) = delete;
  minimizer_t& operator=(const minimizer_t&) = delete;
};

// graph (include/graph.h:9-31), held in device memory.
struct graph {
  mpoa_graph* h = nullptr;
  graph() = default;
  ~graph() { mpoa_graph_destroy(h); }
  graph(const graph&) = delete;
  graph& operator=(const graph&) = delete;

  // graph::init(para) (src/graph.cpp:129-135): the source and the sink.
  void init(para_t* para) {
    mpoa_graph_destroy(h);
    h = nullptr;
    const mpoa_params p = para->c_params();
    check(mpoa_graph_create(context(), &p, &h));
  }
  // graph::add_path (src/graph.cpp:170-235); the sink is always node 1, as main.cpp passes it.
  void add_path(int /*para_m*/, int seq_id, const std::vector<res_t>& res, int /*sink_id*/ = -1) {
    std::vector<mpoa_res> r;
    r.reserve(res.size());
    for (const res_t& e : res) r.push_back(mpoa_res{e.from, e.base, e.aligned_id});
    check(mpoa_graph_add_path(h, seq_id, r.data(), (int64_t)r.size()));
  }
  // graph::topsort(para, op) (src/graph.cpp:21-128)
  void topsort(const para_t* /*para*/, int op) { check(mpoa_graph_topsort(h, op)); }

  // read_gfa's work on a GFA text (src/handle_input.cpp:444-495): segments and '+' paths into
  // this new graph; the paths become the first reads, returned as seq_t with their names and
  // no bases.
  std::vector<seq_t> read_gfa_text(const std::string& text) {
    std::string names(text.size() + 1, '\0');  // a name and its newline fit in its P line
    int64_t len = 0;
    int32_t n = 0;
    check(mpoa_graph_read_gfa(h, text.data(), (int64_t)text.size(), &names[0],
                              (int64_t)names.size(), &len, &n));
    std::vector<seq_t> seqs((size_t)n);
    size_t at = 0;
    for (int i = 0; i < n; i++) {
      const size_t nl = names.find('\n', at);
      seqs[(size_t)i].name = names.substr(at, nl - at);
      at = nl + 1;
    }
    return seqs;
  }

  // The outputs as text, and printed to std::cout as minipoa prints them (graph.cpp:237-433).
  std::string text(int result, const std::vector<seq_t>* seqs = nullptr,
                   const std::vector<int>* rid_to_ord = nullptr) {
    std::string name, comment;
    std::vector<int64_t> name_off, comment_off;
    if (seqs) detail::lists(*seqs, &name, &name_off, &comment, &comment_off);
    int64_t n = 0;
    const char* nm = seqs ? name.data() : nullptr;
    const int64_t* no = seqs ? name_off.data() : nullptr;
    const char* cm = seqs ? comment.data() : nullptr;
    const int64_t* co = seqs ? comment_off.data() : nullptr;
    std::vector<int32_t> map;
    if (rid_to_ord) {
      bool identity = true;
      for (size_t i = 0; i < rid_to_ord->size(); i++) identity &= (*rid_to_ord)[i] == (int)i;
      if (!identity) map.assign(rid_to_ord->begin(), rid_to_ord->end());
    }
    const int32_t* ro = map.empty() ? nullptr : map.data();
    check(mpoa_graph_output(h, result, nm, no, cm, co, ro, nullptr, 0, &n));
    std::string out((size_t)n, '\0');
    check(mpoa_graph_output(h, result, nm, no, cm, co, ro, &out[0], n, &n));
    return out;
  }
  void output_consensus(bool /*needCoverages*/ = false) { std::cout << text(0); }
  void output_rc_msa(para_t* /*para*/, const std::vector<int>& rid_to_ord,
                     const std::vector<seq_t>& seqs) {
    std::cout << text(1, &seqs, &rid_to_ord);
  }
  void output_gfa(const std::vector<int>& rid_to_ord, const std::vector<seq_t>& seqs) {
    std::cout << text(2, &seqs, &rid_to_ord);
  }
};

// readFile (src/handle_input.cpp:13-43): the FASTA/FASTQ(.gz) reads of `path` appended to seqs,
// parsed by the library with kseq's rules. A file that cannot be opened throws, as minipoa's
// readFile throws std::runtime_error.
inline void readFile(std::vector<seq_t>& seqs, const char* path) {
  mpoa_reads* r = nullptr;
  if (mpoa_reads_load(&path, 1, 1, &r) != MPOA_OK)
    throw std::runtime_error(std::string("Error opening file ") + path);
  int32_t n = 0;
  const char *seq, *name, *comment;
  const int64_t *seq_off, *name_off, *comment_off;
  mpoa_reads_view(r, nullptr, nullptr, &n, &seq, &seq_off, &name, &name_off, &comment,
                  &comment_off);
  for (int i = 0; i < n; i++) {
    seq_t s;
    s.name.assign(name + name_off[i], (size_t)(name_off[i + 1] - name_off[i]));
    s.comment.assign(comment + comment_off[i], (size_t)(comment_off[i + 1] - comment_off[i]));
    s.seq.assign(seq + seq_off[i], (size_t)(seq_off[i + 1] - seq_off[i]));
    seqs.push_back(std::move(s));
  }
  mpoa_reads_free(r);
}

// read_gfa (src/handle_input.cpp:444-495, minipoa's -i): the plain-text GFA at `path` into DAG,
// fresh from init; returns the paths as the read set's first seq_t entries. A file that cannot
// be opened gives no entries, as in minipoa (:452-455).
inline std::vector<seq_t> read_gfa(para_t* /*para*/, graph* DAG, const char* path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return {};
  const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  return DAG->read_gfa_text(text);
}

// poa() (include/align.h:15, src/align.cpp:1353-1737); qid and mpool are not used.
inline std::vector<res_t> poa(const para_t* /*para*/, const graph* DAG, int beg_id, int end_id,
                              int /*qid*/, const char* qseq, int qlen, void* /*mpool*/ = nullptr,
                              bool ab_band = true) {
  return detail::res_call(
      [&](mpoa_res* out, int64_t cap, int64_t* n) {
        return mpoa_poa(DAG->h, beg_id, end_id, qseq, qlen, ab_band, out, cap, n, nullptr);
      },
      (int64_t)qlen + 1);
}

// alignment() (include/align.h:16, src/align.cpp:1741-1828); mpool is not used.
inline std::vector<res_t> alignment(const para_t* /*para*/, graph* DAG, minimizer_t* mm, int rid,
                                    const char* qseq, int qlen, void* /*mpool*/ = nullptr) {
  return detail::res_call(
      [&](mpoa_res* out, int64_t cap, int64_t* n) {
        return mpoa_alignment(DAG->h, mm ? mm->h : nullptr, rid, qseq, qlen, out, cap, n);
      },
      (int64_t)qlen + 1);
}

}  // namespace minipoa_amd
