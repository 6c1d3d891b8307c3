// minipoa_amd: minipoa's command line (src/main.cpp:14-35) over the device library.
//
//   minipoa_amd [options] reads.fa[.gz]            one read set, minipoa's stdout
//   minipoa_amd [options] w1.fa w2.fa ...           many read sets as one batch (no -S):
//                                                   their outputs, in argument order
//
// Options as minipoa: -M -X -O -E scores, -b -f band, -B adaptive band, -S seeding with -k -w
// -W, -p progressive order, -r 0|1|2 consensus / row-column MSA / GFA (3: MSA with a consensus
// row), -i graph.gfa[.gz] aligns the reads to an existing graph (src/main.cpp:86-126; one read
// set, call by call through the call-level API, -p ignored as minipoa ignores it), -V 1 prints
// "score:<n>" per DP call to stderr (src/align.cpp:1577; not with -i), -t is accepted and
// ignored (one device). Exit status follows minipoa: 0 after a band violation
// (align.cpp:1584-1588), 1 after a backtrack failure (:1730-1731).
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mpoa.h"

namespace {

struct ReadSet {
  std::string seq, name, comment;
  std::vector<int64_t> seq_off{0}, name_off{0}, comment_off{0};
  int n() const { return (int)seq_off.size() - 1; }
};

// A whole file, plain or gzip.
bool slurp(const char* path, std::string* all) {
  gzFile fp = gzopen(path, "r");
  if (!fp) return false;
  char buf[1 << 16];
  int got;
  while ((got = gzread(fp, buf, sizeof buf)) > 0) all->append(buf, (size_t)got);
  gzclose(fp);
  return true;
}

// Window w of a loaded read set (mpoa_reads_load), copied out with offsets from 0.
ReadSet window_of(const mpoa_reads* r, int w) {
  const int32_t* win_off;
  const char *seq, *name, *comment;
  const int64_t *seq_off, *name_off, *comment_off;
  mpoa_reads_view(r, nullptr, &win_off, nullptr, &seq, &seq_off, &name, &name_off, &comment,
                  &comment_off);
  ReadSet rs;
  const int r0 = win_off[w], r1 = win_off[w + 1];
  for (int i = r0; i < r1; i++) {
    rs.seq.append(seq + seq_off[i], (size_t)(seq_off[i + 1] - seq_off[i]));
    rs.name.append(name + name_off[i], (size_t)(name_off[i + 1] - name_off[i]));
    rs.comment.append(comment + comment_off[i], (size_t)(comment_off[i + 1] - comment_off[i]));
    rs.seq_off.push_back((int64_t)rs.seq.size());
    rs.name_off.push_back((int64_t)rs.name.size());
    rs.comment_off.push_back((int64_t)rs.comment.size());
  }
  return rs;
}

void usage() {
  std::fprintf(stderr,
               "usage: minipoa_amd [options] reads.fa [more.fa ...]\n"
               "  -M INT match (2)      -X INT mismatch (-4)  -O INT gap open (-4)\n"
               "  -E INT gap extend (-2) -b INT band b (100)  -f INT band f (40)\n"
               "  -B adaptive band      -S seeding           -k INT k-mer (19)\n"
               "  -w INT window (10)    -W INT sub-window     -r INT 0 consensus, 1 MSA, 2 GFA, 3 MSA + consensus row\n"
               "  -p progressive order  -V INT verbosity      -t INT threads (ignored)\n"
               "  -i FILE align to the graph in this GFA\n");
}

int exit_status(int rc) {
  if (rc == MPOA_OK) return 0;
  if (rc == MPOA_E_BAND) return 0;
  return 1;
}

// minipoa -i (src/main.cpp:86-126): the GFA's paths are the first reads, then rs's reads are
// aligned one by one, each added and the graph sorted, and the output printed. A GFA file that
// cannot be opened gives an empty graph, as in minipoa (handle_input.cpp:452-455).
int incremental(mpoa_ctx* ctx, const mpoa_params& p, const char* gfa_path, const ReadSet& rs) {
  std::string gfa;
  if (!slurp(gfa_path, &gfa)) gfa.clear();
  mpoa_graph* g = nullptr;
  mpoa_seeder* mm = nullptr;
  int rc = mpoa_graph_create(ctx, &p, &g);
  std::string names(gfa.size() + 1, '\0');  // a path name and its newline fit in its P line
  int64_t names_len = 0;
  int32_t n_paths = 0;
  if (rc == MPOA_OK)
    rc = mpoa_graph_read_gfa(g, gfa.data(), (int64_t)gfa.size(), &names[0],
                             (int64_t)names.size(), &names_len, &n_paths);
  // the read set: paths (a name, no comment, no bases), then the FASTA's reads
  std::string name(names, 0, (size_t)names_len), comment = rs.comment, seq = rs.seq;
  name.erase(std::remove(name.begin(), name.end(), '\n'), name.end());
  std::vector<int64_t> name_off{0}, comment_off, seq_off;
  for (size_t a = 0, k = 0; (int)k < n_paths; k++) {
    const size_t nl = names.find('\n', a);
    name_off.push_back(name_off.back() + (int64_t)(nl - a));
    a = nl + 1;
  }
  const int64_t path_name_len = name_off.back();
  name += rs.name;
  for (int r = 1; r <= rs.n(); r++) name_off.push_back(path_name_len + rs.name_off[r]);
  comment_off.assign((size_t)n_paths, 0);
  seq_off.assign((size_t)n_paths, 0);
  for (int r = 0; r <= rs.n(); r++) {
    comment_off.push_back(rs.comment_off[r]);
    seq_off.push_back(rs.seq_off[r]);
  }
  const int n = n_paths + rs.n();
  if (rc == MPOA_OK) rc = mpoa_seeder_create(&p, n, seq.data(), seq_off.data(), &mm);
  std::vector<mpoa_res> res;
  for (int i = n_paths; i < n && rc == MPOA_OK; i++) {
    const int qlen = (int)(seq_off[i + 1] - seq_off[i]);
    res.resize((size_t)(2 * qlen + 8));
    int64_t got = 0;
    rc = mpoa_alignment(g, mm, i, seq.data() + seq_off[i], qlen, res.data(),
                        (int64_t)res.size(), &got);
    if (rc == MPOA_E_OUT_CAP) {
      res.resize((size_t)got);
      rc = mpoa_alignment(g, mm, i, seq.data() + seq_off[i], qlen, res.data(), got, &got);
    }
    if (rc == MPOA_OK) rc = mpoa_graph_add_path(g, i, res.data(), got);
    if (rc == MPOA_OK) rc = mpoa_graph_topsort(g, i + 1 == n ? 1 : 0);
  }
  std::string text;
  int64_t len = 0;
  if (rc == MPOA_OK)
    rc = mpoa_graph_output(g, p.result, name.data(), name_off.data(), comment.data(),
                           comment_off.data(), nullptr, nullptr, 0, &len);
  if (rc == MPOA_OK) {
    text.resize((size_t)std::max<int64_t>(len, 1));
    rc = mpoa_graph_output(g, p.result, name.data(), name_off.data(), comment.data(),
                           comment_off.data(), nullptr, &text[0], len, &len);
  }
  if (rc == MPOA_OK) std::fwrite(text.data(), 1, (size_t)len, stdout);
  else std::fprintf(stderr, "minipoa_amd: %s\n", mpoa_strerror(rc));
  mpoa_seeder_destroy(mm);
  mpoa_graph_destroy(g);
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  mpoa_params p = {2, -4, -4, -2, 100, 40, 0, 0, 19, 10, 0, 0, 0};
  int verbose = 0;
  std::vector<const char*> files;
  const char* gfa = nullptr;
  for (int i = 1; i < argc; i++) {
    const char* s = argv[i];
    if (s[0] != '-' || s[1] == '\0') {
      files.push_back(s);
      continue;
    }
    const char opt = s[1];
    auto value = [&]() -> int {
      if (s[2]) return std::atoi(s + 2);
      if (i + 1 >= argc) {
        std::fprintf(stderr, "option -%c needs a value\n", opt);
        std::exit(1);
      }
      return std::atoi(argv[++i]);
    };
    switch (opt) {
      case 'M': p.match = value(); break;
      case 'X': p.mismatch = value(); break;
      case 'O': p.gap_open = value(); break;
      case 'E': p.gap_ext = value(); break;
      case 'b': p.band_b = value(); break;
      case 'f': p.band_f = value(); break;
      case 'B': p.ab_band = 1; break;
      case 'S': p.seeding = 1; break;
      case 'p': p.progressive = 1; break;
      case 'k': p.k = value(); break;
      case 'w': p.mm_w = value(); break;
      case 'W': p.poa_w = value(); break;
      case 'r': p.result = value(); break;
      case 'V': verbose = value(); break;
      case 't': (void)value(); break;
      case 'i':
        if (s[2]) gfa = s + 2;
        else if (i + 1 < argc) gfa = argv[++i];
        else {
          std::fprintf(stderr, "option -i needs a value\n");
          return 1;
        }
        break;
      case 'h': usage(); return 0;
      default:
        std::fprintf(stderr, "unknown option -%c\n", opt);
        usage();
        return 1;
    }
  }
  if (files.empty()) {
    usage();
    return 1;
  }
  // readFile for every file (src/handle_input.cpp:13-43), parsed by the library on host threads
  mpoa_reads* loaded = nullptr;
  if (mpoa_reads_load(files.data(), (int32_t)files.size(), 0, &loaded) != MPOA_OK) {
    std::fprintf(stderr, "error read file: %s\n", mpoa_strerror(MPOA_E_ARG));
    return 1;
  }
  std::vector<ReadSet> sets(files.size());
  for (size_t f = 0; f < files.size(); f++) sets[f] = window_of(loaded, (int)f);
  mpoa_reads_free(loaded);
  mpoa_ctx* ctx = nullptr;
  int rc = mpoa_create(0, &ctx);
  if (rc != MPOA_OK) {
    std::fprintf(stderr, "minipoa_amd: %s\n", mpoa_strerror(rc));
    return 1;
  }
  if (gfa) {
    if (sets.size() != 1) {
      std::fprintf(stderr, "minipoa_amd: -i takes one read set\n");
      mpoa_destroy(ctx);
      return 1;
    }
    p.progressive = 0;  // minipoa builds no guide tree with -i (main.cpp:101)
    rc = incremental(ctx, p, gfa, sets[0]);
    mpoa_destroy(ctx);
    return exit_status(rc);
  }
  if (sets.size() == 1 || p.seeding) {
    // minipoa's own run, one read set at a time
    for (const ReadSet& rs : sets) {
      int64_t need = 0;
      int32_t n_calls = 0;
      rc = mpoa_align_reads(ctx, &p, rs.n(), rs.seq.data(), rs.seq_off.data(), rs.name.data(),
                            rs.name_off.data(), rs.comment.data(), rs.comment_off.data(), nullptr,
                            0, &need, nullptr, 0, &n_calls);
      std::vector<mpoa_call_stat> calls((size_t)std::max(n_calls, 1));
      std::string out((size_t)std::max<int64_t>(need, 1), '\0');
      if (rc == MPOA_OK)
        rc = mpoa_align_reads(ctx, &p, rs.n(), rs.seq.data(), rs.seq_off.data(),
                              rs.name.data(), rs.name_off.data(), rs.comment.data(),
                              rs.comment_off.data(), &out[0], need, &need, calls.data(),
                              (int32_t)calls.size(), &n_calls);
      if (verbose)
        for (int c = 0; c < n_calls && c < (int)calls.size(); c++)
          std::fprintf(stderr, "score:%d\n", calls[c].score);
      if (rc != MPOA_OK) {
        std::fprintf(stderr, "minipoa_amd: %s\n", mpoa_strerror(rc));
        mpoa_destroy(ctx);
        return exit_status(rc);
      }
      std::fwrite(out.data(), 1, (size_t)need, stdout);
    }
    mpoa_destroy(ctx);
    return 0;
  }
  // Many read sets: one batch, one device pass.
  std::string seq;
  std::vector<int64_t> seq_off{0};
  std::vector<int32_t> win_off{0};
  for (const ReadSet& rs : sets) {
    for (int r = 0; r < rs.n(); r++) {
      seq.append(rs.seq, (size_t)rs.seq_off[r], (size_t)(rs.seq_off[r + 1] - rs.seq_off[r]));
      seq_off.push_back((int64_t)seq.size());
    }
    win_off.push_back(win_off.back() + rs.n());
  }
  mpoa_batch* b = nullptr;
  rc = mpoa_batch_prepare(ctx, &p, (int32_t)sets.size(), win_off.data(), seq.data(),
                          seq_off.data(), nullptr, &b);
  if (rc == MPOA_OK) rc = mpoa_batch_execute(b);
  std::vector<int64_t> off(sets.size() + 1);
  mpoa_batch_stats st;
  std::string text;
  if (rc == MPOA_OK) rc = mpoa_batch_results(b, nullptr, 0, off.data(), &st);
  if (rc == MPOA_OK) {
    text.resize((size_t)std::max<int64_t>(off.back(), 1));
    rc = mpoa_batch_results(b, &text[0], off.back(), off.data(), &st);
  }
  if (rc != MPOA_OK) {
    std::fprintf(stderr, "minipoa_amd: %s\n", mpoa_strerror(rc));
  } else {
    std::fwrite(text.data(), 1, (size_t)off.back(), stdout);
  }
  if (b) mpoa_batch_free(b);
  mpoa_destroy(ctx);
  return exit_status(rc);
}
