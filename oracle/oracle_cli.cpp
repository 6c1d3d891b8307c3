// Command-line driver of the CPU oracle (TEST INFRASTRUCTURE ONLY): minipoa's flags
// (src/main.cpp:14-35) and stdout, so its output can be diffed against the upstream binary.
// With -V 1 it prints one "score:<H[n-1][m]>" line per poa() call to stderr, the line
// minipoa prints at src/align.cpp:1577.
#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"

namespace {

struct Read {
  std::string name, comment, seq;
};

// FASTA/FASTQ reader with kseq's field split: name up to the first blank, comment = rest of
// the header line, sequence lines concatenated (src/handle_input.cpp:13-43).
bool read_fastx(const char* path, std::vector<Read>& reads) {
  gzFile fp = gzopen(path, "r");
  if (!fp) return false;
  std::string data;
  char buf[1 << 16];
  int n;
  while ((n = gzread(fp, buf, sizeof buf)) > 0) data.append(buf, n);
  gzclose(fp);
  size_t i = 0, L = data.size();
  auto line_end = [&](size_t at) {
    size_t e = data.find('\n', at);
    return e == std::string::npos ? L : e;
  };
  while (i < L && data[i] != '>' && data[i] != '@') i = line_end(i) + 1;
  while (i < L) {
    bool fq = data[i] == '@';
    size_t e = line_end(i);
    std::string head = data.substr(i + 1, e - i - 1);
    if (!head.empty() && head.back() == '\r') head.pop_back();
    Read r;
    size_t sp = head.find_first_of(" \t");
    if (sp == std::string::npos) {
      r.name = head;
    } else {
      r.name = head.substr(0, sp);
      r.comment = head.substr(sp + 1);
    }
    i = e + 1;
    while (i < L && data[i] != '>' && data[i] != '@' && data[i] != '+') {
      size_t le = line_end(i);
      for (size_t k = i; k < le; k++)
        if (data[k] != '\r' && data[k] != ' ' && data[k] != '\t') r.seq.push_back(data[k]);
      i = le + 1;
    }
    if (fq && i < L && data[i] == '+') {
      i = line_end(i) + 1;
      size_t got = 0;
      while (i < L && got < r.seq.size()) {
        size_t le = line_end(i);
        got += le - i;
        i = le + 1;
      }
    }
    reads.push_back(std::move(r));
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  orc_params p = {2, -4, -4, -2, 100, 40, 0, 0, 19, 10, 0, 0, 0};
  int verbose = 0;
  const char* path = nullptr;
  for (int a = 1; a < argc; a++) {
    std::string s = argv[a];
    if (s.size() < 2 || s[0] != '-') {
      path = argv[a];
      continue;
    }
    std::string key, val;
    bool has_val = false;
    if (s[1] == '-') {
      size_t eq = s.find('=');
      key = s.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
      if (eq != std::string::npos) val = s.substr(eq + 1), has_val = true;
    } else {
      key = s.substr(1, 1);
      if (s.size() > 2) val = s.substr(2), has_val = true;
    }
    auto num = [&]() -> int {
      if (!has_val) {
        if (a + 1 >= argc) {
          std::fprintf(stderr, "missing value for %s\n", s.c_str());
          std::exit(1);
        }
        val = argv[++a];
      }
      return std::atoi(val.c_str());
    };
    if (key == "M" || key == "match") p.match = num();
    else if (key == "X" || key == "mismatch") p.mismatch = num();
    else if (key == "O" || key == "gap_open") p.gap_open = num();
    else if (key == "E" || key == "gap_ext") p.gap_ext = num();
    else if (key == "b" || key == "band_b") p.band_b = num();
    else if (key == "f" || key == "band_f") p.band_f = num();
    else if (key == "B" || key == "ab_band") p.ab_band = 1;
    else if (key == "S" || key == "seeding") p.seeding = 1;
    else if (key == "p" || key == "progressive_poa") p.progressive = 1;
    else if (key == "W" || key == "poa_w") p.poa_w = num();
    else if (key == "k" || key == "k_mer") p.k = num();
    else if (key == "w" || key == "mm_w") p.mm_w = num();
    else if (key == "r" || key == "result") p.result = num();
    else if (key == "V" || key == "verbose") verbose = num();
    else if (key == "t" || key == "thread" || key == "s" || key == "sample_num") (void)num();
    else {
      std::fprintf(stderr, "unsupported option %s\n", s.c_str());
      return 1;
    }
  }
  std::vector<Read> reads;
  if (path && !read_fastx(path, reads)) {
    std::fprintf(stderr, "error read file: cannot open %s\n", path);
    return 1;
  }
  std::string seq, name, comment;
  std::vector<int64_t> so{0}, no{0}, co{0};
  for (const Read& r : reads) {
    seq += r.seq;
    so.push_back((int64_t)seq.size());
    name += r.name;
    no.push_back((int64_t)name.size());
    comment += r.comment;
    co.push_back((int64_t)comment.size());
  }
  int nr = (int)reads.size();
  int64_t need = 0;
  int32_t n_calls = 0;
  std::vector<orc_call_stat> st(1 << 16);
  int rc = orc_window(&p, nr, seq.data(), so.data(), name.data(), no.data(), comment.data(),
                      co.data(), nullptr, 0, &need, st.data(), (int32_t)st.size(), &n_calls);
  if (rc == ORC_OK) {
    std::vector<char> out(need + 1);
    rc = orc_window(&p, nr, seq.data(), so.data(), name.data(), no.data(), comment.data(),
                    co.data(), out.data(), need, &need, nullptr, 0, nullptr);
    if (rc == ORC_OK) std::fwrite(out.data(), 1, need, stdout);
  }
  if (verbose)
    for (int c = 0; c < n_calls && c < (int)st.size(); c++) std::fprintf(stderr, "score:%d\n", st[c].score);
  if (rc != ORC_OK) {
    std::fprintf(stderr, "oracle error %d\n", rc);
    return rc == ORC_E_BAND ? 0 : 1;
  }
  return 0;
}
