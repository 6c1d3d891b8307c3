// Read-set ingestion (include/mpoa.h, "Read sets"): minipoa's readFile (src/handle_input.cpp:13-43,
// kseq_read of include/kseq.h:181-223 over gzread) for many files at once, each file a window,
// parsed on a pool of host threads and packed into the arrays mpoa_batch_prepare and
// mpoa_align_reads take (one concatenation of bases with int64 read offsets, window offsets,
// names and comments with their offsets).
//
// The record rules are kseq's: text before the first '>' or '@' is skipped; the name runs to
// the first white-space character (space, \t, \n, \v, \f, \r) and, unless that was the
// newline, the comment is the rest of the line without one trailing \r (a comment of the single
// character \r is kept, as kseq keeps it); sequence lines are joined as they are (blanks and
// all, a trailing \r dropped) up to a line that starts with '>', '@' or '+'; empty lines are
// skipped; after '+' the rest of that line is skipped and quality lines are read until they
// hold as many characters as the sequence. A record whose quality is cut short ends the file
// (kseq_read returns -2 and readFile's loop stops, keeping the records before it).
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "host.h"
#include "mpoa.h"

using mpoa::host::g_err;

struct mpoa_reads {
  std::string seq, name, comment;
  std::vector<int64_t> seq_off{0}, name_off{0}, comment_off{0};
  std::vector<int32_t> win_off{0};
};

namespace {

struct Parsed {
  std::string seq, name, comment;
  std::vector<int64_t> seq_len, name_len, comment_len;
  std::string err;  // empty: parsed
};

inline bool blank(unsigned char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// One file's text (whole) into records, as readFile's kseq_read loop would see it.
void parse_text(const char* t, size_t n, Parsed* out) {
  size_t at = 0;
  // Appends the rest of the current line (up to, not with, the newline, which is consumed) to
  // *s, then drops one trailing \r if *s is longer than one character (ks_getuntil2 with
  // KS_SEP_LINE). Returns false at the end of the text with nothing left to read.
  auto eol = [&](size_t from) {
    const void* nl = std::memchr(t + from, '\n', n - from);
    return nl ? (size_t)(static_cast<const char*>(nl) - t) : n;
  };
  auto line_into = [&](std::string* s) {
    if (at >= n) return false;
    const size_t l0 = at;
    at = eol(at);
    s->append(t + l0, at - l0);
    if (at < n) at++;
    if (s->size() > 1 && s->back() == '\r') s->pop_back();
    return true;
  };
  std::string name, comment, seq, qual;
  out->seq.reserve(out->seq.size() + n);
  bool at_header = false;  // the header character has been read (kseq's last_char)
  for (;;) {
    if (!at_header) {
      while (at < n && t[at] != '>' && t[at] != '@') at++;
      if (at >= n) return;
      at++;
    }
    at_header = false;
    if (at >= n) return;  // nothing after the header character: no record
    name.clear();
    comment.clear();
    seq.clear();
    const size_t n0 = at;
    while (at < n && !blank((unsigned char)t[at])) at++;
    name.assign(t + n0, at - n0);
    if (at < n) {
      const char d = t[at++];  // the white-space character that ended the name
      if (d != '\n') line_into(&comment);
    }
    char c = 0;
    bool more = false;
    while (at < n) {
      c = t[at++];
      if (c == '>' || c == '+' || c == '@') {
        more = true;
        break;
      }
      if (c == '\n') continue;
      seq.push_back(c);
      line_into(&seq);
    }
    if (more && c != '+') {
      at_header = true;
    } else if (more) {
      // FASTQ: skip the rest of the '+' line, then quality lines until they are as long as
      // the sequence (at least one line is read)
      at = eol(at);
      if (at >= n) return;  // no quality string: kseq_read returns -2, readFile stops
      at++;
      qual.clear();
      while (line_into(&qual) && qual.size() < seq.size()) {
      }
      if (qual.size() != seq.size()) return;  // -2 again: the record and the file end here
    }
    out->name += name;
    out->comment += comment;
    out->seq += seq;
    out->name_len.push_back((int64_t)name.size());
    out->comment_len.push_back((int64_t)comment.size());
    out->seq_len.push_back((int64_t)seq.size());
  }
}

// The file's bytes: gzip members inflated by zlib (as gzopen/gzread do for readFile), any other
// file read as it is.
void parse_file(const char* path, Parsed* out) {
  std::string all;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    out->err = std::string("cannot open ") + path;
    return;
  }
  unsigned char magic[2] = {0, 0};
  const size_t m = std::fread(magic, 1, 2, f);
  if (m == 2 && magic[0] == 0x1f && magic[1] == 0x8b) {
    std::fclose(f);
    gzFile fp = gzopen(path, "r");
    if (!fp) {
      out->err = std::string("cannot open ") + path;
      return;
    }
    char buf[1 << 16];
    int got;
    while ((got = gzread(fp, buf, sizeof buf)) > 0) all.append(buf, (size_t)got);
    const bool bad = got < 0;
    gzclose(fp);
    if (bad) {
      out->err = std::string("read error in ") + path;
      return;
    }
  } else {
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size > 0) {
      all.resize((size_t)size);
      all.resize(std::fread(&all[0], 1, (size_t)size, f));
    }
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) {
      out->err = std::string("read error in ") + path;
      return;
    }
  }
  parse_text(all.data(), all.size(), out);
}

void append(mpoa_reads* r, const Parsed& p) {
  size_t s = 0, nm = 0, cm = 0;
  for (size_t i = 0; i < p.seq_len.size(); i++) {
    r->seq.append(p.seq, s, (size_t)p.seq_len[i]);
    r->name.append(p.name, nm, (size_t)p.name_len[i]);
    r->comment.append(p.comment, cm, (size_t)p.comment_len[i]);
    s += (size_t)p.seq_len[i];
    nm += (size_t)p.name_len[i];
    cm += (size_t)p.comment_len[i];
    r->seq_off.push_back((int64_t)r->seq.size());
    r->name_off.push_back((int64_t)r->name.size());
    r->comment_off.push_back((int64_t)r->comment.size());
  }
  r->win_off.push_back(r->win_off.back() + (int32_t)p.seq_len.size());
}

}  // namespace

int mpoa_reads_load(const char* const* paths, int32_t n_files, int32_t n_threads,
                    mpoa_reads** out) {
  if (!out || n_files < 0 || (n_files > 0 && !paths)) return MPOA_E_ARG;
  *out = nullptr;
  std::vector<Parsed> parsed((size_t)n_files);
  int T = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
  T = std::max(1, std::min(T, n_files));
  if (T == 1) {
    for (int f = 0; f < n_files; f++) parse_file(paths[f], &parsed[(size_t)f]);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < T; t++)
      pool.emplace_back([&, t] {
        for (int f = t; f < n_files; f += T) parse_file(paths[f], &parsed[(size_t)f]);
      });
    for (std::thread& th : pool) th.join();
  }
  for (const Parsed& p : parsed)
    if (!p.err.empty()) {
      g_err = p.err;
      return MPOA_E_ARG;
    }
  // Packing: each file's place in the arrays from prefix sums, then the copies on the pool.
  mpoa_reads* r = new mpoa_reads();
  const size_t F = parsed.size();
  std::vector<int64_t> rd(F + 1, 0), sb(F + 1, 0), nb(F + 1, 0), cb(F + 1, 0);
  for (size_t f = 0; f < F; f++) {
    rd[f + 1] = rd[f] + (int64_t)parsed[f].seq_len.size();
    sb[f + 1] = sb[f] + (int64_t)parsed[f].seq.size();
    nb[f + 1] = nb[f] + (int64_t)parsed[f].name.size();
    cb[f + 1] = cb[f] + (int64_t)parsed[f].comment.size();
  }
  r->seq.resize((size_t)sb[F]);
  r->name.resize((size_t)nb[F]);
  r->comment.resize((size_t)cb[F]);
  r->seq_off.assign((size_t)rd[F] + 1, 0);
  r->name_off.assign((size_t)rd[F] + 1, 0);
  r->comment_off.assign((size_t)rd[F] + 1, 0);
  r->win_off.resize(F + 1);
  for (size_t f = 0; f <= F; f++) r->win_off[f] = (int32_t)rd[f];
  auto place = [&](size_t f) {
    const Parsed& p = parsed[f];
    std::memcpy(&r->seq[0] + sb[f], p.seq.data(), p.seq.size());
    std::memcpy(&r->name[0] + nb[f], p.name.data(), p.name.size());
    std::memcpy(&r->comment[0] + cb[f], p.comment.data(), p.comment.size());
    int64_t s = sb[f], nm = nb[f], cm = cb[f];
    for (size_t i = 0; i < p.seq_len.size(); i++) {
      s += p.seq_len[i];
      nm += p.name_len[i];
      cm += p.comment_len[i];
      const size_t k = (size_t)rd[f] + i + 1;
      r->seq_off[k] = s;
      r->name_off[k] = nm;
      r->comment_off[k] = cm;
    }
  };
  if (T == 1) {
    for (size_t f = 0; f < F; f++) place(f);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < T; t++)
      pool.emplace_back([&, t] {
        for (size_t f = (size_t)t; f < F; f += (size_t)T) place(f);
      });
    for (std::thread& th : pool) th.join();
  }
  *out = r;
  return MPOA_OK;
}

int mpoa_reads_parse(const char* text, int64_t len, mpoa_reads** out) {
  if (!out || len < 0 || (len > 0 && !text)) return MPOA_E_ARG;
  Parsed p;
  parse_text(text, (size_t)len, &p);
  mpoa_reads* r = new mpoa_reads();
  append(r, p);
  *out = r;
  return MPOA_OK;
}

void mpoa_reads_free(mpoa_reads* r) { delete r; }

int mpoa_reads_view(const mpoa_reads* r, int32_t* n_windows, const int32_t** win_off,
                    int32_t* n_reads, const char** seq, const int64_t** seq_off,
                    const char** name, const int64_t** name_off, const char** comment,
                    const int64_t** comment_off) {
  if (!r) return MPOA_E_ARG;
  if (n_windows) *n_windows = (int32_t)r->win_off.size() - 1;
  if (win_off) *win_off = r->win_off.data();
  if (n_reads) *n_reads = (int32_t)r->seq_off.size() - 1;
  if (seq) *seq = r->seq.data();
  if (seq_off) *seq_off = r->seq_off.data();
  if (name) *name = r->name.data();
  if (name_off) *name_off = r->name_off.data();
  if (comment) *comment = r->comment.data();
  if (comment_off) *comment_off = r->comment_off.data();
  return MPOA_OK;
}
