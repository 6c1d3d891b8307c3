/* CPU oracle for minipoa's banded partial-order alignment path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg may load it. The product (poa_amd/) never links or calls it.
 *
 * It restates, in plain scalar C++, the semantics of minipoa (NCl3-lhd/POA):
 *   poa()                 src/align.cpp:1353-1737   banded DP + backtrack
 *   alignment()           src/align.cpp:1741-1828   dispatcher, -S sub-windows
 *   graph::add_path       src/graph.cpp:170-235     + node_t::add_in_adj/add_out_adj include/node.h:23-61
 *   graph::topsort        src/graph.cpp:21-128
 *   graph::build_consensus/output_consensus  src/graph.cpp:294-357
 *   graph::output_rc_msa  src/graph.cpp:237-251
 *   graph::output_gfa     src/graph.cpp:358-433
 *   main() loop           src/main.cpp:110-126
 *   seeding (-S)          src/minimizer.cpp:60-135,404-442,493-570, src/lchain.cpp:9-386
 * Each function in oracle.cpp cites the lines it follows.
 */
#ifndef POA_ORACLE_H
#define POA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* minipoa's flags (src/main.cpp:14-35). Values are taken as typed on the command line and
 * normalised as initPara does (src/handle_input.cpp:45-77). */
typedef struct orc_params {
  int32_t match, mismatch, gap_open, gap_ext; /* -M -X -O -E (defaults 2 -4 -4 -2) */
  int32_t band_b, band_f;                     /* -b -f (defaults 100 40) */
  int32_t ab_band;                            /* -B */
  int32_t seeding;                            /* -S */
  int32_t k, mm_w;                            /* -k -w (defaults 19 10) */
  int32_t poa_w;                              /* -W; 0 = flag not given */
  int32_t result;                             /* -r: 0 consensus, 1 row-column MSA, 2 GFA,
                                                 3 the MSA with a consensus row */
  int32_t progressive;                        /* -p: guide-tree read order */
} orc_params;

/* One poa() call: the DP geometry and its optimum. */
typedef struct orc_call_stat {
  int32_t read;          /* read index in input order */
  int32_t rows;          /* n = end_rank - beg_rank + 1 */
  int32_t qlen;          /* query length of this call (m = qlen + 1) */
  int32_t score;         /* H[n-1][m], the value minipoa prints as "score:" under -V 1 */
  int32_t ab_band;       /* 1 if the adaptive band was used */
  int32_t max_width;     /* max over rows of 8*block_num */
  int64_t cells;         /* sum over rows 1..n-1 of 8*block_num (align.cpp:1446) */
  int64_t edge_cells;    /* sum over rows 1..n-1 of (in-window in-degree)*8*block_num */
  int64_t logical_cells; /* sum over rows 1..n-1 of Me-Ms+1 */
  int32_t graph_nodes;   /* graph size when the call ran */
  int32_t graph_edges;
} orc_call_stat;

/* Error codes shared with the product library's C ABI. */
enum {
  ORC_OK = 0,
  ORC_E_ARG = -1,
  ORC_E_BAND = -2,       /* backtrack left the band: minipoa exits with status 0 */
  ORC_E_BACKTRACK = -3,  /* no move found: minipoa prints "backtrack error" and exits 1 */
  ORC_E_OUT_CAP = -4,    /* output buffer too small; *out_len holds the size needed */
  ORC_E_EMPTY_WINDOW = -5/* -S sub-window with end rank before begin rank (minipoa segfaults) */
};

/* minipoa's whole main loop on one set of reads (src/main.cpp:84-126).
 * Reads are the concatenation `seq` with n_reads+1 offsets. name/comment are optional (NULL),
 * laid out the same way. Writes the text minipoa prints to stdout into out[0..out_cap).
 * Per-poa()-call statistics go to stats[0..stats_cap) (may be NULL); *n_stats is the number
 * of calls made. Returns ORC_OK or an error code. */
int orc_window(const orc_params* p, int32_t n_reads, const char* seq, const int64_t* seq_off,
               const char* name, const int64_t* name_off, const char* comment,
               const int64_t* comment_off, char* out, int64_t out_cap, int64_t* out_len,
               orc_call_stat* stats, int32_t stats_cap, int32_t* n_stats);

/* Independent windows, one orc_window each, spread over `threads` OpenMP threads
 * (0 = all). Window w owns reads [win_off[w], win_off[w+1]). Output text of window w is
 * out[out_off[w] .. out_off[w+1]); out_off has n_windows+1 entries. Statistics of every
 * poa() call go to stats (may be NULL) in window order; calls_off (n_windows+1, may be NULL)
 * receives the per-window prefix. Returns ORC_OK or the first error met. */
int orc_batch(const orc_params* p, int32_t n_windows, const int32_t* win_off, const char* seq,
              const int64_t* seq_off, int32_t threads, char* out, int64_t out_cap,
              int64_t* out_off, orc_call_stat* stats, int64_t stats_cap, int64_t* calls_off);

/* Builds the graph of reads[0..n_reads) exactly as orc_window does and exports it:
 * node bases (n_nodes), in-edge CSR (in_off n_nodes+1, in_src, in_w in node.in order),
 * out-edge CSR likewise, rank (node id by rank), hlen/tlen by rank, par_id, and aligned_node
 * (5 per node). Pass NULL arrays to query sizes only. */
int orc_graph(const orc_params* p, int32_t n_reads, const char* seq, const int64_t* seq_off,
              int32_t* n_nodes, int32_t* n_edges, uint8_t* base, int32_t* in_off,
              int32_t* in_src, int32_t* in_w, int32_t* out_off, int32_t* out_dst,
              int32_t* out_w, int32_t* rank_to_id, int32_t* hlen, int32_t* tlen,
              int32_t* par_id, int32_t* aligned);

/* Exports the full H/D/I matrices of the DP that aligns read n_reads-1 to the graph built
 * from reads[0..n_reads-1) (no seeding). Row r (local rank) holds columns
 * [8*Bs[r], 8*Be[r]), stored at offset row_off[r] (n_rows+1 entries) of H, D and I.
 * Pass NULL matrix pointers to query n_rows and total sizes first. */
int orc_dp_matrix(const orc_params* p, int32_t n_reads, const char* seq, const int64_t* seq_off,
                  int32_t* n_rows, int64_t* total, int32_t* Bs, int32_t* Be, int64_t* row_off,
                  int32_t* H, int32_t* D, int32_t* I, int32_t* score);

/* Anchors of read rid on the consensus cons (ASCII): the global chain as alignment()'s -S
 * branch consumes it (x, y pairs into out). */
int orc_anchors(const orc_params* p, int32_t n_reads, const char* seq, const int64_t* seq_off,
                int32_t rid, const char* cons, int32_t cons_len, uint64_t* out, int64_t cap,
                int64_t* n);

/* Call-level mirror of minipoa's functions on one graph, the checker of the library's
 * call-level API. orc_dag_new: graph::init (src/graph.cpp:129-135) and, with seeding, the
 * minimizer_t of the read set (src/minimizer.cpp); res arrays hold 3 int32 per res_t entry
 * (from, base code, aligned_id; include/result.h:3-13). orc_dag_poa: poa()
 * (src/align.cpp:1353-1737); orc_dag_alignment: alignment() (:1741-1828); orc_dag_add_path:
 * graph::add_path(m, seq_id, res, 1) (src/graph.cpp:170-235); orc_dag_topsort:
 * graph::topsort (:21-128); orc_dag_output: the -r 0/1/2 text (:237-433) over the reads
 * added so far; orc_dag_export: the graph as orc_graph exports it. */
typedef struct orc_dag orc_dag;
orc_dag* orc_dag_new(const orc_params* p, int32_t n_reads, const char* seq,
                         const int64_t* seq_off);
void orc_dag_free(orc_dag* g);
/* The read set's order (-p: minimizer_t::get_guide_tree, minimizer.cpp:195-386; else input
 * order): ord[i] is the read added i-th. Returns the number of reads. */
int orc_dag_order(orc_dag* g, int32_t* ord, int32_t cap);
int orc_dag_poa(orc_dag* g, int32_t beg_id, int32_t end_id, const char* q, int32_t qlen,
                  int32_t ab_band, int32_t* res, int64_t cap, int64_t* n_res, int32_t* score);
int orc_dag_alignment(orc_dag* g, int32_t rid, const char* q, int32_t qlen, int32_t* res,
                        int64_t cap, int64_t* n_res);
int orc_dag_add_path(orc_dag* g, int32_t seq_id, const int32_t* res, int64_t n_res);
/* The scores of the DP calls completed on g so far, in order (minipoa's -V 1 "score:" lines);
 * returns their number. */
int orc_dag_scores(orc_dag* g, int32_t* out, int32_t cap);
int orc_dag_topsort(orc_dag* g, int32_t op);
/* read_gfa (src/handle_input.cpp:444-495) into a new graph: the GFA's '+' paths become reads
 * 0, 1, ...; names receives their names, one per line. */
int orc_dag_read_gfa(orc_dag* g, const char* text, int64_t len, char* names,
                     int64_t names_cap, int64_t* names_len, int32_t* n_paths);
int orc_dag_output(orc_dag* g, int32_t result, const char* name, const int64_t* name_off,
                     const char* comment, const int64_t* comment_off, char* out, int64_t out_cap,
                     int64_t* out_len);
int orc_dag_export(orc_dag* g, int32_t* n_nodes, int32_t* n_edges, uint8_t* base,
                     int32_t* in_off, int32_t* in_src, int32_t* in_w, int32_t* out_off,
                     int32_t* out_dst, int32_t* out_w, int32_t* rank_to_id, int32_t* hlen,
                     int32_t* tlen, int32_t* par_id, int32_t* aligned);

#ifdef __cplusplus
}
#endif
#endif
