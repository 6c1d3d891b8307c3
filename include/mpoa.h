/* mpoa: MI355X (gfx950) partial-order alignment, C ABI.
 *
 * Drop-in for minipoa's (NCl3-lhd/POA) alignment path. Every entry point names the upstream
 * interface it replaces. Plain pointers and sizes only; the library owns its device memory
 * through an mpoa_ctx, callers own every host buffer they pass.
 *
 * Scores are minipoa's int32 arithmetic (src/align.cpp:21-22); a batched DP call whose values
 * fit int16 stores its H, D and I matrices as int16 (checked on the device; a call that does not
 * fit is re-run with int32 storage). Results are bit-identical to minipoa on the same inputs and
 * flags.
 */
#ifndef MPOA_H
#define MPOA_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* minipoa's flags (src/main.cpp:14-35, struct para_t in include/parameter.h:5-19), taken
 * as typed on the command line and normalised as initPara does
 * (src/handle_input.cpp:45-77). */
typedef struct mpoa_params {
  int32_t match, mismatch, gap_open, gap_ext; /* -M -X -O -E (defaults 2 -4 -4 -2) */
  int32_t band_b, band_f;                     /* -b -f (defaults 100 40) */
  int32_t ab_band;                            /* -B */
  int32_t seeding;                            /* -S */
  int32_t k, mm_w;                            /* -k -w (defaults 19 10) */
  int32_t poa_w;                              /* -W; 0 = flag not given */
  int32_t result;                             /* -r: 0 consensus, 1 row-column MSA, 2 GFA;
                                                 3 (no minipoa counterpart): the -r 1 MSA,
                                                 then a consensus row, ">consensus sequence",
                                                 each consensus base in its node's column */
  int32_t progressive;                        /* -p: reads added in minipoa's guide-tree order
                                                 (minimizer.cpp:195-386); outputs keep input
                                                 order. Not with result 2: minipoa's GFA writer
                                                 indexes its paths past their ends then
                                                 (graph.cpp:369,423) */
} mpoa_params;

/* Error codes. minipoa terminates the process in these cases; the library returns. */
enum {
  MPOA_OK = 0,
  MPOA_E_ARG = -1,
  MPOA_E_BAND = -2,         /* backtrack left the band (minipoa: exit(0), align.cpp:1584) */
  MPOA_E_BACKTRACK = -3,    /* no move found (minipoa: "backtrack error", exit(1), :1730) */
  MPOA_E_OUT_CAP = -4,      /* caller buffer too small; the size needed is reported */
  MPOA_E_EMPTY_WINDOW = -5, /* -S sub-window ends before it begins, in ranks (minipoa:
                               SIGSEGV) or in the read (minipoa: std::length_error abort) */
  MPOA_E_OOM = -6,          /* device allocation failed (minipoa: throw, mem_alloc_utils.cpp:27) */
  MPOA_E_DEVICE = -7,       /* HIP runtime error, or no GPU */
  MPOA_E_OVERFLOW = -8,     /* a window outgrew its pre-sized arena: internal to the other
                               entry points (they re-run it with exact sizes and never return
                               this); reported by mpoa_align_reads_noalloc */
  MPOA_E_SCRATCH = -9       /* mpoa_align_reads_noalloc: the window needs more than the fixed
                               device scratch (MPOA_NOALLOC_SCRATCH_BYTES) */
};

typedef struct mpoa_ctx mpoa_ctx;

/* Opens the library on HIP device `device`. Fails with MPOA_E_DEVICE when no GPU is present:
 * there is no CPU fallback. */
int mpoa_create(int device, mpoa_ctx** out);
void mpoa_destroy(mpoa_ctx* ctx);
const char* mpoa_strerror(int rc);

/* Statistics of one DP call (one upstream poa() invocation). */
typedef struct mpoa_call_stat {
  int32_t read;         /* read index in input order */
  int32_t rows;         /* n = end_rank - beg_rank + 1 */
  int32_t qlen;
  int32_t score;        /* H[n-1][m]: what minipoa prints as "score:" under -V 1 (align.cpp:1577) */
  int64_t cells;        /* sum over rows 1..n-1 of 8*block_num (align.cpp:1446) */
  int64_t edge_cells;   /* sum over rows of in-window in-degree * 8*block_num */
  int32_t width;        /* band extent: widest row, 8*block_num */
  int32_t nodes;        /* graph nodes and edges when the call ran */
  int32_t edges;
  int32_t window;       /* window index (batch traces) */
  int64_t logical_cells; /* sum over rows 1..n-1 of Me - Ms + 1, the band before the 8-column
                            quantum (align.cpp:1383-1394) */
} mpoa_call_stat;

/* minipoa's whole main loop over one read set: for each read in input order,
 * alignment() -> graph::add_path -> graph::topsort (src/main.cpp:110-121), then the output
 * selected by params->result (main.cpp:124-126: graph::output_consensus, output_rc_msa,
 * output_gfa). Replaces the loop in src/main.cpp and its calls into src/align.cpp and
 * src/graph.cpp. The graph and the DP live on the device.
 *
 * Reads: seq with n_reads+1 offsets. name/comment (optional, NULL) are laid out the same way
 * and only used by the MSA/GFA headers. Output text (exactly minipoa's stdout) goes to
 * out[0..out_cap); *out_len receives its size (also when MPOA_E_OUT_CAP is returned).
 * stats (optional) receives one record per DP call. A call whose text does not fit (out NULL,
 * or out_cap too small) keeps the text and stats in the context until the next call: called
 * again with the same arguments (a buffer of the size learned aside), it returns them without
 * running the alignment again. */
int mpoa_align_reads(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_reads, const char* seq,
                     const int64_t* seq_off, const char* name, const int64_t* name_off,
                     const char* comment, const int64_t* comment_off, char* out,
                     int64_t out_cap, int64_t* out_len, mpoa_call_stat* stats,
                     int32_t stats_cap, int32_t* n_stats);

/* ------------------------------------------------------------------------------------
 * Call-level API: minipoa's own functions (include/align.h:15-16, include/graph.h:16-30) over a
 * graph that stays in device memory between calls, for callers that drive minipoa's per-read
 * loop themselves (src/main.cpp:110-126):
 *     for each read i:  res = alignment(...)  ->  graph::add_path(m, i, res, 1)
 *                       ->  graph::topsort(para, i + 1 == n ? 1 : 0)
 *     then graph::output_consensus / output_rc_msa / output_gfa.
 * include/minipoa_amd.hpp wraps these in minipoa's C++ names. Calls on one graph are serial.
 * ------------------------------------------------------------------------------------ */

/* res_t (include/result.h:3-13): from >= 0 an existing node; from = -1 and aligned_id >= 0 a
 * new node aligned to that column; both -1 a new node. base is the code 0..4 (A C G T N). */
typedef struct mpoa_res {
  int32_t from, base, aligned_id;
} mpoa_res;

typedef struct mpoa_graph mpoa_graph;

/* graph::init (src/graph.cpp:129-135, called at src/main.cpp:99): a graph of the source
 * (id 0) and the sink (id 1) on the context's device. params is the para_t of every later call
 * on the graph (scores, band, -S, -r). */
int mpoa_graph_create(mpoa_ctx* ctx, const mpoa_params* params, mpoa_graph** out);
void mpoa_graph_destroy(mpoa_graph* g);

/* poa(para, DAG, beg_id, end_id, qid, qseq, qlen, mpool, ab_band) (include/align.h:15,
 * src/align.cpp:1353-1737): banded DP of qseq (qlen ASCII bases or codes 0..4) against the
 * ranks [rank(beg_id), rank(end_id)] of the last topological order, then the backtrack.
 * res receives the alignment in forward order (qlen + 1 entries; the first is the sentinel
 * column's match to beg_id), *n_res its length (also when MPOA_E_OUT_CAP is returned), *score
 * H[n-1][m] (minipoa's "score:" under -V 1). The graph must have been sorted (op 0) since its
 * last add_path. Errors: MPOA_E_BAND and MPOA_E_BACKTRACK where minipoa exits (align.cpp:1584,
 * :1730), MPOA_E_EMPTY_WINDOW for rank(end_id) < rank(beg_id) or qlen < 0. */
int mpoa_poa(mpoa_graph* g, int32_t beg_id, int32_t end_id, const char* qseq, int32_t qlen,
             int32_t ab_band, mpoa_res* res, int64_t cap, int64_t* n_res, int32_t* score);

/* minimizer_t (include/minimizer.h:9-32) over the read set, built as src/main.cpp:84-99 does;
 * needed by mpoa_alignment when params->seeding is set. Host only. */
typedef struct mpoa_seeder mpoa_seeder;
int mpoa_seeder_create(const mpoa_params* params, int32_t n_reads, const char* seq,
                       const int64_t* seq_off, mpoa_seeder** out);
void mpoa_seeder_destroy(mpoa_seeder* s);
/* minimizer_t::get_guide_tree (minimizer.cpp:195-386) and its ord: ord[i] is the read added
 * i-th, minipoa's guide-tree order when params->progressive was set, else input order. Writes
 * up to cap entries; returns the number of reads. */
int32_t mpoa_seeder_order(mpoa_seeder* s, int32_t* ord, int32_t cap);

/* alignment(para, DAG, mm, rid, qseq, qlen, mpool) (include/align.h:16,
 * src/align.cpp:1741-1828): the trivial path of a graph without bases, the whole-graph poa(),
 * or with -S the read split at its anchors on the consensus, sub-window poa() calls and the
 * anchor bases spliced between them (the graph is sorted first when it is not, as minipoa
 * does). mm may be NULL without seeding. res / cap / n_res as for mpoa_poa. */
int mpoa_alignment(mpoa_graph* g, mpoa_seeder* mm, int32_t rid, const char* qseq, int32_t qlen,
                   mpoa_res* res, int64_t cap, int64_t* n_res);

/* graph::add_path(m, seq_id, res, 1) (src/graph.cpp:170-235, called at src/main.cpp:118): new
 * nodes and edges of one read's alignment; seq_id is the read's index, and reads are added in
 * order 0, 1, 2, ... (the MSA and GFA rows follow it). Entries are checked on the host (node
 * ids below the node count, bases 0..4): MPOA_E_ARG otherwise. Marks the graph unsorted. */
int mpoa_graph_add_path(mpoa_graph* g, int32_t seq_id, const mpoa_res* res, int64_t n_res);

/* read_gfa (src/handle_input.cpp:444-495, minipoa's -i, called at src/main.cpp:86) into a
 * graph that holds only the source and the sink: S lines add their bases as nodes (in file
 * order), P lines add their paths as reads 0, 1, ... (one read per line, no bases of its own),
 * other lines are skipped; then the op 0 sort. Only forward ('+') paths are read (minipoa's
 * own GFA output has no others): a '-' segment, an unknown or duplicated segment, or a segment
 * without a sequence is MPOA_E_ARG, where minipoa exits. names receives the path names, one per
 * line; *names_len their size (also with MPOA_E_OUT_CAP), *n_paths their number. Reads added
 * afterwards continue the numbering; MSA rows of the paths print under their names. */
int mpoa_graph_read_gfa(mpoa_graph* g, const char* text, int64_t len, char* names,
                        int64_t names_cap, int64_t* names_len, int32_t* n_paths);

/* graph::topsort(para, op) (src/graph.cpp:21-128, called at src/main.cpp:120): op 0, the
 * LIFO topological order with hlen/tlen (and, with -S, build_consensus); op 1, the order of
 * aligned groups (MSA columns). */
int mpoa_graph_topsort(mpoa_graph* g, int32_t op);

/* graph::output_consensus (result 0), output_rc_msa (1) or output_gfa (2), or the MSA with a
 * consensus row (3); rid_to_ord (optional, as output_rc_msa's): read i prints the MSA row of
 * the read added at position rid_to_ord[i] (not with GFA)
 * (src/graph.cpp:237-433, called at src/main.cpp:124-126) over the reads added so far, as the
 * text minipoa prints; names / comments as for mpoa_align_reads. The MSA of a graph not sorted
 * since its last add_path is taken after an op 1 sort, as output_rc_msa does. *out_len
 * receives the size (also with MPOA_E_OUT_CAP); out NULL sizes only. */
int mpoa_graph_output(mpoa_graph* g, int32_t result, const char* name, const int64_t* name_off,
                      const char* comment, const int64_t* comment_off, const int32_t* rid_to_ord,
                      char* out, int64_t out_cap, int64_t* out_len);

/* The graph as host arrays (node_t fields, include/node.h:5-68): node count and in-edge count,
 * then per node id its base, par_id and aligned_node[5]; in- and out-edge CSR in list order
 * with weights; and by rank the node id, hlen and tlen of the last op 0 sort (rank_to_id may
 * be shorter after op 1). Pass NULL arrays to size. */
int mpoa_graph_export(mpoa_graph* g, int32_t* n_nodes, int32_t* n_edges, uint8_t* base,
                      int32_t* in_off, int32_t* in_src, int32_t* in_w, int32_t* out_off,
                      int32_t* out_dst, int32_t* out_w, int32_t* rank_to_id, int32_t* hlen,
                      int32_t* tlen, int32_t* par_id, int32_t* aligned);

/* Allocation-free form of mpoa_align_reads. No counterpart in minipoa: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). One window, no seeding (params->seeding must be 0), consensus (-r 0) or
 * row-column MSA (-r 1) output; the same kernels and results as mpoa_align_reads.
 *
 * Inputs and outputs are device buffers the caller owns: d_seq holds the reads' bases (ASCII),
 * d_seq_off their n_reads + 1 offsets; seq_off is the same offsets in host memory (they size
 * the scratch). d_out (out_cap bytes) receives the consensus bases (result 0) or the n_reads
 * MSA rows back to back, each of the same length L (result 1): minipoa's text without its
 * header lines. d_res (two int64 in device memory) receives the window's status (an MPOA_E_*
 * code, 0 on success) and the output length in bytes.
 *
 * Every other byte the call uses (graph, DP matrices, queue and result records) lies in one
 * fixed device array of MPOA_NOALLOC_SCRATCH_BYTES per device, held until the process ends;
 * nothing is allocated, freed or copied by the host. A window whose capacity estimate does not
 * fit the array is rejected with MPOA_E_SCRATCH before any launch; one that outgrows the
 * estimate on the device reports MPOA_E_OVERFLOW in d_res (mpoa_align_reads runs it). Output
 * longer than out_cap reports MPOA_E_OVERFLOW too. Asynchronous: the work is enqueued on
 * `stream` (a hipStream_t; NULL: the context's stream) and d_res is valid once it completes.
 * Every call on a device shares that device's array, which keeps the last call's contents: the
 * kernels clear what must start empty, and calls that overlap on one device race. */
#define MPOA_NOALLOC_SCRATCH_BYTES (256u << 20)
int mpoa_align_reads_noalloc(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_reads,
                             const char* d_seq, const int64_t* d_seq_off, const int64_t* seq_off,
                             char* d_out, int64_t out_cap, int64_t* d_res, void* stream);

/* ------------------------------------------------------------------------------------
 * Read sets: readFile (src/handle_input.cpp:13-43: kseq_read over gzread, FASTA or FASTQ,
 * plain or gzip) as a library call, for one file or many. Each file is one window; files are
 * parsed on n_threads host threads (0: one per hardware thread) and packed into the arrays
 * mpoa_align_reads and mpoa_batch_prepare take. Names, comments and bases follow kseq's rules
 * exactly (the name up to the first white space, the comment the rest of the header line,
 * sequence lines joined as they are); a FASTQ record whose quality is cut short ends its file
 * with the records before it, as readFile's loop stops there. A file that cannot be opened is
 * MPOA_E_ARG (minipoa: "Error opening file").
 * ------------------------------------------------------------------------------------ */
typedef struct mpoa_reads mpoa_reads;
int mpoa_reads_load(const char* const* paths, int32_t n_files, int32_t n_threads,
                    mpoa_reads** out);
/* One window from a text in memory. */
int mpoa_reads_parse(const char* text, int64_t len, mpoa_reads** out);
void mpoa_reads_free(mpoa_reads* r);
/* The packed arrays, valid until mpoa_reads_free (any pointer may be NULL): n_windows and
 * win_off (n_windows + 1 read offsets), n_reads, seq with seq_off (n_reads + 1 byte offsets),
 * name/name_off, comment/comment_off laid out the same way. */
int mpoa_reads_view(const mpoa_reads* r, int32_t* n_windows, const int32_t** win_off,
                    int32_t* n_reads, const char** seq, const int64_t** seq_off,
                    const char** name, const int64_t** name_off, const char** comment,
                    const int64_t** comment_off);

/* ------------------------------------------------------------------------------------
 * Batched windows. minipoa has no batch interface: it aligns one read set per process
 * (src/main.cpp:110-121). A batch is many independent read sets ("windows"), each the
 * equivalent of one minipoa run, all processed by one persistent device launch.
 * ------------------------------------------------------------------------------------ */
typedef struct mpoa_batch mpoa_batch;

typedef struct mpoa_batch_stats {
  int64_t windows;       /* windows in the batch */
  int64_t dp_calls;      /* DP calls made (reads after the first of each window) */
  int64_t cells;         /* sum of 8*block_num over all DP rows */
  int64_t edge_cells;
  int64_t reruns;        /* windows re-run after an arena overflow */
  double device_ms;      /* device time of the last execute (all launches and copies) */
  double kernel_ms;      /* pipeline-kernel time of the last execute, HIP events on the
                            stream the kernel runs on */
  int32_t launches;      /* pipeline launches in the last execute (1 + re-run passes) */
  int32_t slots;         /* resident workgroups of the first launch */
  int64_t pool_bytes;    /* device memory held by the per-slot arenas */
  int32_t graph_launches; /* graph replays in the last execute (0: eager launches) */
  int32_t graph_builds;   /* graphs built over the batch's life (a rebuild follows any change
                             of a launch's arguments) */
  int64_t logical_cells;  /* sum over all DP rows of Me - Ms + 1 (cells counts 8*block_num) */
  int64_t narrow_cells;   /* cells of the DP calls whose matrices were stored as int16 (the rest
                             as int32, minipoa's width; both give the same results) */
} mpoa_batch_stats;

/* Uploads window w = reads [win_off[w], win_off[w+1]) of seq/seq_off, sizes the per-window
 * arenas and allocates device memory. size_hint (optional, n_windows entries, may be NULL)
 * are predicted DP cells per window, used to order the work queue and to pre-size the DP
 * arenas; any misprediction is caught on the device and that window re-run exactly.
 * With params->seeding (-S) every window runs in one workgroup of a persistent launch that
 * seeds each read itself (minimizers, anchors, chains on the consensus so far) and runs the
 * read's sub-windows in turn; tile, schedule, graph replay and call traces do not apply, and
 * a window past a capacity of that launch runs alone as mpoa_align_reads would run it. */
int mpoa_batch_prepare(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_windows,
                       const int32_t* win_off, const char* seq, const int64_t* seq_off,
                       const double* size_hint, mpoa_batch** out);
/* Tile shape of the DP rows: `threads` per window workgroup (64 = one wavefront; 64, 128, 256
 * or 512) and `cols` DP columns per thread per pass (1, 2, 4 or 8). Not every pair is built;
 * returns MPOA_E_ARG for one that is not. Chosen by the autotuner (poa_amd/autotune.py). */
int mpoa_batch_set_tile(mpoa_batch* b, int32_t threads, int32_t cols);
/* Launch split: windows whose longest read is at most small_qmax, whose band is at most
 * small_width columns and whose graph is predicted to stay within small_nodes nodes run on
 * the small tile; of the others' launch, move_top (-1: half a window per CU) of the costliest
 * small-tile windows join it. Defaults 1500, 360, -1, no node limit. */
int mpoa_batch_set_schedule(mpoa_batch* b, int32_t small_qmax, int32_t small_width,
                            int32_t move_top, int32_t small_nodes);
/* Runs the whole per-window pipeline on the device for every window; inputs are already
 * resident. Synchronous. Returns MPOA_OK when every window finished, else the status of the
 * first window that did not (minipoa's own failures: MPOA_E_BAND, MPOA_E_BACKTRACK,
 * MPOA_E_EMPTY_WINDOW); every window's status is in mpoa_batch_window_stats and the finished
 * windows' texts are in mpoa_batch_results either way. */
int mpoa_batch_execute(mpoa_batch* b);
/* Copies the per-window output text (minipoa's stdout for that window) to out; window w is
 * out[out_off[w] .. out_off[w+1]). */
int mpoa_batch_results(mpoa_batch* b, char* out, int64_t out_cap, int64_t* out_off,
                       mpoa_batch_stats* stats);
void mpoa_batch_free(mpoa_batch* b);

/* Per-window diagnostics of the last execute: final graph size, DP cells and device time per
 * pipeline phase (DP plan, DP rows, backtrack, add_path, topological sort, output; then the
 * sort's two parts: the order itself and hlen/tlen), in ticks
 * of the device wall clock (*tick_hz ticks per second). */
typedef struct mpoa_window_stat {
  int32_t status, nodes, edges, launch;  /* launch: 0 small, 1 medium, 2 re-run, 3 wide tile */
  int64_t cells, edge_cells;
  int64_t ticks[8];
  int64_t t_begin, t_end;                /* device wall clock at the window's start and end */
  int64_t logical_cells;
  int64_t narrow_cells;                  /* cells of its DP calls stored as int16 */
} mpoa_window_stat;
int mpoa_batch_window_stats(mpoa_batch* b, mpoa_window_stat* out, int32_t cap, double* tick_hz);

/* Scheduling predictions, one record of `stride` (>= 4) floats per window: predicted cost
 * (DP cells summed over the window's calls), DP cells of its costliest call, its widest band
 * row (8*block_num) and its final graph node count. They order the work queue (costliest
 * first), choose each window's tile, and size the per-slot graph and DP arenas. They never
 * change a result: a window that outgrows a predicted capacity stops with an overflow flag on
 * the device and is re-run with worst-case capacities at the reference band. */
int mpoa_batch_set_predictions(mpoa_batch* b, const float* pred, int32_t stride);

/* Execution form: MPOA_MODE_BATCHED (default) runs every window in persistent launches with a
 * device work queue; MPOA_MODE_PLAIN launches one workgroup per window, one window after
 * another in input order (the comparison form). Results are identical. */
enum { MPOA_MODE_BATCHED = 0, MPOA_MODE_PLAIN = 1 };
int mpoa_batch_set_mode(mpoa_batch* b, int32_t mode);

/* Test hook (no counterpart in minipoa): runs every read of the batch through the fallback
 * forms the library otherwise takes only for unusual graphs, so that tests can hold each of them
 * to the same output. Bits: MPOA_FORM_SORT_HBM, the sort walk over a CSR in HBM with atomic
 * degree counts (graphs too big for LDS); MPOA_FORM_SORT_LANE, the one-lane sort (a ready stack
 * past 512 nodes); MPOA_FORM_BT_LANE, the one-lane backtrack (a row with more than 63
 * predecessors); MPOA_FORM_ADD_LANE, the one-lane add_path (a repeated edge);
 * MPOA_FORM_CONS_ROUNDS, the -S consensus settling every block in rounds (a rank with more than
 * eight predecessors in its block); MPOA_FORM_WIDE, int32 storage of the DP matrices for every
 * call (a call whose values do not fit int16: minipoa's own width); MPOA_FORM_HOST_SEED, -S
 * reads seeded on the host between launches (a window that outgrows the device seeding's
 * capacities, or ties among more than 64 equal minimizer or anchor keys); MPOA_FORM_BT_RANKS, the
 * wave backtrack's runs of match moves along consecutive ranks instead of along each row's
 * heaviest predecessor (a call whose query and graph do not both fit the LDS left after the
 * DP). 0 (the default) restores
 * the normal choice. Results do not change; only speed does. mpoa_set_forms sets the forms of
 * the batches the context creates afterwards, mpoa_align_reads' included. */
enum {
  MPOA_FORM_SORT_HBM = 1, MPOA_FORM_SORT_LANE = 2, MPOA_FORM_BT_LANE = 4, MPOA_FORM_ADD_LANE = 8,
  MPOA_FORM_CONS_ROUNDS = 16, MPOA_FORM_WIDE = 32, MPOA_FORM_HOST_SEED = 64,
  MPOA_FORM_BT_RANKS = 128
};
int mpoa_batch_set_forms(mpoa_batch* b, int32_t forms);
int mpoa_set_forms(mpoa_ctx* ctx, int32_t forms);
/* Diagnostics (no counterpart in minipoa): how the last seeded (-S) mpoa_align_reads call of
 * the context was seeded: 1 on the device (every read seeded by the workgroup that committed the
 * read before it, no host round trip); 0 on the host by choice (MPOA_FORM_HOST_SEED, fewer than
 * two reads, -k above 28 or -w above 255); above 1 on the host after the device form stopped,
 * the value naming why as bits: 2 the graph outgrew its estimate, 4 a read's minimizers, 8 the
 * consensus length, 16 its minimizers, 32 equal keys among more than 64 consensus minimizers,
 * 64 more than 4,096 anchors, 128 / 256 / 512 equal keys among more than 64 anchors / chain ends
 * / chains (minipoa's radix sort may order them otherwise), 1024 more than 32 sub-windows, 2048
 * their DP pool, 4096 anchor-base splices; -1 before any seeded call. */
int mpoa_last_seeding(const mpoa_ctx* ctx);

/* Graph replay of the batched pass (default on): the concurrent persistent launches, each with
 * its work-queue reset, are captured once into a hipGraph and replayed by every later
 * mpoa_batch_execute whose launch arguments are unchanged; off, they are enqueued one by one.
 * Overflow re-runs are enqueued eagerly in both. Results are identical. No counterpart in
 * minipoa, which has no device launches. */
int mpoa_batch_set_graph(mpoa_batch* b, int32_t enable);

/* Per-DP-call traces: when enabled before mpoa_batch_execute, every DP call records its
 * geometry (rows, query length, band extent, cells, graph size) and score. mpoa_batch_traces
 * copies them in window order (window w's calls are its reads 1..n-1); *n receives the
 * number of records. */
int mpoa_batch_set_traces(mpoa_batch* b, int32_t enable);
int mpoa_batch_traces(mpoa_batch* b, mpoa_call_stat* out, int64_t cap, int64_t* n);

/* Per-launch record of the last execute: tile, resident workgroups, windows and DP cells the
 * launch processed, and its duration from HIP events recorded on the stream it ran on. */
typedef struct mpoa_launch_stat {
  int32_t launch;                        /* 0 small, 1 medium, 2 re-run, 3 wide tile */
  int32_t threads, cols, slots;
  int64_t windows, cells, edge_cells;
  double ms;
} mpoa_launch_stat;
/* Writes up to cap records to out and their number to *n. */
int mpoa_batch_launch_stats(mpoa_batch* b, mpoa_launch_stat* out, int32_t cap, int32_t* n);

/* -S seeding on the host: the anchors of read rid of a window on the consensus cons (ASCII,
 * cons_len bases), as the sub-window split consumes them (minimizer.cpp:537-570 ->
 * lchain.cpp:158-386); out receives (x, y) pairs: x = consensus position of the anchor's last
 * base, y = span << 32 | read position (bit 40: band-mode mark). *n = anchor count. */
int mpoa_seed_anchors(const mpoa_params* params, int32_t n_reads, const char* seq,
                      const int64_t* seq_off, int32_t rid, const char* cons, int32_t cons_len,
                      uint64_t* out, int64_t cap, int64_t* n);

/* The same anchors computed on the device of ctx (the device seeding: minimizers, sorts,
 * anchors, local and global chains as the -S windows run them, on one workgroup), in the same
 * format. MPOA_E_OVERFLOW past the device form's capacities (more than 4,096 anchors, 8,192
 * minimizers in one sort), where mpoa_seed_anchors still answers. */
int mpoa_seed_anchors_device(mpoa_ctx* ctx, const mpoa_params* params, int32_t n_reads,
                             const char* seq, const int64_t* seq_off, int32_t rid,
                             const char* cons, int32_t cons_len, uint64_t* out, int64_t cap,
                             int64_t* n);

/* ------------------------------------------------------------------------------------
 * Learned scheduling. minipoa sizes its DP buffer by growing it on demand per thread
 * (src/mem_alloc_utils.cpp:11-40, used at src/align.cpp:1395-1409) and has no scheduler.
 * Here a small network predicts each DP call's band extent and cell count; the predictions,
 * summed per window, go to mpoa_batch_set_predictions. Results never depend on them.
 * ------------------------------------------------------------------------------------ */

/* Eight features per DP call (window w, read r >= 1), in the order of mpoa_batch_traces:
 * read length, graph node and edge count estimates, chain coverage and anchor count against
 * the window's first read (minimizers, k = 11, w = 5), the previous read's static band
 * extent, the window's error-rate estimate, and the score width in bytes. X receives
 * 8 floats per call; *n_rows the number of calls (pass X = NULL to size). */
int mpoa_sched_features(const mpoa_params* params, int32_t n_windows, const int32_t* win_off,
                        const char* seq, const int64_t* seq_off, float* X, int64_t cap,
                        int64_t* n_rows);

/* The network: 8 -> 64 -> 64 -> 2 with ReLU, FP16 weights, FP32 accumulation. Parameters
 * are FP16 bit patterns in the order W1[8][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]
 * (mpoa_mlp_n_params() of them). Training keeps FP32 master weights (Adam) on the device;
 * the loss is the mean squared error over samples and outputs, halved. All arrays below are
 * host arrays: X has 8 floats per sample, T and Y 2. */
typedef struct mpoa_mlp mpoa_mlp;
int mpoa_mlp_create(int device, uint64_t seed, mpoa_mlp** out);
void mpoa_mlp_destroy(mpoa_mlp* m);
int32_t mpoa_mlp_n_params(void);
int mpoa_mlp_get_weights(mpoa_mlp* m, uint16_t* out, int32_t cap);
int mpoa_mlp_set_weights(mpoa_mlp* m, const uint16_t* in, int32_t n);
int mpoa_mlp_forward(mpoa_mlp* m, const float* X, int64_t n, float* Y);
/* Gradient of the loss over the n samples with respect to every parameter (FP32, same
 * order as the weights), and the loss itself; the weights are not changed. */
int mpoa_mlp_gradients(mpoa_mlp* m, const float* X, const float* T, int64_t n, float* grad,
                       float* loss);
/* Adam (lr, beta 0.9/0.999) over `epochs` shuffled passes of minibatches of `batch`
 * samples; epoch_loss (optional) receives the mean minibatch loss of each epoch. */
int mpoa_mlp_train(mpoa_mlp* m, const float* X, const float* T, int64_t n, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, float* epoch_loss);
const char* mpoa_mlp_error(void);

#ifdef __cplusplus
}
#endif
#endif
