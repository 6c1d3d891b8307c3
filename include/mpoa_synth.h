/* Seeded synthetic input for the benchmarks and tests: uniform ACGT templates and reads
 * derived by i.i.d. per-base substitution / insertion / deletion. Not part of minipoa's
 * interface (minipoa reads FASTA files); it produces the workloads the GPU path is measured on.
 */
#ifndef MPOA_SYNTH_H
#define MPOA_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reads of one window. seq may be NULL to query sizes (off still filled, n_reads+1 entries).
 * geo_p > 0 draws indel lengths from a geometric law with P(L=1) = geo_p; 0 means length 1.
 * Returns 0, -1 on bad arguments, -4 when cap is too small. */
int mpoa_synth_window(uint64_t seed, int32_t tlen, int32_t n_reads, double sub, double ins,
                      double del, double geo_p, char* seq, int64_t cap, int64_t* off);

/* Uneven batch of independent windows: reads per window uniform in [min_reads, max_reads];
 * template length round(exp(N(ln median_len, sigma))) clipped to [min_len, max_len]; per-window
 * error rate uniform in [err_lo, err_hi] split sub:ins:del = split_sub:split_ins:split_del
 * (split_sum is their sum). */
typedef struct mpoa_synth_batch_spec {
  int32_t min_reads, max_reads;
  double median_len, sigma;
  int32_t min_len, max_len;
  double err_lo, err_hi;
  double split_sub, split_ins, split_del, split_sum;
  double geo_p;
} mpoa_synth_batch_spec;

/* win_off: n_windows+1 read offsets; seq_off: total_reads+1 base offsets. Pass NULL buffers
 * to query total_reads and total_bases first. */
int mpoa_synth_batch(uint64_t seed, const mpoa_synth_batch_spec* spec, int32_t n_windows,
                     int32_t* win_off, char* seq, int64_t cap, int64_t* seq_off,
                     int64_t* total_reads, int64_t* total_bases);

#ifdef __cplusplus
}
#endif
#endif
