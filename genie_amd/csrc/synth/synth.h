/* Deterministic synthetic workloads for the match-count search (tests and
 * benchmarks).  Counter-based PRNG (splitmix64 of seed, stream and index), so
 * every value depends only on (seed, stream, index) and generation can be
 * split over threads without changing the data. */
#ifndef GENIE_SYNTH_H_
#define GENIE_SYNTH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* value models of a single-value table */
#define SYNTH_NORMAL 0   /* clamp(round(p0 + p1 * N(0,1)), lo, hi) */
#define SYNTH_ZIPF_PERM 1 /* per dim: pi_d(rank - 1), rank ~ Zipf(p0) over 1..n_values */
#define SYNTH_UNIFORM 2  /* U[0, n_values) */

/* rows[n][d] int32. */
int synth_rows(uint64_t seed, uint32_t n, uint32_t d, int model, double p0, double p1, int32_t lo, int32_t hi,
               uint32_t n_values, int32_t* rows);

/* Query values [q][d]: a random source row (out_src) with either per-dim
 * integer noise U{-noise..noise} clamped to [lo, hi] (redraw_p < 0) or each
 * entry redrawn U[0, n_values) with probability redraw_p. */
int synth_queries(uint64_t seed, const int32_t* rows, uint32_t n, uint32_t d, uint32_t q, int32_t noise,
                  int32_t lo, int32_t hi, double redraw_p, uint32_t n_values, int32_t* out, uint32_t* out_src);

/* Set-valued rows: lengths U[len_lo, len_hi], tokens rank-1 with rank ~
 * Zipf(s) over 1..vocab (duplicates kept).  Call with vals == NULL to get
 * row_ptr (n+1) only, then again with vals sized row_ptr[n]. */
int synth_sets(uint64_t seed, uint32_t n, uint32_t len_lo, uint32_t len_hi, uint32_t vocab, double s,
               int64_t* row_ptr, int32_t* vals);

/* Uniform integers in [lo, hi] (weights, k choices). */
int synth_uniform_int(uint64_t seed, uint64_t stream, uint64_t n, int32_t lo, int32_t hi, int32_t* out);

/* Exponential inter-arrival gaps with the given rate (Poisson arrivals), seconds. */
int synth_exponential(uint64_t seed, uint64_t n, double rate, double* out);

#ifdef __cplusplus
}
#endif
#endif
