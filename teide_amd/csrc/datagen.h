/*
 * datagen.h -- deterministic synthetic column generators.
 *
 * Every generator is a pure function of (seed, stream, row index range): rows
 * are produced in chunks of TDG_CHUNK rows and chunk k draws from its own
 * xoshiro256** state, seeded through SplitMix64 from (seed, stream, k).  The
 * output is therefore identical for any thread count, on any host that shares
 * the C library's exp/log/sqrt.
 *
 * Distributions: Lemire's nearly-divisionless bounded integers, uniform
 * doubles from the top 53 bits, Marsaglia's polar normal, lognormal as
 * exp(mu + sigma * N), Zipf by Hormann-Derflinger rejection-inversion, and a
 * Fisher-Yates permutation (single stream; sequential by nature).
 *
 * No symbol here starts with td_, so this file links next to either the
 * reference library or the new one.
 */
#ifndef TDG_DATAGEN_H
#define TDG_DATAGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDG_CHUNK (1 << 20)

/* Uniform integers in [lo, hi] (inclusive). */
void tdg_i64_uniform(int64_t* out, int64_t n, int64_t lo, int64_t hi,
                     uint64_t seed, uint32_t stream, int nthreads);
void tdg_i32_uniform(int32_t* out, int64_t n, int32_t lo, int32_t hi,
                     uint64_t seed, uint32_t stream, int nthreads);
void tdg_u32_uniform(uint32_t* out, int64_t n, uint32_t lo, uint32_t hi,
                     uint64_t seed, uint32_t stream, int nthreads);
/* Uniform doubles in [lo, hi). */
void tdg_f64_uniform(double* out, int64_t n, double lo, double hi,
                     uint64_t seed, uint32_t stream, int nthreads);
void tdg_f64_normal(double* out, int64_t n, double mu, double sigma,
                    uint64_t seed, uint32_t stream, int nthreads);
void tdg_f64_lognormal(double* out, int64_t n, double mu, double sigma,
                       uint64_t seed, uint32_t stream, int nthreads);
/* Zipf(s) ranks in [1, n_values]; mapped to keys by tdg_zipf_key(rank). */
void tdg_i64_zipf_keys(int64_t* out, int64_t n, int64_t n_values, double s,
                       uint64_t seed, uint32_t stream, int nthreads);
int64_t tdg_zipf_key(int64_t rank);
/* A uniformly random permutation of base + [0, n). */
void tdg_i64_permutation(int64_t* out, int64_t n, int64_t base,
                         uint64_t seed, uint32_t stream);
/* out[i] = keys[u] for u uniform in [0, n_keys) (foreign keys into a key set). */
void tdg_i64_pick(int64_t* out, int64_t n, const int64_t* keys, int64_t n_keys,
                  uint64_t seed, uint32_t stream, int nthreads);

/* Scalar helpers for small per-query parameters (threshold, sizes). */
uint64_t tdg_hash_u64(uint64_t seed, uint64_t a, uint64_t b);
double   tdg_unit(uint64_t seed, uint64_t a, uint64_t b);

#ifdef __cplusplus
}
#endif

/* ------------------------------------------------------------------------
 * Workload recipes.  One definition of each benchmark configuration's
 * columns, used by the reference harness, the tests and the benchmark, so
 * all three see the same bytes for the same (sizes, seed).  Streams:
 *   1xx filter+group, 2xx join+group, 3xx high-cardinality group,
 *   4xx batched queries, 5xx served tables.
 * ------------------------------------------------------------------------ */
#ifdef __cplusplus
extern "C" {
#endif
/* Filter+group: flag i64 in {0,1}; region i64 in [0,16); amount f64 in [0,1000). */
void tdg_recipe_fg(int64_t n, uint64_t seed, int nthreads,
                   int64_t* flag, int64_t* region, double* amount);
/* Join dimension: pk = base + permutation of [0,n_dim); cat i32 in [0,n_cat). */
void tdg_recipe_dim(int64_t n_dim, int64_t n_cat, uint64_t seed, uint32_t stream_base,
                    int nthreads, int64_t* pk, int32_t* cat);
/* Join fact: fk uniform over [0,n_dim); qty i64 in [1,100]; price lognormal(3,1). */
void tdg_recipe_fact(int64_t n_fact, int64_t n_dim, uint64_t seed, uint32_t stream_base,
                     int nthreads, int64_t* fk, int64_t* qty, double* price);
/* High-cardinality group: key = zipf(s) over n_values ranks; sym u32 in
 * [0,n_sym); v_i64 in [-1e9,1e9]; v_f64 normal(0,1). */
void tdg_recipe_hc(int64_t n, int64_t n_values, double s, uint32_t n_sym, uint64_t seed,
                   int nthreads, int64_t* key, uint32_t* sym, int64_t* v_i64, double* v_f64);
#ifdef __cplusplus
}
#endif

#endif /* TDG_DATAGEN_H */
