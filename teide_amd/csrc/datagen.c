/*
 * datagen.c -- deterministic synthetic column generators (see datagen.h).
 */
#define _GNU_SOURCE
#include "datagen.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ PRNG core -- */

static inline uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; int have_spare; double spare; } rng_t;

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t* r) {          /* xoshiro256** */
    uint64_t* s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

uint64_t tdg_hash_u64(uint64_t seed, uint64_t a, uint64_t b) {
    uint64_t x = seed ^ (a * 0xD1B54A32D192ED03ull) ^ (b * 0x8CB92BA72F3D8DD7ull);
    splitmix64(&x);
    return splitmix64(&x);
}

static void rng_seed(rng_t* r, uint64_t seed, uint32_t stream, uint64_t chunk) {
    uint64_t x = tdg_hash_u64(seed, (uint64_t)stream + 1, chunk + 1);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
    r->have_spare = 0;
    r->spare = 0.0;
}

static inline double rng_unit(rng_t* r) {            /* [0, 1) */
    return (double)(rng_next(r) >> 11) * 0x1.0p-53;
}

double tdg_unit(uint64_t seed, uint64_t a, uint64_t b) {
    return (double)(tdg_hash_u64(seed, a, b) >> 11) * 0x1.0p-53;
}

/* Lemire, "Fast random integer generation in an interval" (2019). */
static inline uint64_t rng_bounded(rng_t* r, uint64_t range) {
    if (range == 0) return rng_next(r);              /* full 64-bit range */
    __uint128_t m = (__uint128_t)rng_next(r) * range;
    uint64_t l = (uint64_t)m;
    if (l < range) {
        uint64_t t = (0 - range) % range;
        while (l < t) {
            m = (__uint128_t)rng_next(r) * range;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

/* Marsaglia polar method: two normals per accepted pair. */
static inline double rng_normal(rng_t* r) {
    if (r->have_spare) { r->have_spare = 0; return r->spare; }
    double u, v, s;
    do {
        u = 2.0 * rng_unit(r) - 1.0;
        v = 2.0 * rng_unit(r) - 1.0;
        s = u * u + v * v;
    } while (s >= 1.0 || s == 0.0);
    double f = sqrt(-2.0 * log(s) / s);
    r->spare = v * f;
    r->have_spare = 1;
    return u * f;
}

/* -------------------------------------------------- chunked parallel map -- */

typedef void (*chunk_fn)(void* ctx, rng_t* r, int64_t start, int64_t end);

typedef struct {
    chunk_fn fn; void* ctx; int64_t n; uint64_t seed; uint32_t stream;
    int64_t next_chunk; pthread_mutex_t mu;
} par_t;

static void* par_worker(void* arg) {
    par_t* p = (par_t*)arg;
    int64_t nchunks = (p->n + TDG_CHUNK - 1) / TDG_CHUNK;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        int64_t k = p->next_chunk++;
        pthread_mutex_unlock(&p->mu);
        if (k >= nchunks) break;
        rng_t r;
        rng_seed(&r, p->seed, p->stream, (uint64_t)k);
        int64_t s = k * TDG_CHUNK, e = s + TDG_CHUNK;
        if (e > p->n) e = p->n;
        p->fn(p->ctx, &r, s, e);
    }
    return NULL;
}

static void run_chunks(chunk_fn fn, void* ctx, int64_t n, uint64_t seed,
                       uint32_t stream, int nthreads) {
    if (n <= 0) return;
    par_t p = { fn, ctx, n, seed, stream, 0, PTHREAD_MUTEX_INITIALIZER };
    int64_t nchunks = (n + TDG_CHUNK - 1) / TDG_CHUNK;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    if (nthreads > nchunks) nthreads = (int)nchunks;
    pthread_t th[64];
    int started = 0;
    for (int i = 1; i < nthreads; i++)
        if (pthread_create(&th[started], NULL, par_worker, &p) == 0) started++;
    par_worker(&p);
    for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
}

/* ----------------------------------------------------------- generators -- */

typedef struct { void* out; int64_t lo; uint64_t range; double dlo, dscale; int esz; } uni_ctx;

static void uni_int_fn(void* c_, rng_t* r, int64_t s, int64_t e) {
    uni_ctx* c = (uni_ctx*)c_;
    for (int64_t i = s; i < e; i++) {
        int64_t v = (int64_t)((uint64_t)c->lo + rng_bounded(r, c->range));
        switch (c->esz) {
            case 8: ((int64_t*)c->out)[i] = v; break;
            case 4: ((int32_t*)c->out)[i] = (int32_t)v; break;
            default: ((uint32_t*)c->out)[i] = (uint32_t)v; break;
        }
    }
}

void tdg_i64_uniform(int64_t* out, int64_t n, int64_t lo, int64_t hi,
                     uint64_t seed, uint32_t stream, int nthreads) {
    uni_ctx c = { out, lo, (uint64_t)hi - (uint64_t)lo + 1, 0, 0, 8 };
    run_chunks(uni_int_fn, &c, n, seed, stream, nthreads);
}

void tdg_i32_uniform(int32_t* out, int64_t n, int32_t lo, int32_t hi,
                     uint64_t seed, uint32_t stream, int nthreads) {
    uni_ctx c = { out, lo, (uint64_t)((int64_t)hi - (int64_t)lo + 1), 0, 0, 4 };
    run_chunks(uni_int_fn, &c, n, seed, stream, nthreads);
}

void tdg_u32_uniform(uint32_t* out, int64_t n, uint32_t lo, uint32_t hi,
                     uint64_t seed, uint32_t stream, int nthreads) {
    uni_ctx c = { out, (int64_t)lo, (uint64_t)hi - (uint64_t)lo + 1, 0, 0, 5 };
    run_chunks(uni_int_fn, &c, n, seed, stream, nthreads);
}

static void uni_f64_fn(void* c_, rng_t* r, int64_t s, int64_t e) {
    uni_ctx* c = (uni_ctx*)c_;
    double* o = (double*)c->out;
    for (int64_t i = s; i < e; i++) o[i] = c->dlo + rng_unit(r) * c->dscale;
}

void tdg_f64_uniform(double* out, int64_t n, double lo, double hi,
                     uint64_t seed, uint32_t stream, int nthreads) {
    uni_ctx c = { out, 0, 0, lo, hi - lo, 8 };
    run_chunks(uni_f64_fn, &c, n, seed, stream, nthreads);
}

typedef struct { double* out; double mu, sigma; int logn; } norm_ctx;

static void norm_fn(void* c_, rng_t* r, int64_t s, int64_t e) {
    norm_ctx* c = (norm_ctx*)c_;
    for (int64_t i = s; i < e; i++) {
        double z = c->mu + c->sigma * rng_normal(r);
        c->out[i] = c->logn ? exp(z) : z;
    }
}

void tdg_f64_normal(double* out, int64_t n, double mu, double sigma,
                    uint64_t seed, uint32_t stream, int nthreads) {
    norm_ctx c = { out, mu, sigma, 0 };
    run_chunks(norm_fn, &c, n, seed, stream, nthreads);
}

void tdg_f64_lognormal(double* out, int64_t n, double mu, double sigma,
                       uint64_t seed, uint32_t stream, int nthreads) {
    norm_ctx c = { out, mu, sigma, 1 };
    run_chunks(norm_fn, &c, n, seed, stream, nthreads);
}

/* Zipf by rejection-inversion: W. Hormann and G. Derflinger, "Rejection-
 * inversion to generate variates from monotone discrete distributions",
 * ACM TOMACS 6(3), 1996.  With h(x) = x^-s and its antiderivative
 * H(x) = (x^(1-s) - 1) / (1 - s), draw u uniformly in [H(N+1/2), H(3/2) - 1],
 * invert to x = H^-1(u), round to k, and accept k when k - x <= c or
 * u >= H(k + 1/2) - h(k).  The helpers evaluate log1p(y)/y and expm1(y)/y
 * stably near y = 0. */
static double zf_helper1(double y) {                 /* log1p(y) / y */
    if (fabs(y) > 1e-8) return log1p(y) / y;
    return 1.0 - y * (0.5 - y * (1.0 / 3.0 - 0.25 * y));
}
static double zf_helper2(double y) {                 /* expm1(y) / y */
    if (fabs(y) > 1e-8) return expm1(y) / y;
    return 1.0 + y * 0.5 * (1.0 + y * (1.0 / 3.0) * (1.0 + 0.25 * y));
}
typedef struct { double s, hx1, hn, c; int64_t n; int64_t* out; } zipf_ctx;
static double zf_h(const zipf_ctx* z, double x) { return exp(-z->s * log(x)); }
static double zf_H(const zipf_ctx* z, double x) {
    double lx = log(x);
    return zf_helper2((1.0 - z->s) * lx) * lx;
}
static double zf_Hinv(const zipf_ctx* z, double x) {
    double t = x * (1.0 - z->s);
    if (t < -1.0) t = -1.0;                          /* rounding guard */
    return exp(zf_helper1(t) * x);
}

int64_t tdg_zipf_key(int64_t rank) {
    /* Bijective 64-bit mix of the rank: heavy hitters get scattered i64
     * keys (possibly negative) instead of 1, 2, 3 ... */
    uint64_t z = (uint64_t)rank;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (int64_t)z;
}

static void zipf_fn(void* c_, rng_t* r, int64_t s, int64_t e) {
    zipf_ctx* z = (zipf_ctx*)c_;
    for (int64_t i = s; i < e; i++) {
        int64_t k;
        for (;;) {
            double u = z->hn + rng_unit(r) * (z->hx1 - z->hn);
            double x = zf_Hinv(z, u);
            k = (int64_t)(x + 0.5);
            if (k < 1) k = 1;
            else if (k > z->n) k = z->n;
            if ((double)k - x <= z->c || u >= zf_H(z, (double)k + 0.5) - zf_h(z, (double)k))
                break;
        }
        z->out[i] = tdg_zipf_key(k);
    }
}

void tdg_i64_zipf_keys(int64_t* out, int64_t n, int64_t n_values, double s,
                       uint64_t seed, uint32_t stream, int nthreads) {
    zipf_ctx z;
    z.s = s; z.n = n_values; z.out = out;
    z.hx1 = zf_H(&z, 1.5) - 1.0;
    z.hn = zf_H(&z, (double)n_values + 0.5);
    z.c = 2.0 - zf_Hinv(&z, zf_H(&z, 2.5) - zf_h(&z, 2.0));
    run_chunks(zipf_fn, &z, n, seed, stream, nthreads);
}

void tdg_i64_permutation(int64_t* out, int64_t n, int64_t base,
                         uint64_t seed, uint32_t stream) {
    for (int64_t i = 0; i < n; i++) out[i] = base + i;
    rng_t r;
    rng_seed(&r, seed, stream, 0);
    for (int64_t i = n - 1; i > 0; i--) {            /* Fisher-Yates */
        int64_t j = (int64_t)rng_bounded(&r, (uint64_t)i + 1);
        int64_t t = out[i]; out[i] = out[j]; out[j] = t;
    }
}

typedef struct { int64_t* out; const int64_t* keys; int64_t nk; } pick_ctx;

static void pick_fn(void* c_, rng_t* r, int64_t s, int64_t e) {
    pick_ctx* c = (pick_ctx*)c_;
    for (int64_t i = s; i < e; i++)
        c->out[i] = c->keys[rng_bounded(r, (uint64_t)c->nk)];
}

void tdg_i64_pick(int64_t* out, int64_t n, const int64_t* keys, int64_t n_keys,
                  uint64_t seed, uint32_t stream, int nthreads) {
    pick_ctx c = { out, keys, n_keys };
    run_chunks(pick_fn, &c, n, seed, stream, nthreads);
}

/* ------------------------------------------------------------- recipes -- */

void tdg_recipe_fg(int64_t n, uint64_t seed, int nthreads,
                   int64_t* flag, int64_t* region, double* amount) {
    tdg_i64_uniform(flag, n, 0, 1, seed, 101, nthreads);
    tdg_i64_uniform(region, n, 0, 15, seed, 102, nthreads);
    tdg_f64_uniform(amount, n, 0.0, 1000.0, seed, 103, nthreads);
}

void tdg_recipe_dim(int64_t n_dim, int64_t n_cat, uint64_t seed, uint32_t stream_base,
                    int nthreads, int64_t* pk, int32_t* cat) {
    tdg_i64_permutation(pk, n_dim, 0, seed, stream_base + 1);
    tdg_i32_uniform(cat, n_dim, 0, (int32_t)(n_cat - 1), seed, stream_base + 2, nthreads);
}

void tdg_recipe_fact(int64_t n_fact, int64_t n_dim, uint64_t seed, uint32_t stream_base,
                     int nthreads, int64_t* fk, int64_t* qty, double* price) {
    tdg_i64_uniform(fk, n_fact, 0, n_dim - 1, seed, stream_base + 3, nthreads);
    tdg_i64_uniform(qty, n_fact, 1, 100, seed, stream_base + 4, nthreads);
    tdg_f64_lognormal(price, n_fact, 3.0, 1.0, seed, stream_base + 5, nthreads);
}

void tdg_recipe_hc(int64_t n, int64_t n_values, double s, uint32_t n_sym, uint64_t seed,
                   int nthreads, int64_t* key, uint32_t* sym, int64_t* v_i64, double* v_f64) {
    tdg_i64_zipf_keys(key, n, n_values, s, seed, 301, nthreads);
    tdg_u32_uniform(sym, n, 0, n_sym - 1, seed, 302, nthreads);
    tdg_i64_uniform(v_i64, n, -1000000000LL, 1000000000LL, seed, 303, nthreads);
    tdg_f64_normal(v_f64, n, 0.0, 1.0, seed, 304, nthreads);
}
