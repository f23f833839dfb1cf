/*
 * k_pgroup.hip -- radix-partitioned direct-array group-by for group tables
 * too large for LDS.
 *
 * A direct-array group-by folds every row into its group's cells.  When the
 * table fits LDS each workgroup keeps a private copy (k_da_lds); beyond that
 * the fold goes to a global table with atomics, and global 64-bit atomics
 * run at ~26 G/s on MI355X whether or not the table stays in L2 -- for
 * 300 M rows x 7 cells that is ~80 ms.  Here the rows are instead split by
 * the high bits of their group slot, in passes of up to 6 bits, until each
 * partition's slot range has a table that fits LDS; every partition is then
 * folded in LDS and merged into the global table once per workgroup.
 *
 * A pass is the counting scatter of a radix sort: a job (a range of items
 * of one parent partition, handled by one workgroup) first counts its items
 * per sub-partition (k_pg_hist), an exclusive scan orders those counts
 * parent by parent, sub-partition by sub-partition, job by job, and the
 * scatter (k_pg_pass) writes each job's items at its offsets -- ranks within
 * a wave from ballots, running cursors in LDS, no global atomics -- so
 * records keep their row order inside each partition.  Each round of 2,048
 * items is placed in LDS in sub-partition order first and then written out
 * by consecutive threads as contiguous runs: a wave's store touches a few
 * cache lines instead of up to 64 (measured on MI355X, 300 M rows in 256
 * sub-partitions: 6.2 -> 4.75 ms per pass).  A record is the
 * group slot (32 bits) plus one 64-bit word per aggregate input (the column
 * value widened as the group-by reads it, or the row index for FIRST /
 * LAST).  The leaf fold (k_pg_agg) is the group-by's own cell fold, with
 * the inputs taken from the record words.
 */
#include <type_traits>
#include <vector>

#include "dev_common.h"
#include "group_plan.h"

#define PG_T     512                      /* threads per workgroup */
#define PG_W     (PG_T / WAVE)
#define PG_R     4                        /* items per lane per scatter round */
#define PG_MAXB  8                        /* bits per pass: up to 256 sub-partitions */
#define PG_MAXP  (1 << PG_MAXB)
#define PG_MAXV  4                        /* value words per record */
#define PG_JOB   ((int64_t)1 << 17)       /* items per job */
#define PG_U     4                        /* records per thread per leaf step */

struct PGBuf {                            /* records of one pass (SoA) */
    uint32_t* slot;                       /* direct array: the group slot */
    uint64_t* key;                        /* hash: the 64-bit group key */
    uint64_t* val[PG_MAXV];
};

struct PGJob {
    int64_t  r0, r1;                      /* item range */
    int64_t  parent;                      /* partition the items belong to */
    int32_t  js, nj;                      /* first job of the parent, jobs of the parent */
};

struct PGRows {                           /* the first pass reads the group's rows */
    GroupPlan P;
    const uint64_t* mask;                 /* selection bits, NULL: every row */
    int64_t  nrows;
    int8_t   vsrc[PG_MAXV];               /* value word v: aggregate index, -1: row index */
};

/* Item i of a job: group key (the slot, or the hash key) and value words,
 * from rows or from records. */
template <bool ROWS, bool HK, int NV, int KE = 0>
__device__ __forceinline__ bool pg_item(const PGRows& S, const PGBuf& in, int64_t i, int64_t r1,
                                        uint64_t& key, uint64_t (&v)[PG_MAXV]) {
    if (i >= r1) return false;
    if (ROWS && KE != 0) {
        /* one key of KE bytes, 8-byte value columns: plain loads, no width
         * switch, so the loads of a round are in flight together */
        if (S.mask && !((S.mask[i >> 6] >> (i & 63)) & 1)) return false;
        const tdd_col_t& kc = S.P.keys[0];
        const int64_t raw = KE == 8 ? ((const int64_t*)kc.data)[i]
                          : kc.type == TD_SYM ? (int64_t)((const uint32_t*)kc.data)[i]
                                              : (int64_t)((const int32_t*)kc.data)[i];
        key = HK && S.P.kstride[0] == 0 ? (uint64_t)raw
                                        : (uint64_t)((raw - S.P.kmin[0]) * S.P.kstride[0]);
#pragma unroll
        for (int k = 0; k < NV; k++)
            v[k] = S.vsrc[k] >= 0 ? (uint64_t)((const int64_t*)S.P.aggs[S.vsrc[k]].data)[i]
                                  : (uint64_t)i;
        return true;
    }
    if (ROWS) {
        if (S.mask && !((S.mask[i >> 6] >> (i & 63)) & 1)) return false;
        key = HK ? ht_key(S.P, i) : (uint64_t)da_slot2(S.P, i, i);
#pragma unroll
        for (int k = 0; k < NV; k++)
            v[k] = S.vsrc[k] >= 0 ? (uint64_t)ld_i64(S.P.aggs[S.vsrc[k]], i) : (uint64_t)i;
    } else {
        key = HK ? in.key[i] : (uint64_t)in.slot[i];
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = in.val[k][i];
    }
    return true;
}

/* The bits a pass partitions on: the slot itself, or the key's hash (whose
 * low bits then index the leaf's LDS hash table, independent of these). */
template <bool HK>
__device__ __forceinline__ uint64_t pg_bits(uint64_t key) {
    return HK ? mix64(key) : key;
}

/* Counts of one job's items per sub-partition, written where the exclusive
 * scan turns them into write offsets: parent block (js * PL), then
 * sub-partition s (s * nj), then job (j - js). */
template <bool ROWS, bool HK, int KE = 0>
__global__ __launch_bounds__(PG_T) void k_pg_hist(PGRows S, PGBuf in, const PGJob* jobs, int njobs,
                                                  int shift, int PL, uint32_t* Q) {
    __shared__ uint32_t h[PG_MAXP];
    for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
        const PGJob J = jobs[j];
        for (int s = threadIdx.x; s < PL; s += blockDim.x) h[s] = 0;
        __syncthreads();
        uint64_t v[PG_MAXV];
        const int lane = lane_id();
        /* every lane of a wave runs the same trip count (i advances by the
         * block), so the wave can count a run of one sub-partition at once */
        for (int64_t i0 = J.r0; i0 < J.r1; i0 += blockDim.x) {
            const int64_t i = i0 + threadIdx.x;
            uint64_t key = 0;
            const bool live = pg_item<ROWS, HK, 0, KE>(S, in, i, J.r1, key, v);
            const int s = live ? (int)((pg_bits<HK>(key) >> shift) & (PL - 1)) : -1;
            const uint64_t lm = __ballot(live);
            if (!lm) continue;
            const int first = __ffsll((unsigned long long)lm) - 1;
            const int s0 = __shfl(s, first);
            if (__ballot(live && s != s0) == 0) {
                if (lane == first) atomicAdd(&h[s0], (uint32_t)__popcll(lm));
            } else if (live) {
                atomicAdd(&h[s], 1u);
            }
        }
        __syncthreads();
        for (int s = threadIdx.x; s < PL; s += blockDim.x)
            Q[(int64_t)J.js * PL + (int64_t)s * J.nj + (j - J.js)] = h[s];
        __syncthreads();
    }
}

/* The scatter of one pass. */
template <bool ROWS, bool HK, int NV, int R, int KE = 0>
__global__ __launch_bounds__(PG_T) void k_pg_pass(PGRows S, PGBuf in, PGBuf out, const PGJob* jobs,
                                                  int njobs, int shift, int PL, const int64_t* O,
                                                  int64_t nout) {
    __shared__ int64_t cur[PG_MAXP];
    __shared__ int64_t dbase[PG_MAXP];        /* round: global position of LDS record 0 of s */
    __shared__ uint32_t tot[PG_MAXP];
    __shared__ uint32_t rbase[PG_MAXP];       /* round: first LDS record of sub-partition s */
    __shared__ uint32_t wcnt[PG_W][PG_MAXP];
    __shared__ uint32_t ntot;
    /* the round's records in sub-partition order, written out as runs */
    __shared__ typename std::conditional<HK, uint64_t, uint32_t>::type skey[(PG_T * R)];
    __shared__ uint64_t sval[NV > 0 ? NV : 1][(PG_T * R)];
    __shared__ uint8_t ssub[(PG_T * R)];
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
        const PGJob J = jobs[j];
        for (int s = threadIdx.x; s < PL; s += blockDim.x)
            cur[s] = O[(int64_t)J.js * PL + (int64_t)s * J.nj + (j - J.js)];
        for (int64_t b = J.r0; b < J.r1; b += (PG_T * R)) {
            uint64_t slot[R];
            int32_t sub[R], rk[R];
            uint64_t v[R][PG_MAXV];
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int64_t i = b + (int64_t)(wv * R + k) * 64 + lane;
                const bool live = pg_item<ROWS, HK, NV, KE>(S, in, i, J.r1, slot[k], v[k]);
                sub[k] = live ? (int32_t)((pg_bits<HK>(slot[k]) >> shift) & (PL - 1)) : -1;
            }
            /* ranks within the wave, per sub-partition: the wave's own row of
             * LDS counters (its LDS operations complete in program order) */
            for (int s2 = lane; s2 < PL; s2 += WAVE) wcnt[wv][s2] = 0;   /* PL may exceed 64 */
#pragma unroll
            for (int k = 0; k < R; k++) {
                /* a wave whose items all go to one sub-partition takes its
                 * ranks with one counter update */
                const uint64_t lm = __ballot(sub[k] >= 0);
                const int first = lm ? __ffsll((unsigned long long)lm) - 1 : 0;
                const int s0 = __shfl(sub[k], first);
                if (lm && __ballot(sub[k] >= 0 && sub[k] != s0) == 0) {
                    uint32_t base = 0;
                    if (lane == first) base = atomicAdd(&wcnt[wv][s0], (uint32_t)__popcll(lm));
                    base = __shfl(base, first);
                    rk[k] = sub[k] >= 0 ? (int32_t)(base + lanes_below(lm)) : 0;
                } else {
                    rk[k] = sub[k] >= 0 ? (int32_t)atomicAdd(&wcnt[wv][sub[k]], 1u) : 0;
                }
            }
            __syncthreads();
            /* wave offsets within the round, then the round's totals */
            if (threadIdx.x < PL) {
                const int s = threadIdx.x;
                uint32_t run = 0;
#pragma unroll
                for (int w = 0; w < PG_W; w++) {
                    const uint32_t c = wcnt[w][s];
                    wcnt[w][s] = run;
                    run += c;
                }
                tot[s] = run;
            }
            __syncthreads();
            /* wave 0: sub-partition bases within the round (exclusive scan of
             * the totals, 4 per lane), their global positions, next cursors */
            if (wv == 0) {
                uint32_t t4[PG_MAXP / WAVE], sum = 0;
#pragma unroll
                for (int q = 0; q < PG_MAXP / WAVE; q++) {
                    const int s = lane * (PG_MAXP / WAVE) + q;
                    t4[q] = s < PL ? tot[s] : 0u;
                    sum += t4[q];
                }
                uint32_t inc = sum;
#pragma unroll
                for (int d = 1; d < WAVE; d <<= 1) {
                    const uint32_t y = __shfl_up(inc, d);
                    if (lane >= d) inc += y;
                }
                uint32_t ex = inc - sum;
#pragma unroll
                for (int q = 0; q < PG_MAXP / WAVE; q++) {
                    const int s = lane * (PG_MAXP / WAVE) + q;
                    if (s < PL) {
                        rbase[s] = ex;
                        dbase[s] = cur[s] - (int64_t)ex;
                        cur[s] += t4[q];
                    }
                    ex += t4[q];
                }
                if (lane == WAVE - 1) ntot = inc;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < R; k++) {
                if (sub[k] < 0) continue;
                const uint32_t lp = rbase[sub[k]] + wcnt[wv][sub[k]] + (uint32_t)rk[k];
                skey[lp] = (typename std::conditional<HK, uint64_t, uint32_t>::type)slot[k];
#pragma unroll
                for (int q = 0; q < NV; q++) sval[q][lp] = v[k][q];
                ssub[lp] = (uint8_t)sub[k];
            }
            __syncthreads();
            /* consecutive threads write consecutive records of a run; the
             * next round's first barrier orders these reads before the
             * staging is rewritten */
            const uint32_t n_r = ntot;
            for (uint32_t t = threadIdx.x; t < n_r; t += PG_T) {
                const int64_t at = dbase[ssub[t]] + (int64_t)t;
                if (at < 0 || at >= nout) continue;          /* never taken when counts agree */
                if (HK) out.key[at] = skey[t];
                else out.slot[at] = (uint32_t)skey[t];
#pragma unroll
                for (int q = 0; q < NV; q++) out.val[q][at] = sval[q][t];
            }
        }
        __syncthreads();                      /* cur of the next job */
    }
}

/* Fold one record (value words v) into cells c, LDS or global. */
template <int NV, bool SKIP>
__device__ __forceinline__ void pg_fold(uint64_t* c, const GroupPlan& P, const int8_t* cw,
                                        const uint8_t* vf, const uint64_t (&v)[PG_MAXV]) {
    atomicAdd((unsigned long long*)&c[0], 1ull);
    for (int k = 1; k < P.ncells; k++) {
        uint64_t x;
        switch (cw[k]) {                          /* compile-time register per word */
        case 0: x = v[0]; break;
        case 1: x = v[NV > 1 ? 1 : 0]; break;
        case 2: x = v[NV > 2 ? 2 : 0]; break;
        default: x = v[NV > 3 ? 3 : 0]; break;
        }
        const bool f = vf[cw[k]];
        const double d = f ? __longlong_as_double((long long)x) : (double)(int64_t)x;
        switch (P.kind[k]) {
        case C_ADD_F64: atomicAdd((double*)&c[k], d); break;
        case C_ADD_I64: atomicAdd((unsigned long long*)&c[k], (unsigned long long)x); break;
        case C_MIN_I64: case C_MIN_ROW: min_cell<SKIP>(&c[k], (int64_t)x); break;
        case C_MAX_I64: case C_MAX_ROW: max_cell<SKIP>(&c[k], (int64_t)x); break;
        case C_MIN_F64: min_cell<SKIP>(&c[k], f64_ord(d)); break;
        case C_MAX_F64: max_cell<SKIP>(&c[k], f64_ord(d)); break;
        default: atomicAdd((double*)&c[k], d * d); break;
        }
    }
}

/* Leaf fold: job j covers records of one leaf partition (slot range
 * [parent << lbits, +2^lbits)), folded into an LDS table, then merged into
 * the global direct-array table. */
/* Cell k of U records (bits of `live`) at LDS table rows loc[u], the input
 * from record word Q: one dispatch on the cell's kind for all U records. */
template <int U, int Q>
__device__ __forceinline__ void pg_cell_rows(uint64_t* lt, int nc, int k, int kind, bool f,
                                             const uint32_t (&loc)[U], uint32_t live,
                                             const uint64_t (&v)[U][PG_MAXV]) {
#define PG_ROWS(BODY)                                                      \
    _Pragma("unroll") for (int u = 0; u < U; u++) {                        \
        if (!((live >> u) & 1)) continue;                                  \
        uint64_t* c = &lt[(int64_t)loc[u] * nc + k];                       \
        const uint64_t x = v[u][Q];                                        \
        const double d = f ? __longlong_as_double((long long)x) : (double)(int64_t)x; \
        (void)d;                                                           \
        BODY;                                                              \
    }
    switch (kind) {
    case C_ADD_F64: PG_ROWS(atomicAdd((double*)c, d)) break;
    case C_ADD_I64: PG_ROWS(atomicAdd((unsigned long long*)c, (unsigned long long)x)) break;
    case C_MIN_I64: case C_MIN_ROW: PG_ROWS(cell_min(c, (int64_t)x)) break;
    case C_MAX_I64: case C_MAX_ROW: PG_ROWS(cell_max(c, (int64_t)x)) break;
    case C_MIN_F64: PG_ROWS(cell_min(c, f64_ord(d))) break;
    case C_MAX_F64: PG_ROWS(cell_max(c, f64_ord(d))) break;
    default: PG_ROWS(atomicAdd((double*)c, d * d)) break;
    }
#undef PG_ROWS
}

/* U records into the LDS table: COUNT, then cell by cell. */
template <int U, int NV>
__device__ __forceinline__ void pg_fold_rows(uint64_t* lt, const GroupPlan& P, const int8_t* cw,
                                             const uint8_t* vf, const uint32_t (&loc)[U],
                                             uint32_t live, const uint64_t (&v)[U][PG_MAXV]) {
    const int nc = P.ncells;
#pragma unroll
    for (int u = 0; u < U; u++)
        if ((live >> u) & 1) atomicAdd((unsigned long long*)&lt[(int64_t)loc[u] * nc], 1ull);
    for (int k = 1; k < nc; k++) {
        const int q = cw[k];
        const bool f = vf[q];
        switch (q) {                              /* compile-time record word */
        case 0: pg_cell_rows<U, 0>(lt, nc, k, P.kind[k], f, loc, live, v); break;
        case 1: pg_cell_rows<U, (NV > 1 ? 1 : 0)>(lt, nc, k, P.kind[k], f, loc, live, v); break;
        case 2: pg_cell_rows<U, (NV > 2 ? 2 : 0)>(lt, nc, k, P.kind[k], f, loc, live, v); break;
        default: pg_cell_rows<U, (NV > 3 ? 3 : 0)>(lt, nc, k, P.kind[k], f, loc, live, v); break;
        }
    }
}

template <int NV>
__global__ __launch_bounds__(PG_T) void k_pg_agg(GroupPlan P, PGBuf in, const PGJob* jobs,
                                                 int njobs, int lbits, const int8_t* cw_dev,
                                                 const uint8_t* vf64_dev, uint64_t* tab,
                                                 int64_t n_slots) {
    extern __shared__ uint64_t lt[];
    __shared__ int8_t cw[GP_MAX_CELLS];
    __shared__ uint8_t vf[PG_MAXV];
    if (threadIdx.x < GP_MAX_CELLS) cw[threadIdx.x] = cw_dev[threadIdx.x];
    if (threadIdx.x < PG_MAXV) vf[threadIdx.x] = vf64_dev[threadIdx.x];
    const int nc = P.ncells;
    const int64_t nl = (int64_t)1 << lbits;
    for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
        const PGJob J = jobs[j];
        for (int64_t i = threadIdx.x; i < nl * nc; i += blockDim.x)
            lt[i] = cell_identity(P.kind[i % nc]);
        __syncthreads();
        const uint32_t base = (uint32_t)(J.parent << lbits);
        for (int64_t i0 = J.r0 + threadIdx.x; i0 < J.r1; i0 += (int64_t)blockDim.x * PG_U) {
            uint32_t sl[PG_U];
            uint64_t v[PG_U][PG_MAXV];
#pragma unroll
            for (int u = 0; u < PG_U; u++) {           /* every load of the step first */
                const int64_t i = i0 + (int64_t)u * blockDim.x;
                sl[u] = i < J.r1 ? in.slot[i] : 0u;
#pragma unroll
                for (int q = 0; q < NV; q++) v[u][q] = i < J.r1 ? in.val[q][i] : 0ull;
            }
            uint32_t loc[PG_U];
            uint32_t live = 0;
#pragma unroll
            for (int u = 0; u < PG_U; u++) {
                loc[u] = sl[u] - base;
                if (i0 + (int64_t)u * blockDim.x < J.r1 && loc[u] < (uint32_t)nl)  /* always, */
                    live |= 1u << u;                                  /* for a correct split */
            }
            pg_fold_rows<PG_U, NV>(lt, P, cw, vf, loc, live, v);
        }
        __syncthreads();
        for (int64_t s = threadIdx.x; s < nl; s += blockDim.x) {
            const uint64_t* c = &lt[s * nc];
            if (c[0] == 0 || (int64_t)base + s >= n_slots) continue;
            cells_merge(&tab[((int64_t)base + s) * nc], c, P);
        }
        __syncthreads();
    }
}

/* Hash leaf: job j covers records of one hash partition, folded into an
 * LDS open-addressing table (the low bits of the key's hash pick the
 * entry), then each entry merged into the global hash table.  A key the LDS
 * table cannot place within PH_PROBES probes (and the all-ones key) folds
 * straight into the global table. */
#define PH_PROBES 32

struct PHList {                           /* groups of parts folded by a single job */
    unsigned long long* keys;
    uint64_t*           cells;            /* [cap][ncells] */
    unsigned long long* count;
    int64_t             cap;
};

template <int NV>
__global__ __launch_bounds__(PG_T) void k_pg_hagg(GroupPlan P, PGBuf in, const PGJob* jobs,
                                                  int njobs, int lcap, const int8_t* cw_dev,
                                                  const uint8_t* vf64_dev,
                                                  unsigned long long* gkeys, uint64_t gmask,
                                                  uint64_t* gcells, PHList L, int* overflow) {
    extern __shared__ uint64_t sm[];
    unsigned long long* lk = (unsigned long long*)sm;
    uint64_t* lc = sm + lcap;
    uint32_t* lidx = (uint32_t*)(lc + (int64_t)lcap * P.ncells);
    __shared__ uint32_t nloc;
    __shared__ unsigned long long lbase;
    __shared__ int8_t cw[GP_MAX_CELLS];
    __shared__ uint8_t vf[PG_MAXV];
    if (threadIdx.x < GP_MAX_CELLS) cw[threadIdx.x] = cw_dev[threadIdx.x];
    if (threadIdx.x < PG_MAXV) vf[threadIdx.x] = vf64_dev[threadIdx.x];
    const int nc = P.ncells;
    const uint32_t lmask = (uint32_t)lcap - 1;
    for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
        const PGJob J = jobs[j];
        for (int i = threadIdx.x; i < lcap; i += blockDim.x) lk[i] = EMPTY_KEY;
        for (int64_t i = threadIdx.x; i < (int64_t)lcap * nc; i += blockDim.x)
            lc[i] = cell_identity(P.kind[i % nc]);
        __syncthreads();
        for (int64_t i0 = J.r0 + threadIdx.x; i0 < J.r1; i0 += (int64_t)blockDim.x * PG_U) {
          uint64_t keys[PG_U];
          uint64_t vs[PG_U][PG_MAXV];
#pragma unroll
          for (int u = 0; u < PG_U; u++) {             /* every load of the step first */
              const int64_t i = i0 + (int64_t)u * blockDim.x;
              keys[u] = i < J.r1 ? in.key[i] : EMPTY_KEY;
#pragma unroll
              for (int q = 0; q < NV; q++) vs[u][q] = i < J.r1 ? in.val[q][i] : 0ull;
          }
#pragma unroll
          for (int u = 0; u < PG_U; u++) {
            if (i0 + (int64_t)u * blockDim.x >= J.r1) break;
            const uint64_t key = keys[u];
            const uint64_t (&v)[PG_MAXV] = vs[u];
            int e = -1;
            if (key != EMPTY_KEY) {
                uint32_t x = (uint32_t)mix64(key) & lmask;
                for (int p = 0; p < PH_PROBES; p++) {
                    const unsigned long long k = lk[x];
                    if (k == key) { e = (int)x; break; }
                    if (k == EMPTY_KEY) {
                        const unsigned long long prev = atomicCAS(&lk[x], EMPTY_KEY, key);
                        if (prev == EMPTY_KEY || prev == key) { e = (int)x; break; }
                    }
                    x = (x + 1) & lmask;
                }
            }
            if (e >= 0) {
                pg_fold<NV, true>(&lc[(int64_t)e * nc], P, cw, vf, v);
            } else {
                const int64_t g = ht_find(gkeys, gmask, key);
                if (g < 0) { *overflow = 1; continue; }
                pg_fold<NV, false>(&gcells[g * nc], P, cw, vf, v);
            }
          }
        }
        __syncthreads();
        if (J.nj == 1) {
            /* the part's only job holds every group of its keys: append
             * them to the list, one reservation per job */
            if (threadIdx.x == 0) nloc = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < lcap; i += blockDim.x)
                if (lk[i] != EMPTY_KEY && lc[(int64_t)i * nc] != 0) lidx[i] = atomicAdd(&nloc, 1u);
            __syncthreads();
            if (threadIdx.x == 0) lbase = atomicAdd(L.count, (unsigned long long)nloc);
            __syncthreads();
            for (int i = threadIdx.x; i < lcap; i += blockDim.x) {
                if (lk[i] == EMPTY_KEY || lc[(int64_t)i * nc] == 0) continue;
                const int64_t at = (int64_t)(lbase + lidx[i]);
                if (at >= L.cap) { overflow[1] = 1; continue; }   /* the list */
                L.keys[at] = lk[i];
                for (int c = 0; c < nc; c++) L.cells[at * nc + c] = lc[(int64_t)i * nc + c];
            }
        } else {
            for (int i = threadIdx.x; i < lcap; i += blockDim.x) {
                if (lk[i] == EMPTY_KEY || lc[(int64_t)i * nc] == 0) continue;
                const int64_t g = ht_find(gkeys, gmask, lk[i]);
                if (g < 0) { *overflow = 1; continue; }
                cells_merge(&gcells[g * nc], &lc[(int64_t)i * nc], P);
            }
        }
        __syncthreads();
    }
}

/* Hot-key fold (the pre-aggregation step of the hash group-by): workgroup
 * b takes the 64-row words [b * wpb, (b + 1) * wpb) and reads their rows
 * the way the first scatter pass does (plain loads when the key and values
 * allow), PG_U words per wave per step; rows of hot keys fold into this
 * workgroup's LDS accumulator block of their key, every other selected
 * row is written as a record (key + value words) to the workgroup's region
 * [b * wpb * 64, ...) of `out`, in row order; counts[b] = its records.
 * The blocks merge into the global hot cells once per workgroup. */
#define PH_SET 2048                       /* LDS hash set of hot keys (load <= 0.25) */
#define PH_STEP (PG_W * PG_U)             /* words per workgroup step */

template <int NV, int KE>
__global__ __launch_bounds__(PG_T) void k_pg_hotfold(PGRows S, const unsigned long long* hot,
                                                     int nhot, const int8_t* cw_dev,
                                                     const uint8_t* vf_dev, int64_t wpb,
                                                     PGBuf out, int64_t* counts,
                                                     uint64_t* gcells) {
    extern __shared__ uint64_t sm[];
    unsigned long long* skey = (unsigned long long*)sm;      /* PH_SET */
    uint16_t* sidx = (uint16_t*)(sm + PH_SET);               /* PH_SET */
    uint64_t* cells = sm + PH_SET + PH_SET / 4;              /* nhot * ncells */
    __shared__ int8_t cw[GP_MAX_CELLS];
    __shared__ uint8_t vf[PG_MAXV];
    __shared__ uint32_t wcnt[PH_STEP];
    __shared__ uint32_t cursor;
    const GroupPlan& P = S.P;
    const int nc = P.ncells;
    if (threadIdx.x < GP_MAX_CELLS) cw[threadIdx.x] = cw_dev[threadIdx.x];
    if (threadIdx.x < PG_MAXV) vf[threadIdx.x] = vf_dev[threadIdx.x];
    if (threadIdx.x == 0) cursor = 0;
    for (int i = threadIdx.x; i < PH_SET; i += PG_T) skey[i] = EMPTY_KEY;
    for (int i = threadIdx.x; i < nhot * nc; i += PG_T) cells[i] = cell_identity(P.kind[i % nc]);
    __syncthreads();
    for (int h = threadIdx.x; h < nhot; h += PG_T) {        /* distinct keys */
        uint32_t x = (uint32_t)(mix64(hot[h]) >> 32) & (PH_SET - 1);
        while (atomicCAS(&skey[x], EMPTY_KEY, hot[h]) != EMPTY_KEY) x = (x + 1) & (PH_SET - 1);
        sidx[x] = (uint16_t)h;
    }
    __syncthreads();
    PGBuf none;
    memset(&none, 0, sizeof(none));
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int64_t nw = (S.nrows + 63) >> 6;
    const int64_t w_begin = (int64_t)blockIdx.x * wpb;
    const int64_t w_end = w_begin + wpb < nw ? w_begin + wpb : nw;
    const int64_t region = w_begin * 64;
    for (int64_t ws = w_begin; ws < w_end; ws += PH_STEP) {   /* uniform across the block */
        uint64_t key[PG_U];
        uint64_t v[PG_U][PG_MAXV];
        bool live[PG_U];
        uint64_t m[PG_U];
#pragma unroll
        for (int u = 0; u < PG_U; u++) {           /* every load of the step first */
            const int64_t w = ws + wv * PG_U + u;
            live[u] = w < w_end && pg_item<true, true, NV, KE>(S, none, w * 64 + lane, S.nrows,
                                                                key[u], v[u]);
        }
#pragma unroll
        for (int u = 0; u < PG_U; u++) {
            int h = -1;
            if (live[u] && key[u] != EMPTY_KEY) {
                uint32_t x = (uint32_t)(mix64(key[u]) >> 32) & (PH_SET - 1);
                for (;;) {
                    const unsigned long long k = skey[x];
                    if (k == key[u]) { h = sidx[x]; break; }
                    if (k == EMPTY_KEY) break;
                    x = (x + 1) & (PH_SET - 1);
                }
            }
            if (h >= 0) pg_fold<NV, true>(&cells[(int64_t)h * nc], P, cw, vf, v[u]);
            m[u] = __ballot(live[u] && h < 0);
            if (lane == 0) wcnt[wv * PG_U + u] = (uint32_t)__popcll(m[u]);
        }
        __syncthreads();
        uint32_t pre = cursor;                    /* records of the step's earlier words */
        for (int i = 0; i < wv * PG_U; i++) pre += wcnt[i];
#pragma unroll
        for (int u = 0; u < PG_U; u++) {
            if ((m[u] >> lane) & 1) {
                const int64_t at = region + pre + lanes_below(m[u]);
                out.key[at] = key[u];
#pragma unroll
                for (int q = 0; q < NV; q++) out.val[q][at] = v[u][q];
            }
            pre += (uint32_t)__popcll(m[u]);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int i = 0; i < PH_STEP; i++) t += wcnt[i];
            cursor += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = cursor;
    for (int h = threadIdx.x; h < nhot; h += PG_T)
        if (cells[(int64_t)h * nc] != 0)
            cells_merge(&gcells[(int64_t)h * nc], &cells[(int64_t)h * nc], P);
}

/* ------------------------------------------------------------------ host -- */

/* Host: the element bytes ld_i64 reads a column with when that is a plain
 * signed 8- or 4-byte load (4-byte SYM: unsigned), else 0. */
static int pg_plain(const tdd_col_t& c) {
    switch (c.type) {
    case TD_I64: case TD_TIMESTAMP: case TD_F64: return 8;
    case TD_I32: case TD_DATE: case TD_TIME: return 4;
    case TD_SYM:
        switch (c.attrs & TD_SYM_W_MASK) {
        case TD_SYM_W32: return 4;
        case TD_SYM_W8: case TD_SYM_W16: return 0;
        default: return 8;
        }
    default: return 0;
    }
}

template <bool HK>
static int pg_hist_and_offsets(bool rows, const PGRows& S, const PGBuf& in, const PGJob* d_jobs,
                               int njobs, int shift, int PL, uint32_t* Q, int64_t* O) {
    const int grid = stream_grid(njobs, 4);
    const int ke = rows && S.P.n_keys == 1 && S.P.keys[0].data ? pg_plain(S.P.keys[0]) : 0;
    if (rows && ke == 8)
        LAUNCH("group_pg_hist", (k_pg_hist<true, HK, 8>), grid, PG_T, 0, S, in, d_jobs, njobs, shift, PL, Q);
    else if (rows && ke == 4)
        LAUNCH("group_pg_hist", (k_pg_hist<true, HK, 4>), grid, PG_T, 0, S, in, d_jobs, njobs, shift, PL, Q);
    else if (rows)
        LAUNCH("group_pg_hist", (k_pg_hist<true, HK>), grid, PG_T, 0, S, in, d_jobs, njobs, shift, PL, Q);
    else
        LAUNCH("group_pg_hist", (k_pg_hist<false, HK>), grid, PG_T, 0, S, in, d_jobs, njobs, shift, PL, Q);
    return tdd::scan_u32_i64(Q, (int64_t)njobs * PL, O);
}


template <bool ROWS, bool HK>
static int pg_pass_launch(int nv, const PGRows& S, const PGBuf& in, const PGBuf& out,
                          const PGJob* d_jobs, int njobs, int shift, int PL, const int64_t* O,
                          int64_t nout) {
    const int grid = stream_grid(njobs, 2);
    /* rows of one 4- or 8-byte key and 8-byte value columns: the plain-load form */
    int ke = 0;
    if (ROWS && S.P.n_keys == 1 && S.P.keys[0].data) {
        ke = pg_plain(S.P.keys[0]);
        for (int k = 0; k < nv; k++)
            if (S.vsrc[k] >= 0 && pg_plain(S.P.aggs[S.vsrc[k]]) != 8) ke = 0;
    }
#define PGP1(NV_, KE_) LAUNCH("group_pg_pass", (k_pg_pass<ROWS, HK, NV_, PG_R, KE_>), grid, PG_T, 0, \
                              S, in, out, d_jobs, njobs, shift, PL, O, nout)
#define PGP(NV_)                                      \
    do {                                              \
        if (ROWS && ke == 8) PGP1(NV_, (ROWS ? 8 : 0)); \
        else if (ROWS && ke == 4) PGP1(NV_, (ROWS ? 4 : 0)); \
        else PGP1(NV_, 0);                            \
    } while (0)
    switch (nv) {
    case 0: PGP(0); break;
    case 1: PGP(1); break;
    case 2: PGP(2); break;
    case 3: PGP(3); break;
    default: PGP(4); break;
    }
#undef PGP
#undef PGP1
    return TD_OK;
}

/* Jobs over the partitions of a finished pass: partition q = parent * PL + s
 * holds items [start(q), start(q) + count(q)). */
static void pg_jobs_from(const std::vector<int64_t>& start, const std::vector<int64_t>& count,
                         std::vector<PGJob>& jobs) {
    jobs.clear();
    for (size_t q = 0; q < start.size(); q++) {
        if (count[q] == 0) continue;
        const int js = (int)jobs.size();
        const int nj = (int)ceil_div(count[q], PG_JOB);
        for (int k = 0; k < nj; k++) {
            PGJob J;
            J.r0 = start[q] + (int64_t)k * PG_JOB;
            J.r1 = J.r0 + PG_JOB < start[q] + count[q] ? J.r0 + PG_JOB : start[q] + count[q];
            J.parent = (int64_t)q;
            J.js = js;
            J.nj = nj;
            jobs.push_back(J);
        }
    }
}

/* Value words of a plan: each aggregate input column once, and the row
 * index for FIRST / LAST; cw[i] = word of cell i, vf[w] = word w is f64. */
static int pg_words(const GroupPlan& P, PGRows& S, int8_t* cw, uint8_t* vf) {
    int nv = 0;
    memset(cw, 0, GP_MAX_CELLS);
    memset(vf, 0, PG_MAXV);
    for (int i = 1; i < P.ncells; i++) {
        const bool row = P.kind[i] == C_MIN_ROW || P.kind[i] == C_MAX_ROW;
        const int src = row ? -1 : P.col[i];
        int w = -1;
        for (int k = 0; k < nv; k++) {
            const bool krow = S.vsrc[k] < 0;
            if (krow ? row : (!row && P.aggs[S.vsrc[k]].data == P.aggs[src].data &&
                              P.aggs[S.vsrc[k]].type == P.aggs[src].type))
                w = k;
        }
        if (w < 0) {
            if (nv >= PG_MAXV) return -1;
            w = nv++;
            S.vsrc[w] = (int8_t)src;
            vf[w] = !row && P.aggs[src].type == TD_F64;
        }
        cw[i] = (int8_t)w;
    }
    return nv;
}

/* Record buffers for n items: key (4-byte slot or 8-byte hash key) + words. */
struct PGBufs {
    void* a = nullptr;
    void* b = nullptr;
    ~PGBufs() { if (a) td__arena_free(a); if (b) td__arena_free(b); }
};

static bool pg_alloc(PGBufs& m, int64_t n, int nv, bool hk, PGBuf& A, PGBuf& B) {
    const size_t kb = hk ? 8 : 4;
    const size_t bytes = (size_t)(n > 0 ? n : 1) * (kb + (size_t)nv * 8) + 256;  /* + alignment */
    m.a = td__arena_alloc(bytes);
    m.b = td__arena_alloc(bytes);
    if (!m.a || !m.b) return false;
    memset(&A, 0, sizeof(A));
    memset(&B, 0, sizeof(B));
    PGBuf* bs[2] = {&A, &B};
    void* ms[2] = {m.a, m.b};
    for (int t = 0; t < 2; t++) {
        uint8_t* base = (uint8_t*)ms[t];
        if (hk) bs[t]->key = (uint64_t*)base;
        else bs[t]->slot = (uint32_t*)base;
        uint8_t* v = base + (size_t)(n > 0 ? n : 1) * kb;
        v = (uint8_t*)(((uintptr_t)v + 255) & ~(uintptr_t)255);   /* n * 4 may leave 8-byte misalignment */
        for (int k = 0; k < nv; k++) bs[t]->val[k] = (uint64_t*)v + (size_t)k * (n > 0 ? n : 1);
    }
    return true;
}

/* The scatter passes: pbits[] bits each, from bit `top` down.  On return
 * `jobs` covers the leaf partitions and `res` names the buffer holding
 * their records. */
template <bool HK>
static int pg_scatter(const PGRows& S, int nv, int64_t nrows, const int* pbits, int npass, int top,
                      const PGBuf& A, const PGBuf& B, std::vector<PGJob>& jobs, PGBuf& res,
                      const std::vector<PGJob>* first = nullptr) {
    jobs.clear();
    if (first) {
        jobs = *first;                   /* the first pass reads records of A */
    } else {
        const int nj0 = (int)ceil_div(nrows, PG_JOB);
        for (int k = 0; k < nj0; k++) {
            PGJob J;
            J.r0 = (int64_t)k * PG_JOB;
            J.r1 = J.r0 + PG_JOB < nrows ? J.r0 + PG_JOB : nrows;
            J.parent = 0;
            J.js = 0;
            J.nj = nj0;
            jobs.push_back(J);
        }
    }
    PGBuf in = A, outb = B;          /* the first pass reads rows, or A's records */
    int shift = top;
    int64_t parents = 1;
    for (int pass = 0; pass < npass; pass++) {
        POLL_CANCEL();
        const int PL = 1 << pbits[pass];
        shift -= pbits[pass];
        const int njobs = (int)jobs.size();
        SCRATCH(djobs, sizeof(PGJob) * (size_t)(njobs > 0 ? njobs : 1));
        SCRATCH(Q, (size_t)(njobs > 0 ? njobs : 1) * PL * 4);
        SCRATCH(O, ((size_t)njobs * PL + 1) * 8);
        if (njobs > 0)
            HC(hipMemcpyAsync(djobs.p, jobs.data(), sizeof(PGJob) * (size_t)njobs,
                              hipMemcpyHostToDevice, tdd::g_stream));
        const bool rows = pass == 0 && !first;
        int rc = pg_hist_and_offsets<HK>(rows, S, in, djobs.as<PGJob>(), njobs, shift, PL,
                                         Q.as<uint32_t>(), O.as<int64_t>());
        if (rc) return rc;
        if (njobs > 0) {
            rc = rows ? pg_pass_launch<true, HK>(nv, S, in, outb, djobs.as<PGJob>(), njobs, shift,
                                                 PL, O.as<int64_t>(), nrows)
                      : pg_pass_launch<false, HK>(nv, S, in, outb, djobs.as<PGJob>(), njobs, shift,
                                                  PL, O.as<int64_t>(), nrows);
            if (rc) return rc;
        }
        /* partition bounds: (parent, s) starts at the offset of the parent's
         * first job in sub-partition s */
        std::vector<int64_t> off((size_t)njobs * PL + 1);
        HC(hipMemcpyAsync(off.data(), O.p, off.size() * 8, hipMemcpyDeviceToHost, tdd::g_stream));
        HC(hipStreamSynchronize(tdd::g_stream));
        const int64_t nparts = parents * PL;
        std::vector<int64_t> start((size_t)nparts, 0), count((size_t)nparts, 0);
        std::vector<int64_t> pstart((size_t)parents, -1);
        for (const PGJob& J : jobs)
            if (pstart[(size_t)J.parent] < 0) pstart[(size_t)J.parent] = J.js;
        for (int64_t p = 0; p < parents; p++) {
            if (pstart[(size_t)p] < 0) continue;               /* empty parent */
            const int js = (int)pstart[(size_t)p];
            const int nj = jobs[(size_t)js].nj;
            for (int s2 = 0; s2 < PL; s2++) {
                const int64_t q0 = (int64_t)js * PL + (int64_t)s2 * nj;
                start[(size_t)(p * PL + s2)] = off[(size_t)q0];
                count[(size_t)(p * PL + s2)] = off[(size_t)(q0 + nj)] - off[(size_t)q0];
            }
        }
        pg_jobs_from(start, count, jobs);
        parents = nparts;
        std::swap(in, outb);
    }
    res = in;
    return TD_OK;
}

static void pg_split_bits(int upper, int* pbits, int* npass) {
    int mb = TUNE_INT("TEIDE_PG_BITS", 8);           /* bits per pass */
    if (mb < 1) mb = 1;
    if (mb > PG_MAXB) mb = PG_MAXB;
    while ((upper + mb - 1) / mb > 8 && mb < PG_MAXB) mb++;     /* at most 8 passes */
    *npass = (upper + mb - 1) / mb;
    for (int k = 0; k < *npass; k++) pbits[k] = upper / *npass + (k < upper % *npass ? 1 : 0);
}

/* Partitioned direct-array group-by of the rows selected by (mask, nrows)
 * into n_slots slots of plan P, with leaf tables of at most leaf_bytes.
 * TD_ERR_NYI when the shape is not taken. */
int run_pda(const GroupPlan& P, const uint64_t* mask, int64_t nrows, int64_t n_slots,
            int64_t leaf_bytes, tdd_group_result_t* out) {
    if (n_slots <= 1 || n_slots > ((int64_t)1 << 31) || nrows <= 0) return TD_ERR_NYI;
    PGRows S;
    memset(&S, 0, sizeof(S));
    S.P = P;
    S.mask = mask;
    S.nrows = nrows;
    int8_t cw[GP_MAX_CELLS];
    uint8_t vf[PG_MAXV];
    const int nv = pg_words(P, S, cw, vf);
    if (nv < 0) return TD_ERR_NYI;
    /* leaf: a slot range whose table fits the LDS budget; passes above it */
    int total_bits = 0;
    while (total_bits < 40 && (((int64_t)1 << total_bits) < n_slots)) total_bits++;
    int lbits = 0;
    while (lbits < total_bits && (((int64_t)2 << lbits) * P.ncells * 8) <= leaf_bytes) lbits++;
    const int upper = total_bits - lbits;
    if (upper <= 0) return TD_ERR_NYI;                 /* fits LDS: the plain LDS path */
    int pbits[8], npass;
    pg_split_bits(upper, pbits, &npass);

    PGBufs mem;
    PGBuf A, B, in;
    if (!pg_alloc(mem, nrows, nv, false, A, B)) return TD_ERR_OOM;
    std::vector<PGJob> jobs;
    int rc = pg_scatter<false>(S, nv, nrows, pbits, npass, total_bits, A, B, jobs, in);
    if (rc) return rc;
    /* leaf fold into the global table */
    SCRATCH(tab, (size_t)n_slots * P.ncells * 8);
    rc = da_table_init(P, tab.as<uint64_t>(), n_slots);
    if (rc) return rc;
    const int njobs = (int)jobs.size();
    if (njobs > 0) {
        SCRATCH(djobs, sizeof(PGJob) * (size_t)njobs);
        SCRATCH(meta, GP_MAX_CELLS + PG_MAXV);
        uint8_t hm[GP_MAX_CELLS + PG_MAXV];
        memcpy(hm, cw, GP_MAX_CELLS);
        memcpy(hm + GP_MAX_CELLS, vf, PG_MAXV);
        HC(hipMemcpyAsync(djobs.p, jobs.data(), sizeof(PGJob) * (size_t)njobs,
                          hipMemcpyHostToDevice, tdd::g_stream));
        HC(hipMemcpyAsync(meta.p, hm, sizeof(hm), hipMemcpyHostToDevice, tdd::g_stream));
        const size_t lds = ((size_t)1 << lbits) * P.ncells * 8;
        const int grid = resident_grid(k_pg_agg<4>, PG_T, lds, njobs);
        const int8_t* dcw = (const int8_t*)meta.p;
        const uint8_t* dvf = meta.as<uint8_t>() + GP_MAX_CELLS;
#define PGA(NV_) LAUNCH("group_pg_agg", k_pg_agg<NV_>, grid, PG_T, lds, P, in, djobs.as<PGJob>(), \
                        njobs, lbits, dcw, dvf, tab.as<uint64_t>(), n_slots)
        switch (nv) {
        case 0: PGA(0); break;
        case 1: PGA(1); break;
        case 2: PGA(2); break;
        case 3: PGA(3); break;
        default: PGA(4); break;
        }
#undef PGA
    }
    rc = da_emit(P, tab.as<uint64_t>(), n_slots, out);
    if (rc) return rc;
    out->path = 4;
    return TD_OK;
}

/* ------------------------------------------------- partitioned hash group -- */

struct tdd_pht {
    PGBufs            mem;
    PGBuf             res;
    std::vector<PGJob> jobs;
    GroupPlan         P;
    int               nv, lcap;
    int64_t           rows;            /* rows the partition passes carry */
    int8_t            cw[GP_MAX_CELLS];
    uint8_t           vf[PG_MAXV];
};

/* Scatter the rows of a hash group-by into partitions of the key's hash
 * whose expected distinct keys (from the estimate) fill half of an LDS
 * table; TD_ERR_NYI when one LDS table already holds them all. */
/* The hot-key fold of the rows into A (see k_pg_hotfold); *first receives
 * the first pass's jobs over the workgroups' record regions.  TD_ERR_NYI
 * when the LDS blocks do not fit. */
static int pht_hot_fold(const PGRows& S, int nv, const int8_t* cw, const uint8_t* vf,
                        const unsigned long long* hot, int nhot, uint64_t* hot_cells,
                        const PGBuf& A, std::vector<PGJob>* first, int64_t* cold) {
    const GroupPlan& P = S.P;
    const size_t lds = (size_t)PH_SET * 8 + (size_t)PH_SET * 2 + (size_t)nhot * P.ncells * 8;
    if (lds > 64 * 1024) return TD_ERR_NYI;
    int ke = P.n_keys == 1 && P.keys[0].data ? pg_plain(P.keys[0]) : 0;
    for (int k = 0; k < nv; k++)
        if (S.vsrc[k] >= 0 && pg_plain(P.aggs[S.vsrc[k]]) != 8) ke = 0;
    SCRATCH(meta, GP_MAX_CELLS + PG_MAXV);
    uint8_t hm[GP_MAX_CELLS + PG_MAXV];
    memcpy(hm, cw, GP_MAX_CELLS);
    memcpy(hm + GP_MAX_CELLS, vf, PG_MAXV);
    HC(hipMemcpyAsync(meta.p, hm, sizeof(hm), hipMemcpyHostToDevice, tdd::g_stream));
    const int8_t* dcw = (const int8_t*)meta.p;
    const uint8_t* dvf = meta.as<uint8_t>() + GP_MAX_CELLS;
    const int64_t nw = (S.nrows + 63) >> 6;
    const int want = resident_grid(k_pg_hotfold<2, 8>, PG_T, lds, ceil_div(nw, PH_STEP));
    const int64_t wpb = ceil_div(ceil_div(nw, want), PH_STEP) * PH_STEP;
    const int grid = (int)ceil_div(nw, wpb);
    SCRATCH(cnt, (size_t)grid * 8);
#define PGF1(NV_, KE_) LAUNCH("group_hot_fold", (k_pg_hotfold<NV_, KE_>), grid, PG_T, lds, S, hot, \
                              nhot, dcw, dvf, wpb, A, cnt.as<int64_t>(), hot_cells)
#define PGF(NV_)                                      \
    do {                                              \
        if (ke == 8) PGF1(NV_, 8);                    \
        else if (ke == 4) PGF1(NV_, 4);               \
        else PGF1(NV_, 0);                            \
    } while (0)
    switch (nv) {
    case 0: PGF(0); break;
    case 1: PGF(1); break;
    case 2: PGF(2); break;
    case 3: PGF(3); break;
    default: PGF(4); break;
    }
#undef PGF
#undef PGF1
    std::vector<int64_t> c((size_t)grid);
    HC(hipMemcpyAsync(c.data(), cnt.p, c.size() * 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    first->clear();
    *cold = 0;
    for (int b = 0; b < grid; b++) *cold += c[(size_t)b];
    for (int b = 0; b < grid; b++)
        for (int64_t r = 0; r < c[(size_t)b]; r += PG_JOB) {
            PGJob J;
            J.r0 = (int64_t)b * wpb * 64 + r;
            J.r1 = J.r0 + (c[(size_t)b] - r < PG_JOB ? c[(size_t)b] - r : PG_JOB);
            J.parent = 0;
            J.js = 0;
            first->push_back(J);
        }
    for (PGJob& J : *first) J.nj = (int32_t)first->size();
    return TD_OK;
}

int pht_partition(const GroupPlan& P, const uint64_t* mask, int64_t nrows, double est,
                  int64_t lds_bytes, bool forced, const unsigned long long* hot, int nhot,
                  uint64_t* hot_cells, tdd_pht** out) {
    *out = nullptr;
    if (nrows <= 0) return TD_ERR_NYI;
    const int64_t per = 8 + (int64_t)P.ncells * 8;
    int lcap = 1;
    while ((int64_t)(lcap * 2) * per <= lds_bytes) lcap *= 2;
    if (lcap < 16) return TD_ERR_NYI;
    int bits = 0;
    while (bits < 24 && est / (double)((int64_t)1 << bits) > lcap / 2) bits++;
    if (bits == 0 && forced) bits = 1;
    if (bits == 0) return TD_ERR_NYI;
    tdd_pht* h = new tdd_pht();
    PGRows S;
    memset(&S, 0, sizeof(S));
    S.P = P;
    S.mask = mask;
    S.nrows = nrows;
    h->P = P;
    h->lcap = lcap;
    h->nv = pg_words(P, S, h->cw, h->vf);
    if (h->nv < 0) { delete h; return TD_ERR_NYI; }
    PGBuf A, B;
    if (!pg_alloc(h->mem, nrows, h->nv, true, A, B)) { delete h; return TD_ERR_OOM; }
    int pbits[8], npass;
    pg_split_bits(bits, pbits, &npass);
    std::vector<PGJob> first;
    bool from_records = false;
    h->rows = -1;                                   /* every selected row */
    if (nhot > 0) {
        int frc = pht_hot_fold(S, h->nv, h->cw, h->vf, hot, nhot, hot_cells, A, &first,
                               &h->rows);
        if (frc && frc != TD_ERR_NYI) { delete h; return frc; }
        from_records = frc == 0;
    }
    if (nhot > 0 && !from_records) { delete h; return TD_ERR_NYI; }   /* the hot rows need the fold */
    int rc = pg_scatter<true>(S, h->nv, nrows, pbits, npass, 64, A, B, h->jobs, h->res,
                              from_records ? &first : nullptr);
    if (rc) { delete h; return rc; }
    *out = h;
    return TD_OK;
}

/* Fold the partitions: groups of a part folded by a single job are appended
 * to the list (lkeys / lcells, lcap_out entries, *lcount zeroed); groups of
 * parts split over several jobs merge into the global table (keys / cells,
 * mask + 1 slots plus the extra all-ones slot, already initialised).
 * overflow[0] is set when the global table has no slot for a key,
 * overflow[1] when the list holds fewer than *lcount entries. */
int pht_fold(tdd_pht* h, unsigned long long* gkeys, uint64_t gmask, uint64_t* gcells,
             unsigned long long* lkeys, uint64_t* lcells, unsigned long long* lcount,
             int64_t lcap_out, int* overflow) {
    const int njobs = (int)h->jobs.size();
    if (njobs == 0) return TD_OK;
    SCRATCH(djobs, sizeof(PGJob) * (size_t)njobs);
    SCRATCH(meta, GP_MAX_CELLS + PG_MAXV);
    uint8_t hm[GP_MAX_CELLS + PG_MAXV];
    memcpy(hm, h->cw, GP_MAX_CELLS);
    memcpy(hm + GP_MAX_CELLS, h->vf, PG_MAXV);
    HC(hipMemcpyAsync(djobs.p, h->jobs.data(), sizeof(PGJob) * (size_t)njobs,
                      hipMemcpyHostToDevice, tdd::g_stream));
    HC(hipMemcpyAsync(meta.p, hm, sizeof(hm), hipMemcpyHostToDevice, tdd::g_stream));
    const size_t lds = (size_t)h->lcap * (8 + (size_t)h->P.ncells * 8 + 4);
    const int grid = resident_grid(k_pg_hagg<4>, PG_T, lds, njobs);
    PHList L;
    L.keys = lkeys;
    L.cells = lcells;
    L.count = lcount;
    L.cap = lcap_out;
    const int8_t* dcw = (const int8_t*)meta.p;
    const uint8_t* dvf = meta.as<uint8_t>() + GP_MAX_CELLS;
#define PGH(NV_) LAUNCH("group_pg_hagg", k_pg_hagg<NV_>, grid, PG_T, lds, h->P, h->res, \
                        djobs.as<PGJob>(), njobs, h->lcap, dcw, dvf, gkeys, gmask, gcells, L, \
                        overflow)
    switch (h->nv) {
    case 0: PGH(0); break;
    case 1: PGH(1); break;
    case 2: PGH(2); break;
    case 3: PGH(3); break;
    default: PGH(4); break;
    }
#undef PGH
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}

/* Keys the global table can expect from a partitioned fold: the parts
 * split over several jobs merge there (at most one leaf table of keys per
 * job of such a part), plus an allowance for keys that find no room in
 * their part's leaf table. */
/* Keys the global table takes: every key of a part split over several
 * jobs (at most a table each), plus an allowance for keys that find no
 * slot within the probe limit of their part's table. */
int64_t pht_rows(const tdd_pht* h) { return h->rows; }

int64_t pht_global_keys(const tdd_pht* h) {
    int64_t n = 4096 + 16 * (int64_t)h->jobs.size();
    for (size_t j = 0; j < h->jobs.size(); j++)
        if (h->jobs[j].nj > 1) n += h->lcap;
    return n;
}

void pht_free(tdd_pht* h) {
    if (!h) return;
    (void)hipStreamSynchronize(tdd::g_stream);
    delete h;
}

