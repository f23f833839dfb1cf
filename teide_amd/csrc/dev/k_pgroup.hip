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
 * A pass is the counting scatter of a radix sort over tiles of items
 * (k_px_hist counts, an exclusive scan places, k_px_pass scatters; see
 * the pass section below for the runs per XCD that keep its writes whole
 * cache lines).  A record is the
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
#define PG_MAXB  8                        /* bits per pass: up to 256 sub-partitions */
#define PG_MAXP  (1 << PG_MAXB)
#define PG_MAXV  4                        /* value words per record */
#define PG_JOB   ((int64_t)1 << 17)       /* items per job (hash leaf) */
#define PG_JOB_DA ((int64_t)1 << 20)      /* items per job (direct-array leaf): a part of
                                             fewer items is one job, which stores its
                                             slots instead of merging them with atomics */
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

/* A pass over tiles of PX_TILE consecutive items, each tile within one
 * parent partition.  Pass 1 (one parent) appends each sub-partition's items
 * to one run per XCD: tile t is handled by a workgroup on XCD t mod 8 (the
 * grid is a multiple of 8 and workgroup b runs on XCD b mod 8), so the open
 * end of every run -- a cache line per sub-partition and record array,
 * ~100 KB per XCD for 256 sub-partitions -- stays in that XCD's L2 until it
 * is whole, instead of one open run per workgroup and sub-partition.  Later
 * passes take every tile of a parent on one XCD (the tile list is laid out
 * so) and keep one run per (parent, sub-partition).  A histogram pass counts
 * the items of every run; an exclusive scan of the counts, ordered parent,
 * sub-partition, run, gives the runs' starts, so each partition is
 * contiguous.  The scatter counting-sorts a tile in LDS, reserves each of
 * its sub-partitions' runs with one atomic, and stages the records through
 * LDS one array at a time so consecutive lanes write consecutive
 * addresses of a run. */
#define PX_R    8                         /* items per lane per tile */
#define PX_TILE (PG_T * PX_R)
#define PX_XCD  8

struct PXTile {
    int64_t r0, r1;                       /* items [r0, r1) */
    int64_t parent;
};

struct PXFirst {                          /* pass-1 tiles laid out on the device */
    const PXTile* tiles;
    const int*    count;                  /* their number, on the device */
    int64_t       bound;                  /* at most this many */
    int64_t       items;                  /* at most this many records */
};

template <bool HK>
__device__ __forceinline__ uint32_t px_digit(uint64_t key, int shift, int PL) {
    return (uint32_t)(pg_bits<HK>(key) >> shift) & (uint32_t)(PL - 1);
}

/* Value word q of item i: a record word, or the row's aggregate input (its
 * index for FIRST / LAST); KE != 0: the value columns are plain 8-byte. */
template <bool ROWS, int KE>
__device__ __forceinline__ uint64_t px_val(const PGRows& S, const PGBuf& in, int q, int64_t i) {
    if (!ROWS) return in.val[q][i];
    if (S.vsrc[q] < 0) return (uint64_t)i;
    const tdd_col_t& c = S.P.aggs[S.vsrc[q]];
    return KE != 0 ? (uint64_t)((const int64_t*)c.data)[i] : (uint64_t)ld_i64(c, i);
}

/* Items per run: C[(parent * PL + s) * XR + run]. */
template <bool ROWS, bool HK, int KE = 0>
__global__ __launch_bounds__(PG_T) void k_px_hist(PGRows S, PGBuf in, const PXTile* tiles,
                                                  int ntiles, const int* dn, int shift, int PL,
                                                  int XR, uint32_t* C) {
    __shared__ uint32_t h[PG_MAXP];
    if (dn) ntiles = *dn;                                 /* a list laid out on the device */
    const int lane = lane_id();
    for (int s = threadIdx.x; s < PL; s += PG_T) h[s] = 0;
    __syncthreads();
    int64_t par = -1;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const PXTile T = tiles[t];                        /* uniform across the block */
        if (T.parent != par) {
            if (par >= 0) {                               /* later passes: flush a parent */
                __syncthreads();
                for (int s = threadIdx.x; s < PL; s += PG_T)
                    if (h[s]) { atomicAdd(&C[((int64_t)par * PL + s) * XR], h[s]); h[s] = 0; }
                __syncthreads();
            }
            par = T.parent;
        }
        uint64_t key[PX_R], vd[PG_MAXV];
        bool live[PX_R];
        const tdd_col_t& kc = S.P.keys[0];
        if (ROWS && KE == 4 && !S.mask && T.r1 - T.r0 == PX_TILE && ((uintptr_t)kc.data & 15) == 0) {
            /* a full tile of 4-byte keys, no selection: 8 consecutive keys a
             * lane in two 16-byte loads (the count does not care which lane
             * holds which row) */
            const uint4* p = (const uint4*)((const uint32_t*)kc.data + T.r0 +
                                            (int64_t)threadIdx.x * PX_R);
            const uint4 a = p[0], b = p[1];
            const uint32_t w[PX_R] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int k = 0; k < PX_R; k++) {
                const int64_t raw = kc.type == TD_SYM ? (int64_t)w[k] : (int64_t)(int32_t)w[k];
                key[k] = HK && S.P.kstride[0] == 0 ? (uint64_t)raw
                                                   : (uint64_t)((raw - S.P.kmin[0]) * S.P.kstride[0]);
                live[k] = true;
            }
        } else {
#pragma unroll
            for (int k = 0; k < PX_R; k++)                /* every load of the tile first */
                live[k] = pg_item<ROWS, HK, 0, KE>(S, in, T.r0 + (int64_t)k * PG_T + threadIdx.x,
                                                   T.r1, key[k], vd);
        }
#pragma unroll
        for (int k = 0; k < PX_R; k++) {
            const int s = live[k] ? (int)px_digit<HK>(key[k], shift, PL) : -1;
            const uint64_t lm = __ballot(live[k]);
            if (!lm) continue;
            /* a wave whose items all go to one sub-partition counts them at once */
            const int first = __ffsll((unsigned long long)lm) - 1;
            const int s0 = __shfl(s, first);
            if (__ballot(live[k] && s != s0) == 0) {
                if (lane == first) atomicAdd(&h[s0], (uint32_t)__popcll(lm));
            } else if (live[k]) {
                atomicAdd(&h[s], 1u);
            }
        }
    }
    __syncthreads();
    const int run = XR > 1 ? (int)(blockIdx.x % PX_XCD) : 0;
    if (par >= 0)
        for (int s = threadIdx.x; s < PL; s += PG_T)
            if (h[s]) atomicAdd(&C[((int64_t)par * PL + s) * XR + run], h[s]);
}

/* The scatter of one pass: runs' cursors in `cur` (the scanned counts). */
template <bool ROWS, bool HK, int NV, int KE = 0>
__global__ __launch_bounds__(PG_T) void k_px_pass(PGRows S, PGBuf in, PGBuf out,
                                                  const PXTile* tiles, int ntiles, const int* dn,
                                                  int shift, int PL, int XR,
                                                  unsigned long long* cur, int64_t nout) {
    __shared__ uint32_t hist[PG_MAXP], loff[PG_MAXP];
    if (dn) ntiles = *dn;                                 /* a list laid out on the device */
    __shared__ unsigned long long gbase[PG_MAXP];
    __shared__ uint8_t spart[PX_TILE];
    __shared__ uint64_t stage[PX_TILE];
    const int lane = lane_id();
    const int run = XR > 1 ? (int)(blockIdx.x % PX_XCD) : 0;
    /* rows of 4-byte keys and 8-byte values, no selection, 16-byte aligned:
     * a full tile's lane takes 8 consecutive rows with 16-byte loads */
    bool vcols = ROWS && KE == 4 && !S.mask && ((uintptr_t)S.P.keys[0].data & 15) == 0;
    for (int q = 0; q < NV; q++)
        if (S.vsrc[q] >= 0 && ((uintptr_t)S.P.aggs[S.vsrc[q]].data & 15) != 0) vcols = false;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const PXTile T = tiles[t];                        /* uniform across the block */
        for (int s = threadIdx.x; s < PL; s += PG_T) hist[s] = 0;
        __syncthreads();
        uint64_t key[PX_R], vd[PG_MAXV];
        uint32_t dig[PX_R], at[PX_R], live = 0;
        bool vrow = T.r1 - T.r0 == PX_TILE && (ROWS ? vcols : true);
        if (!ROWS && vrow) {                              /* records: 16-byte aligned at r0 */
            vrow = (HK ? ((uintptr_t)(in.key + T.r0) & 15) : ((uintptr_t)(in.slot + T.r0) & 15)) == 0;
            for (int q = 0; q < NV; q++)
                if ((uintptr_t)(in.val[q] + T.r0) & 15) vrow = false;
        }
        const int64_t i0 = vrow ? T.r0 + (int64_t)threadIdx.x * PX_R : T.r0 + threadIdx.x;
        const int64_t istep = vrow ? 1 : PG_T;            /* item k of the lane: i0 + k * istep */
        if (vrow && !ROWS) {
            if (HK) {
                const ulonglong2* p = (const ulonglong2*)(in.key + i0);
#pragma unroll
                for (int h = 0; h < PX_R / 2; h++) {
                    const ulonglong2 x = p[h];
                    key[2 * h] = x.x;
                    key[2 * h + 1] = x.y;
                }
            } else {
                const uint4* p = (const uint4*)(in.slot + i0);
                const uint4 a = p[0], b = p[1];
                key[0] = a.x; key[1] = a.y; key[2] = a.z; key[3] = a.w;
                key[4] = b.x; key[5] = b.y; key[6] = b.z; key[7] = b.w;
            }
            live = (1u << PX_R) - 1;
        } else if (vrow) {
            const tdd_col_t& kc = S.P.keys[0];
            const uint4* p = (const uint4*)((const uint32_t*)kc.data + i0);
            const uint4 a = p[0], b = p[1];
            const uint32_t w[PX_R] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int k = 0; k < PX_R; k++) {
                const int64_t raw = kc.type == TD_SYM ? (int64_t)w[k] : (int64_t)(int32_t)w[k];
                key[k] = HK && S.P.kstride[0] == 0 ? (uint64_t)raw
                                                   : (uint64_t)((raw - S.P.kmin[0]) * S.P.kstride[0]);
            }
            live = (1u << PX_R) - 1;
        } else {
#pragma unroll
            for (int k = 0; k < PX_R; k++)
                if (pg_item<ROWS, HK, 0, KE>(S, in, i0 + (int64_t)k * istep, T.r1, key[k], vd))
                    live |= 1u << k;
        }
#pragma unroll
        for (int k = 0; k < PX_R; k++) {
            dig[k] = at[k] = 0;
            if ((live >> k) & 1) {
                dig[k] = px_digit<HK>(key[k], shift, PL);
                at[k] = atomicAdd(&hist[dig[k]], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < WAVE) {      /* wave 0: tile offsets and run reservations, 4 a lane */
            uint32_t c4[PG_MAXP / WAVE], sum = 0;
#pragma unroll
            for (int q = 0; q < PG_MAXP / WAVE; q++) {
                const int s = lane * (PG_MAXP / WAVE) + q;
                c4[q] = s < PL ? hist[s] : 0u;
                sum += c4[q];
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
                    loff[s] = ex;
                    gbase[s] = c4[q] ? atomicAdd(&cur[((int64_t)T.parent * PL + s) * XR + run],
                                                 (unsigned long long)c4[q])
                                     : 0ull;
                }
                ex += c4[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PX_R; k++)
            if ((live >> k) & 1) {
                at[k] += loff[dig[k]];                    /* sorted place in the tile */
                spart[at[k]] = (uint8_t)dig[k];
                stage[at[k]] = key[k];
            }
        __syncthreads();
        const int cnt = (int)(loff[PL - 1] + hist[PL - 1]);
        int64_t pos[PX_R];
#pragma unroll
        for (int k = 0; k < PX_R; k++) {
            const int i = threadIdx.x + k * PG_T;
            pos[k] = -1;
            if (i < cnt) {
                const int d = spart[i];
                pos[k] = (int64_t)gbase[d] + (i - (int)loff[d]);
                if (pos[k] >= nout) pos[k] = -1;          /* never taken when counts agree */
            }
        }
#pragma unroll
        for (int k = 0; k < PX_R; k++) {
            if (pos[k] < 0) continue;
            const uint64_t x = stage[threadIdx.x + k * PG_T];
            if (HK) out.key[pos[k]] = x;
            else out.slot[pos[k]] = (uint32_t)x;
        }
#pragma unroll
        for (int q = 0; q < NV; q++) {
            uint64_t v[PX_R];
            if (vrow && (!ROWS || S.vsrc[q] >= 0)) {      /* 8 consecutive values, 16-byte loads */
                const ulonglong2* p =
                    (const ulonglong2*)(ROWS ? (const uint64_t*)S.P.aggs[S.vsrc[q]].data + i0
                                             : in.val[q] + i0);
#pragma unroll
                for (int h = 0; h < PX_R / 2; h++) {
                    const ulonglong2 x = p[h];
                    v[2 * h] = x.x;
                    v[2 * h + 1] = x.y;
                }
            } else {
#pragma unroll
                for (int k = 0; k < PX_R; k++)            /* the word's loads first */
                    v[k] = (live >> k) & 1 ? px_val<ROWS, KE>(S, in, q, i0 + (int64_t)k * istep)
                                           : 0ull;
            }
            __syncthreads();                              /* the previous array is out */
#pragma unroll
            for (int k = 0; k < PX_R; k++)
                if ((live >> k) & 1) stage[at[k]] = v[k];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PX_R; k++)
                if (pos[k] >= 0) out.val[q][pos[k]] = stage[threadIdx.x + k * PG_T];
        }
        __syncthreads();                                  /* the tile's LDS is free */
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
        if (J.nj == 1) {                       /* the part's only job: its slots are its own */
            for (int64_t w = threadIdx.x; w < nl * nc; w += blockDim.x) {
                const int64_t s = w / nc;
                if (lt[s * nc] == 0 || (int64_t)base + s >= n_slots) continue;
                tab[(int64_t)base * nc + w] = lt[w];
            }
        } else {
            for (int64_t s = threadIdx.x; s < nl; s += blockDim.x) {
                const uint64_t* c = &lt[s * nc];
                if (c[0] == 0 || (int64_t)base + s >= n_slots) continue;
                cells_merge(&tab[((int64_t)base + s) * nc], c, P);
            }
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
__global__ __launch_bounds__(PG_T) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_pg_hagg(GroupPlan P, PGBuf in, const PGJob* jobs,
                                                  int njobs, int lcap, const int8_t* cw_dev,
                                                  const uint8_t* vf64_dev,
                                                  unsigned long long* gkeys, uint64_t gmask,
                                                  uint64_t* gcells, PHList L, int* overflow) {
    extern __shared__ uint64_t sm[];
    unsigned long long* lk = (unsigned long long*)sm;
    uint64_t* lc = sm + lcap;
    uint32_t* linv = (uint32_t*)(lc + (int64_t)lcap * P.ncells);   /* list order -> entry */
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
                if (lk[i] != EMPTY_KEY && lc[(int64_t)i * nc] != 0) linv[atomicAdd(&nloc, 1u)] = i;
            __syncthreads();
            if (threadIdx.x == 0) lbase = atomicAdd(L.count, (unsigned long long)nloc);
            __syncthreads();
            /* the job's groups are one run of the list: keys, then the cells
             * word by word, consecutive lanes on consecutive words */
            const int64_t n = lbase + nloc <= (unsigned long long)L.cap ? (int64_t)nloc
                            : lbase < (unsigned long long)L.cap ? L.cap - (int64_t)lbase : 0;
            if (threadIdx.x == 0 && n < (int64_t)nloc) overflow[1] = 1;   /* the list */
            for (int64_t g = threadIdx.x; g < n; g += blockDim.x) L.keys[lbase + g] = lk[linv[g]];
            for (int64_t w = threadIdx.x; w < n * nc; w += blockDim.x) {
                const int64_t g = w / nc, c = w - g * nc;
                L.cells[(int64_t)lbase * nc + w] = lc[(int64_t)linv[g] * nc + c];
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
#define PH_U    4                         /* 64-row words per wave per step */
#define PH_STEP (PG_W * PH_U)             /* words per workgroup step */

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
        uint64_t key[PH_U];
        uint64_t v[PH_U][PG_MAXV];
        bool live[PH_U];
        uint64_t m[PH_U];
#pragma unroll
        for (int u = 0; u < PH_U; u++) {           /* every load of the step first */
            const int64_t w = ws + wv * PH_U + u;
            live[u] = w < w_end && pg_item<true, true, NV, KE>(S, none, w * 64 + lane, S.nrows,
                                                                key[u], v[u]);
        }
#pragma unroll
        for (int u = 0; u < PH_U; u++) {
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
            if (lane == 0) wcnt[wv * PH_U + u] = (uint32_t)__popcll(m[u]);
        }
        __syncthreads();
        uint32_t pre = cursor;                    /* records of the step's earlier words */
        for (int i = 0; i < wv * PH_U; i++) pre += wcnt[i];
#pragma unroll
        for (int u = 0; u < PH_U; u++) {
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
    /* this workgroup's hot cells, row b of the partials (k_pg_hotreduce
     * merges them: 1,024 workgroups merging into the same few thousand
     * cells with atomics queue on those addresses) */
    uint64_t* part = gcells + (int64_t)blockIdx.x * nhot * nc;
    for (int64_t w = threadIdx.x; w < (int64_t)nhot * nc; w += PG_T) part[w] = cells[w];
}

/* Hot cells: partials [nwg][nhot * ncells] folded per cell over a chunk of
 * workgroups (blockIdx.y), then merged into the hot cells with one atomic
 * per (chunk, cell). */
__global__ __launch_bounds__(256) void k_pg_hotreduce(GroupPlan P, const uint64_t* part, int nwg,
                                                      int64_t ncell, int per, uint64_t* gcells) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ncell) return;
    const int kind = P.kind[i % P.ncells];
    const int w0 = blockIdx.y * per, w1 = w0 + per < nwg ? w0 + per : nwg;
    uint64_t a = cell_identity(kind);
    double af = 0.0;
    for (int w = w0; w < w1; w++) {
        const uint64_t s = part[(int64_t)w * ncell + i];
        switch (kind) {
        case C_ADD_F64: case C_SQ_F64: af += __longlong_as_double((long long)s); break;
        case C_MIN_I64: case C_MIN_F64: case C_MIN_ROW: a = (int64_t)s < (int64_t)a ? s : a; break;
        case C_MAX_I64: case C_MAX_F64: case C_MAX_ROW: a = (int64_t)s > (int64_t)a ? s : a; break;
        default: a += s; break;                              /* count, integer sums */
        }
    }
    switch (kind) {
    case C_ADD_F64: case C_SQ_F64: atomicAdd((double*)&gcells[i], af); break;
    case C_MIN_I64: case C_MIN_F64: case C_MIN_ROW: cell_min(&gcells[i], (int64_t)a); break;
    case C_MAX_I64: case C_MAX_F64: case C_MAX_ROW: cell_max(&gcells[i], (int64_t)a); break;
    default: atomicAdd((unsigned long long*)&gcells[i], (unsigned long long)a); break;
    }
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

/* The next pass's tile list from a finished pass's run starts O (parent
 * p = the finished pass's partition, items [O[p * XR], O[(p + 1) * XR])):
 * the tiles of parent p go to XCD p mod 8, tile t of the list to XCD t mod 8
 * (row i of XCD x at t = i * 8 + x), empty tiles pad the shorter XCD lists.
 * k_px_tscan (one workgroup): each parent's first row on its XCD (base) and
 * the list length (*ntiles = 8 x the longest XCD list); k_px_twrite (one
 * workgroup per parent, then one per XCD for the padding) writes it. */
__global__ __launch_bounds__(1024) void k_px_tscan(const int64_t* O, int64_t nparts, int XR,
                                                   int64_t* base, int* ntiles) {
    __shared__ int64_t v[1024];
    __shared__ int64_t run[PX_XCD];
    if (threadIdx.x < PX_XCD) run[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t c = 0; c < nparts; c += 1024) {       /* chunks of 1,024 parents */
        const int64_t p = c + threadIdx.x;
        const int64_t nt = p < nparts ? (O[(p + 1) * XR] - O[p * XR] + PX_TILE - 1) / PX_TILE : 0;
        v[threadIdx.x] = nt;
        __syncthreads();
        for (int d = PX_XCD; d < 1024; d <<= 1) {      /* inclusive scan within p mod 8 */
            const int64_t add = (int)threadIdx.x >= d ? v[threadIdx.x - d] : 0;
            __syncthreads();
            v[threadIdx.x] += add;
            __syncthreads();
        }
        const int x = (int)(threadIdx.x % PX_XCD);
        if (p < nparts) base[p] = run[x] + v[threadIdx.x] - nt;
        __syncthreads();
        if (threadIdx.x >= 1024 - PX_XCD) run[x] += v[threadIdx.x];   /* each class's total */
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int64_t longest = 0;
        for (int x = 0; x < PX_XCD; x++) longest = run[x] > longest ? run[x] : longest;
        base[nparts + PX_XCD] = longest;
        for (int x = 0; x < PX_XCD; x++) base[nparts + x] = run[x];
        *ntiles = (int)(longest * PX_XCD);
    }
}

__global__ __launch_bounds__(256) void k_px_twrite(const int64_t* O, int64_t nparts, int XR,
                                                   const int64_t* base, PXTile* tl) {
    const int64_t b = blockIdx.x;
    if (b < nparts) {                                   /* parent b's tiles */
        const int64_t r0 = O[b * XR], r1 = O[(b + 1) * XR], x = b % PX_XCD;
        const int64_t nt = (r1 - r0 + PX_TILE - 1) / PX_TILE;
        for (int64_t k = threadIdx.x; k < nt; k += 256) {
            const int64_t a = r0 + k * PX_TILE;
            tl[(base[b] + k) * PX_XCD + x] = PXTile{a, a + PX_TILE < r1 ? a + PX_TILE : r1, b};
        }
    } else {                                            /* XCD x's padding */
        const int64_t x = b - nparts, longest = base[nparts + PX_XCD];
        for (int64_t i = base[nparts + x] + threadIdx.x; i < longest; i += 256)
            tl[i * PX_XCD + x] = PXTile{0, 0, 0};
    }
}

/* Pass-1 tiles over the hot fold's record regions (its workgroup b wrote
 * counts[b] records from b * span on): one workgroup lays out every
 * region's tiles back to back and their count. */
__global__ __launch_bounds__(1024) void k_px_ftiles(const int64_t* counts, int nreg, int64_t span,
                                                    PXTile* tl, int* ntiles) {
    __shared__ int64_t v[1024];
    __shared__ int64_t run;
    if (threadIdx.x == 0) run = 0;
    __syncthreads();
    for (int c = 0; c < nreg; c += 1024) {
        const int b = c + (int)threadIdx.x;
        const int64_t cnt = b < nreg ? counts[b] : 0;
        const int64_t nt = (cnt + PX_TILE - 1) / PX_TILE;
        v[threadIdx.x] = nt;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {           /* inclusive scan of the tile counts */
            const int64_t add = (int)threadIdx.x >= d ? v[threadIdx.x - d] : 0;
            __syncthreads();
            v[threadIdx.x] += add;
            __syncthreads();
        }
        const int64_t base = run + v[threadIdx.x] - nt, r0 = (int64_t)b * span;
        for (int64_t k = 0; k < nt; k++)
            tl[base + k] = PXTile{r0 + k * PX_TILE, r0 + ((k + 1) * PX_TILE < cnt ? (k + 1) * PX_TILE : cnt), 0};
        __syncthreads();
        if (threadIdx.x == 1023) run += v[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *ntiles = (int)run;
}

/* Histogram and scatter of one pass over the tiles at d_tiles (grids a
 * multiple of 8, so a tile's run is the XCD its workgroup runs on); dn: the
 * tile count on the device (a list laid out there), ntiles its bound. */
template <bool ROWS, bool HK>
static int px_launch(int nv, const PGRows& S, const PGBuf& in, const PGBuf& out,
                     const PXTile* d_tiles, int ntiles, const int* dn, int shift, int PL, int XR,
                     uint32_t* C, int64_t* O, unsigned long long* cur, int64_t nrun, int64_t nout) {
    /* rows of one 4- or 8-byte key and 8-byte value columns: plain loads */
    int ke = 0;
    if (ROWS && S.P.n_keys == 1 && S.P.keys[0].data) {
        ke = pg_plain(S.P.keys[0]);
        for (int k = 0; k < nv; k++)
            if (S.vsrc[k] >= 0 && pg_plain(S.P.aggs[S.vsrc[k]]) != 8) ke = 0;
    }
    const int hke = ROWS && S.P.n_keys == 1 && S.P.keys[0].data ? pg_plain(S.P.keys[0]) : 0;
    auto grid8 = [&](int g) { g = g / PX_XCD * PX_XCD; return g < PX_XCD ? PX_XCD : g; };
    const int gh = grid8(stream_grid(ntiles, 4));
    int rc = tdd_fill(C, 0, (size_t)nrun * 4);
    if (rc) return rc;
#define PXH(KE_) LAUNCH("group_pg_hist", (k_px_hist<ROWS, HK, KE_>), gh, PG_T, 0, S, in, d_tiles, \
                        ntiles, dn, shift, PL, XR, C)
    if (ROWS && hke == 8) PXH((ROWS ? 8 : 0));
    else if (ROWS && hke == 4) PXH((ROWS ? 4 : 0));
    else PXH(0);
#undef PXH
    rc = tdd::scan_u32_i64_dup(C, nrun, O, (int64_t*)cur);   /* run starts, and cursors */
    if (rc) return rc;
    const int gp = grid8(resident_grid(k_px_pass<ROWS, HK, 2, 0>, PG_T, 0, ntiles));
#define PXP1(NV_, KE_) LAUNCH("group_pg_pass", (k_px_pass<ROWS, HK, NV_, KE_>), gp, PG_T, 0, S, in, \
                              out, d_tiles, ntiles, dn, shift, PL, XR, cur, nout)
#define PXP(NV_)                                          \
    do {                                                  \
        if (ROWS && ke == 8) PXP1(NV_, (ROWS ? 8 : 0));   \
        else if (ROWS && ke == 4) PXP1(NV_, (ROWS ? 4 : 0)); \
        else PXP1(NV_, 0);                                \
    } while (0)
    switch (nv) {
    case 0: PXP(0); break;
    case 1: PXP(1); break;
    case 2: PXP(2); break;
    case 3: PXP(3); break;
    default: PXP(4); break;
    }
#undef PXP
#undef PXP1
    return TD_OK;
}

/* Jobs over the partitions of a finished pass: partition q = parent * PL + s
 * holds items [start(q), start(q) + count(q)). */
static void pg_jobs_from(const std::vector<int64_t>& start, const std::vector<int64_t>& count,
                         std::vector<PGJob>& jobs, int64_t per) {
    size_t total = 0;
    for (size_t q = 0; q < start.size(); q++) total += (size_t)ceil_div(count[q], per);
    jobs.resize(total);
    PGJob* J = jobs.data();
    for (size_t q = 0; q < start.size(); q++) {
        if (count[q] == 0) continue;
        const int js = (int)(J - jobs.data());
        const int nj = (int)ceil_div(count[q], per);
        const int64_t end = start[q] + count[q];
        for (int k = 0; k < nj; k++, J++) {
            J->r0 = start[q] + (int64_t)k * per;
            J->r1 = J->r0 + per < end ? J->r0 + per : end;
            J->parent = (int64_t)q;
            J->js = js;
            J->nj = nj;
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
                      const PXFirst* f0 = nullptr) {
    /* pass-1 tiles: the rows, or (f0) tiles over records in A laid out on
     * the device */
    std::vector<PXTile> tl;
    if (!f0)
        for (int64_t r = 0; r < nrows; r += PX_TILE)
            tl.push_back(PXTile{r, r + PX_TILE < nrows ? r + PX_TILE : nrows, 0});
    PGBuf in = A, outb = B;              /* the first pass reads rows, or A's records */
    int shift = top;
    int64_t parents = 1;
    std::vector<int64_t> O;
    int XR = PX_XCD;
    /* pass 1's tiles from the host; later passes' lists are laid out on the
     * device from the finished pass's run starts, with no host round trip */
    int64_t items = 0;
    for (const PXTile& T : tl) items += T.r1 - T.r0;
    Scratch dtl(sizeof(PXTile) * (tl.empty() ? 1 : tl.size()));
    Scratch dn(8);
    if (!dtl.ok() || !dn.ok()) return TD_ERR_OOM;
    if (!tl.empty())
        HC(hipMemcpyAsync(dtl.p, tl.data(), sizeof(PXTile) * tl.size(), hipMemcpyHostToDevice,
                          tdd::g_stream));
    int ntiles = (int)tl.size();
    const int* dcount = nullptr;                      /* pass 1: the host's count */
    const PXTile* t0 = dtl.as<PXTile>();
    if (f0) {                                         /* pass 1: tiles laid out on the device */
        t0 = f0->tiles;
        ntiles = (int)f0->bound;
        dcount = f0->count;
        items = f0->items;
    }
    for (int pass = 0; pass < npass; pass++) {
        POLL_CANCEL();
        const int PL = 1 << pbits[pass];
        shift -= pbits[pass];
        XR = pass == 0 ? PX_XCD : 1;
        const int64_t nrun = parents * PL * XR;
        SCRATCH(C, (size_t)nrun * 4);
        Scratch Od(((size_t)nrun + 1) * 8);
        SCRATCH(cur, (size_t)nrun * 8);
        if (!Od.ok()) return TD_ERR_OOM;
        const bool rows = pass == 0 && !f0;
        const PXTile* tp = pass == 0 ? t0 : dtl.as<PXTile>();
        int rc = rows ? px_launch<true, HK>(nv, S, in, outb, tp, ntiles, dcount,
                                            shift, PL, XR, C.as<uint32_t>(), Od.as<int64_t>(),
                                            cur.as<unsigned long long>(), nrun, nrows)
                      : px_launch<false, HK>(nv, S, in, outb, tp, ntiles, dcount,
                                             shift, PL, XR, C.as<uint32_t>(), Od.as<int64_t>(),
                                             cur.as<unsigned long long>(), nrun, nrows);
        if (rc) return rc;
        const int64_t nparts = parents * PL;
        if (pass + 1 < npass) {
            /* bound of the next list: every XCD list at most the items' tiles
             * plus one partial tile per parent */
            const int64_t bound = (ceil_div(items, (int64_t)PX_TILE) + nparts) * PX_XCD;
            Scratch base(((size_t)nparts + PX_XCD + 1) * 8);
            if (!base.ok() || !dtl.reset(sizeof(PXTile) * (size_t)bound)) return TD_ERR_OOM;
            LAUNCH("group_pg_tiles", k_px_tscan, 1, 1024, 0, Od.as<int64_t>(), nparts, XR,
                   base.as<int64_t>(), dn.as<int>());
            LAUNCH("group_pg_tiles", k_px_twrite, (int)(nparts + PX_XCD), 256, 0,
                   Od.as<int64_t>(), nparts, XR, base.as<int64_t>(), dtl.as<PXTile>());
            ntiles = (int)bound;
            dcount = dn.as<int>();
        } else {
            O.resize((size_t)nrun + 1);
            HC(hipMemcpyAsync(O.data(), Od.p, O.size() * 8, hipMemcpyDeviceToHost, tdd::g_stream));
            HC(hipStreamSynchronize(tdd::g_stream));
        }
        parents = nparts;
        std::swap(in, outb);
    }
    /* leaf jobs: partition q holds [O[q * XR], O[(q + 1) * XR]) */
    std::vector<int64_t> start((size_t)parents), count((size_t)parents);
    for (int64_t q = 0; q < parents; q++) {
        start[(size_t)q] = O[(size_t)(q * XR)];
        count[(size_t)q] = O[(size_t)((q + 1) * XR)] - start[(size_t)q];
    }
    pg_jobs_from(start, count, jobs, HK ? PG_JOB : PG_JOB_DA);
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
/* The hot-key fold of the rows into A (see k_pg_hotfold); *f0 receives the
 * first pass's tiles over the workgroups' record regions, laid out on the
 * device (tiles and count in ftl / fcnt).  TD_ERR_NYI when the LDS blocks
 * do not fit. */
static int pht_hot_fold(const PGRows& S, int nv, const int8_t* cw, const uint8_t* vf,
                        const unsigned long long* hot, int nhot, uint64_t* hot_cells,
                        const PGBuf& A, Scratch& ftl, Scratch& fcnt, PXFirst* f0) {
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
    SCRATCH(hpart, (size_t)grid * nhot * P.ncells * 8);   /* per workgroup hot cells */
#define PGF1(NV_, KE_) LAUNCH("group_hot_fold", (k_pg_hotfold<NV_, KE_>), grid, PG_T, lds, S, hot, \
                              nhot, dcw, dvf, wpb, A, cnt.as<int64_t>(), hpart.as<uint64_t>())
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
    {
        const int64_t ncell = (int64_t)nhot * P.ncells;
        const int per = 64;
        dim3 rg((unsigned)ceil_div(ncell, 256), (unsigned)ceil_div(grid, per));
        LAUNCH("group_hot_fold", k_pg_hotreduce, rg, 256, 0, P, (const uint64_t*)hpart.as<uint64_t>(),
               grid, ncell, per, hot_cells);
    }
    /* workgroup b's region holds at most wpb * 64 records */
    const int64_t bound = (int64_t)grid * ceil_div(wpb * 64, (int64_t)PX_TILE);
    if (!ftl.reset(sizeof(PXTile) * (size_t)bound) || !fcnt.reset(8)) return TD_ERR_OOM;
    LAUNCH("group_pg_tiles", k_px_ftiles, 1, 1024, 0, (const int64_t*)cnt.as<int64_t>(), grid,
           wpb * 64, ftl.as<PXTile>(), fcnt.as<int>());
    f0->tiles = ftl.as<PXTile>();
    f0->count = fcnt.as<int>();
    f0->bound = bound;
    f0->items = S.nrows;
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
    /* parts of at most fill% of a table's entries (by the estimate): 512
     * entries at 60 % (~300 keys a part) leave room for four leaf
     * workgroups per CU, which hide the per-part latency (table set-up,
     * the part's loads, the list append) better than two with 1,024 */
    const double fill = TUNE_INT("TEIDE_PH_FILL_PCT", 60) / 100.0;
    int bits = 0;
    while (bits < 24 && est / (double)((int64_t)1 << bits) > lcap * fill) bits++;
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
    bool from_records = false;
    h->rows = -1;                                   /* every selected row */
    Scratch ftl(8), fcnt(8);
    PXFirst f0;
    memset(&f0, 0, sizeof(f0));
    if (nhot > 0) {
        int frc = pht_hot_fold(S, h->nv, h->cw, h->vf, hot, nhot, hot_cells, A, ftl, fcnt, &f0);
        if (frc && frc != TD_ERR_NYI) { delete h; return frc; }
        from_records = frc == 0;
    }
    if (nhot > 0 && !from_records) { delete h; return TD_ERR_NYI; }   /* the hot rows need the fold */
    int rc = pg_scatter<true>(S, h->nv, nrows, pbits, npass, 64, A, B, h->jobs, h->res,
                              from_records ? &f0 : nullptr);
    if (rc) { delete h; return rc; }
    if (from_records) {                             /* the rows the passes carried: the cold ones */
        h->rows = 0;
        for (const PGJob& J : h->jobs) h->rows += J.r1 - J.r0;
    }
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
    return TD_OK;                                   /* the caller reads the flags back */
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

