/*
 * k_batch.hip -- many fused filter -> join -> group-by queries per launch.
 *
 * A batch holds queries of one shape (same group keys, aggregate ops and
 * column types); each has its own probe and build tables, predicate
 * constants and key ranges.  Every phase is one launch over a flat list of
 * work tiles (query, row range), so a batch of thousands of small queries
 * costs a handful of launches instead of thousands:
 *
 *   map      one fill of all slot maps, one scatter of every build row's
 *            group slot, one count of occupied entries per query (unique
 *            build keys <=> count == build rows);
 *   probe    a persistent grid walks the probe tiles: star predicate, slot
 *            map gather, and a fold into the query's group table -- in LDS
 *            per tile when the table is small (flushed with one atomic per
 *            touched cell); a query whose table does not fit writes its
 *            passing rows as records (slot + aggregate inputs) and counts
 *            them per group slice instead;
 *   bucket   those records counting-sorted by slice (a slice's table fits
 *            LDS), then one workgroup per (query, slice) folds its records
 *            in LDS and stores the slice of the table -- no global atomics
 *            per row (they bound the global-table fold at ~26 G/s);
 *   emit     one workgroup per query counts non-empty slots, then writes its
 *            groups in ascending key order into shared output columns at the
 *            query's offset.
 *
 * After the first run the launch sequence (which needs no host decisions
 * any more) is captured as a hipGraph and replayed.  Tile size and block
 * size come from a small autotuner that times the probe launch on the
 * batch itself the first time a batch of a given size class is prepared.
 */
#include "group_plan.h"
#include "tile.h"
#include "star.h"
#include "join_keys.h"

#include <map>
#include <vector>

#define BQ_R 4                      /* rows per lane per tile step */
#define BQ_LDS_MAX (48 * 1024)      /* per-tile LDS group table limit */
#define BQ_MAX_SL  4096             /* group slices of a bucketed query */
#define BQ_SC      4096             /* records per bucket-scatter tile */
#define BQ_SCT     256              /* threads of a bucket-scatter workgroup */

/* Per-query device descriptor. */
struct BQuery {
    GroupPlan P;                    /* build-side keys, probe-side aggregates */
    TileCols  T;                    /* key slot 0, predicate slots, aggregate sources */
    StarPred  SP;
    uint8_t   cell_slot[GP_MAX_CELLS];
    int32_t   map_u16, lds;
    int32_t   active, pad;          /* 1: batched (else run on its own) */
    int64_t   nl, nr;
    int64_t   kmin, range;          /* build key bounds */
    tdd_col_t rk;
    int64_t   map_off;              /* bytes into the map arena */
    int64_t   tab_off;              /* cells into the table arena */
    int64_t   n_slots;
    int32_t   nsl, sg;              /* bucketed fold: group slices, slots per slice (0: none) */
    int64_t   rec_off;              /* its record region in the record buffers */
    int64_t   sl_off;               /* its nsl + 1 slice entries */
    int32_t   nrw, pad2;            /* record value words */
    uint8_t   rw_slot[TL_NC];       /* word w holds tile slot rw_slot[w] */
    uint8_t   cell_rw[GP_MAX_CELLS];/* cell c reads record word cell_rw[c] */
    int32_t   hash, hbits;          /* hash join (build keys not a dense unique range), its radix bits */
    int64_t   hb_off, hp_off;       /* its (1 << hbits) + 1 partition entries per side */
    int64_t   hbr_off, hpr_off;     /* its build / probe records */
};

struct BRec {                       /* records of bucketed queries (SoA) */
    uint32_t* slot;
    uint64_t* val[TL_NC - 1];
};

struct BTile {
    int32_t q, pad;
    int64_t r0, r1;                 /* row range within the query */
};

/* ---------------------------------------------------------------- maps -- */

__global__ __launch_bounds__(256) void k_bq_scatter(const BQuery* Q, const BTile* tiles,
                                                    int64_t n_tiles, uint8_t* maps) {
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        for (int64_t r = W.r0 + threadIdx.x; r < W.r1; r += blockDim.x) {
            const int64_t i = ld_i64(q.rk, r) - q.kmin;
            int64_t slot = 0;
            for (int k = 0; k < q.P.n_keys; k++)
                slot += (ld_i64(q.P.keys[k], r) - q.P.kmin[k]) * q.P.kstride[k];
            if (q.map_u16) ((uint16_t*)(maps + q.map_off))[i] = (uint16_t)slot;
            else ((int32_t*)(maps + q.map_off))[i] = (int32_t)slot;
        }
    }
}

/* Occupied entries per query, over tiles of map entries. */
__global__ __launch_bounds__(256) void k_bq_count(const BQuery* Q, const BTile* tiles,
                                                  int64_t n_tiles, const uint8_t* maps,
                                                  unsigned long long* occupied) {
    __shared__ unsigned long long ws[256 / WAVE];
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        unsigned long long c = 0;
        for (int64_t i = W.r0 + threadIdx.x; i < W.r1; i += blockDim.x)
            c += q.map_u16 ? ((const uint16_t*)(maps + q.map_off))[i] != 0xffff
                           : ((const int32_t*)(maps + q.map_off))[i] != -1;
        for (int d = WAVE / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, WAVE);
        if (lane_id() == 0) ws[threadIdx.x / WAVE] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long s = 0;
            for (int w = 0; w < 256 / WAVE; w++) s += ws[w];
            if (s) atomicAdd(&occupied[W.q], s);
        }
        __syncthreads();
    }
}

/* Identity-initialise every query's group table. */
__global__ __launch_bounds__(256) void k_bq_init(const BQuery* Q, const BTile* tiles,
                                                 int64_t n_tiles, uint64_t* tabs) {
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        for (int64_t i = W.r0 + threadIdx.x; i < W.r1; i += blockDim.x)
            tabs[q.tab_off + i] = cell_identity(q.P.kind[i % q.P.ncells]);
    }
}

/* --------------------------------------------------------------- probe -- */

__device__ __forceinline__ uint32_t bq_pred(const BQuery& q, const int64_t (&cv)[TL_NC][BQ_R]) {
    uint32_t m = (1u << BQ_R) - 1;
    if (q.SP.npred >= 1) m &= range_mask<BQ_R>(q.SP, 0, q.T.col[1].type == TD_F64, cv[1]);
    if (q.SP.npred >= 2) m &= range_mask<BQ_R>(q.SP, 1, q.T.col[2].type == TD_F64, cv[2]);
    return m;
}

/* Fold a cell table `src` (LDS) into the query's global table. */
__device__ __forceinline__ void bq_flush(const GroupPlan& P, const uint64_t* src, uint64_t* dst,
                                         int64_t n_slots) {
    for (int64_t s = threadIdx.x; s < n_slots; s += blockDim.x) {
        const uint64_t* c = src + s * P.ncells;
        if (c[0] == 0) continue;
        cells_merge(dst + s * P.ncells, c, P);
    }
}

/* The star predicate and the slot of each of a lane's rows (-1: none). */
__device__ __forceinline__ void bq_slots(const BQuery& q, const uint8_t* maps, int64_t base,
                                         int64_t r1, int64_t (&cv)[TL_NC][BQ_R],
                                         int32_t (&s)[BQ_R]) {
    load_tile<BQ_R>(q.T, base, r1, cv);
    const uint32_t m = tile_live<BQ_R>(nullptr, base, r1) & bq_pred(q, cv);
    const uint16_t* m16 = (const uint16_t*)(maps + q.map_off);
    const int32_t* m32 = (const int32_t*)(maps + q.map_off);
#pragma unroll
    for (int j = 0; j < BQ_R; j++) {
        s[j] = -1;
        const uint64_t i = (uint64_t)(cv[0][j] - q.kmin);
        if (((m >> j) & 1) && i < (uint64_t)q.range) {
            if (q.map_u16) {
                uint16_t v = m16[i];
                s[j] = v == 0xffff ? -1 : (int32_t)v;
            } else {
                s[j] = m32[i];
            }
        }
    }
}

template <int J>
__device__ __forceinline__ uint64_t bq_word(const int64_t (&cv)[TL_NC][BQ_R], int slot) {
    switch (slot) {
    case 1: return (uint64_t)cv[1][J];
    case 2: return (uint64_t)cv[2][J];
    case 3: return (uint64_t)cv[3][J];
    default: return (uint64_t)cv[0][J];
    }
}

template <int J>
__device__ __forceinline__ void bq_put(const BQuery& q, const BRec& R, uint32_t* h,
                                       const int64_t (&cv)[TL_NC][BQ_R], const int32_t (&s)[BQ_R],
                                       int64_t& at) {
    if (s[J] < 0) return;
    R.slot[at] = (uint32_t)s[J];
    for (int w = 0; w < q.nrw; w++) R.val[w][at] = bq_word<J>(cv, q.rw_slot[w]);
    atomicAdd(&h[s[J] / q.sg], 1u);
    at++;
}

/* A wave's passing rows as records: one reservation per wave step, the
 * slice counts of the tile in LDS (h).  Every lane of the wave calls it. */
__device__ __forceinline__ void bq_records(const BQuery& q, const BRec& R,
                                           unsigned long long* cur, uint32_t* h,
                                           const int64_t (&cv)[TL_NC][BQ_R],
                                           const int32_t (&s)[BQ_R], int lane) {
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < BQ_R; j++) c += s[j] >= 0;
    uint32_t inc = c;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
    }
    const uint32_t tot = __shfl(inc, WAVE - 1);
    if (!tot) return;
    unsigned long long b = 0;
    if (lane == WAVE - 1) b = atomicAdd(cur, (unsigned long long)tot);
    b = __shfl(b, WAVE - 1);
    int64_t at = q.rec_off + (int64_t)b + (int64_t)(inc - c);
    bq_put<0>(q, R, h, cv, s, at);
    bq_put<1>(q, R, h, cv, s, at);
    bq_put<2>(q, R, h, cv, s, at);
    bq_put<3>(q, R, h, cv, s, at);
}

/* Probe tiles of bucketed queries: passing rows to records, slice counts. */
__global__ __launch_bounds__(512) void k_bq_precord(const BQuery* Q, const BTile* tiles,
                                                    int64_t n_tiles, const uint8_t* maps, BRec rec,
                                                    unsigned long long* qcur, uint32_t* sl_cnt) {
    extern __shared__ uint32_t h[];
    const int lane = lane_id();
    const int nwaves = blockDim.x / WAVE, wv = threadIdx.x / WAVE;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        for (int i = threadIdx.x; i < q.nsl; i += blockDim.x) h[i] = 0;
        __syncthreads();
        for (int64_t wb = W.r0 + (int64_t)wv * WAVE * BQ_R; wb < W.r1;
             wb += (int64_t)nwaves * WAVE * BQ_R) {            /* the same trips for a wave */
            int64_t cv[TL_NC][BQ_R];
            int32_t s[BQ_R];
            bq_slots(q, maps, wb + (int64_t)lane * BQ_R, W.r1, cv, s);
            bq_records(q, rec, qcur + W.q, h, cv, s, lane);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < q.nsl; i += blockDim.x)
            if (h[i]) atomicAdd(&sl_cnt[q.sl_off + i], h[i]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(512) void k_bq_probe(const BQuery* Q, const BTile* tiles,
                                                  int64_t n_tiles, const uint8_t* maps,
                                                  uint64_t* tabs) {
    extern __shared__ uint64_t lt[];
    const int lane = lane_id();
    const int nwaves = blockDim.x / WAVE, wv = threadIdx.x / WAVE;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        const GroupPlan& P = q.P;
        uint64_t* gtab = tabs + q.tab_off;
        const int64_t ncell_total = q.n_slots * P.ncells;
        uint64_t* tab = q.lds ? lt : gtab;
        if (q.lds) {
            for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x)
                lt[i] = cell_identity(P.kind[i % P.ncells]);
            __syncthreads();
        }
        const uint16_t* m16 = (const uint16_t*)(maps + q.map_off);
        const int32_t* m32 = (const int32_t*)(maps + q.map_off);
        for (int64_t base = W.r0 + ((int64_t)wv * WAVE + lane) * BQ_R; base < W.r1;
             base += (int64_t)nwaves * WAVE * BQ_R) {
            int64_t cv[TL_NC][BQ_R];
            load_tile<BQ_R>(q.T, base, W.r1, cv);
            const uint32_t m = tile_live<BQ_R>(nullptr, base, W.r1) & bq_pred(q, cv);
            int32_t s[BQ_R];
#pragma unroll
            for (int j = 0; j < BQ_R; j++) {
                s[j] = -1;
                const uint64_t i = (uint64_t)(cv[0][j] - q.kmin);
                if (((m >> j) & 1) && i < (uint64_t)q.range) {
                    if (q.map_u16) {
                        uint16_t v = m16[i];
                        s[j] = v == 0xffff ? -1 : (int32_t)v;
                    } else {
                        s[j] = m32[i];
                    }
                }
            }
            acc_tile<BQ_R>(tab, P, q.cell_slot, q.T, cv, s);
        }
        if (q.lds) {
            __syncthreads();
            bq_flush(P, lt, gtab, q.n_slots);
            __syncthreads();
        }
    }
}

/* -------------------------------------------------------------- bucket -- */

/* Slice starts of each bucketed query (record positions), and cursors. */
__global__ void k_bq_bscan(const BQuery* Q, const uint32_t* cnt, int64_t* pos,
                           unsigned long long* cur) {
    const BQuery& q = Q[blockIdx.x];
    if (!q.active || !q.nsl || threadIdx.x) return;
    int64_t a = q.rec_off;
    for (int i = 0; i < q.nsl; i++) {
        pos[q.sl_off + i] = a;
        cur[q.sl_off + i] = (unsigned long long)a;
        a += cnt[q.sl_off + i];
    }
    pos[q.sl_off + q.nsl] = a;
}

/* Records of a tile into slice order: ranks from LDS counters, one
 * reservation per (tile, slice). */
__global__ __launch_bounds__(BQ_SCT) void k_bq_bscatter(const BQuery* Q, const BTile* tiles,
                                                        int64_t n_tiles,
                                                        const unsigned long long* qcur, BRec A,
                                                        BRec B, unsigned long long* cur) {
    extern __shared__ uint32_t sh[];
    constexpr int K = BQ_SC / BQ_SCT;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        const int64_t nrec = (int64_t)qcur[W.q];
        const int64_t r1 = W.r1 < nrec ? W.r1 : nrec;
        if (W.r0 >= r1) continue;                 /* uniform */
        uint32_t* cnt = sh;
        unsigned long long* base = (unsigned long long*)(sh + ((q.nsl + 1) & ~1));
        for (int i = threadIdx.x; i < q.nsl; i += BQ_SCT) cnt[i] = 0;
        __syncthreads();
        int32_t sl[K];
        uint32_t rk[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int64_t r = W.r0 + threadIdx.x + (int64_t)k * BQ_SCT;
            sl[k] = r < r1 ? (int32_t)(A.slot[q.rec_off + r] / (uint32_t)q.sg) : -1;
            rk[k] = sl[k] >= 0 ? atomicAdd(&cnt[sl[k]], 1u) : 0u;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < q.nsl; i += BQ_SCT)
            if (cnt[i]) base[i] = atomicAdd(&cur[q.sl_off + i], (unsigned long long)cnt[i]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (sl[k] < 0) continue;
            const int64_t from = q.rec_off + W.r0 + threadIdx.x + (int64_t)k * BQ_SCT;
            const int64_t to = (int64_t)base[sl[k]] + rk[k];
            B.slot[to] = A.slot[from];
            for (int w = 0; w < q.nrw; w++) B.val[w][to] = A.val[w][from];
        }
        __syncthreads();
    }
}

/* Cell k of U records (bits of `live`) at LDS rows loc[u], the input from
 * record word W: one dispatch on the cell's kind for all U records. */
template <int U, int W>
__device__ __forceinline__ void bq_fold_cells(uint64_t* lt, int nc, int k, int kind, bool f,
                                              const uint32_t (&loc)[U], uint32_t live,
                                              const uint64_t (&v)[TL_NC - 1][U]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (!((live >> u) & 1)) continue;
        uint64_t* c = &lt[(int64_t)loc[u] * nc + k];
        const uint64_t x = v[W][u];
        const double d = f ? __longlong_as_double((long long)x) : (double)(int64_t)x;
        switch (kind) {                           /* uniform */
        case C_ADD_F64: atomicAdd((double*)c, d); break;
        case C_SQ_F64: atomicAdd((double*)c, d * d); break;
        case C_ADD_I64: atomicAdd((unsigned long long*)c, (unsigned long long)x); break;
        case C_MIN_I64: cell_min(c, (int64_t)x); break;
        case C_MAX_I64: cell_max(c, (int64_t)x); break;
        case C_MIN_F64: cell_min(c, f64_ord(d)); break;
        case C_MAX_F64: cell_max(c, f64_ord(d)); break;
        default: break;
        }
    }
}

#define BQ_FU 4                             /* records per thread per fold step */

/* One workgroup per (query, slice): the slice's records folded in LDS, then
 * the slice of the table stored (this workgroup alone writes it). */
__global__ __launch_bounds__(512) void k_bq_bfold(const BQuery* Q, const BTile* tiles,
                                                  int64_t n_tiles, const int64_t* pos, BRec B,
                                                  uint64_t* tabs) {
    extern __shared__ uint64_t lt[];
    __shared__ int8_t ckind[GP_MAX_CELLS], cword[GP_MAX_CELLS], cf64[GP_MAX_CELLS];
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        const GroupPlan& P = q.P;
        const int nc = P.ncells, nrw = q.nrw;
        if (threadIdx.x < GP_MAX_CELLS && (int)threadIdx.x < nc) {
            const int k = threadIdx.x;
            ckind[k] = (int8_t)P.kind[k];
            cword[k] = (int8_t)q.cell_rw[k];
            cf64[k] = q.T.col[q.rw_slot[q.cell_rw[k]]].type == TD_F64;
        }
        const int64_t s0 = W.r0 * q.sg;
        const int64_t ns = s0 + q.sg <= q.n_slots ? q.sg : q.n_slots - s0;
        for (int64_t i = threadIdx.x; i < ns * nc; i += blockDim.x)
            lt[i] = cell_identity(P.kind[i % nc]);
        __syncthreads();
        const int64_t beg = pos[q.sl_off + W.r0], end = pos[q.sl_off + W.r0 + 1];
        for (int64_t r0 = beg + threadIdx.x; r0 < end; r0 += (int64_t)blockDim.x * BQ_FU) {
            uint32_t loc[BQ_FU];
            uint64_t v[TL_NC - 1][BQ_FU];
            uint32_t live = 0;
#pragma unroll
            for (int u = 0; u < BQ_FU; u++) {      /* every load of the step first */
                const int64_t r = r0 + (int64_t)u * blockDim.x;
                const bool in = r < end;
                loc[u] = in ? (uint32_t)((int64_t)B.slot[r] - s0) : 0u;
#pragma unroll
                for (int w = 0; w < TL_NC - 1; w++) v[w][u] = in && w < nrw ? B.val[w][r] : 0ull;
                live |= (uint32_t)in << u;
            }
#pragma unroll
            for (int u = 0; u < BQ_FU; u++)
                if ((live >> u) & 1) atomicAdd((unsigned long long*)&lt[(int64_t)loc[u] * nc], 1ull);
            for (int k = 1; k < nc; k++) {
                switch (cword[k]) {                       /* compile-time record word */
                case 0: bq_fold_cells<BQ_FU, 0>(lt, nc, k, ckind[k], cf64[k], loc, live, v); break;
                case 1: bq_fold_cells<BQ_FU, 1>(lt, nc, k, ckind[k], cf64[k], loc, live, v); break;
                default: bq_fold_cells<BQ_FU, 2>(lt, nc, k, ckind[k], cf64[k], loc, live, v); break;
                }
            }
        }
        __syncthreads();
        uint64_t* dst = tabs + q.tab_off + s0 * nc;
        for (int64_t i = threadIdx.x; i < ns * nc; i += blockDim.x) dst[i] = lt[i];
        __syncthreads();
    }
}

/* ---------------------------------------------------------------- hash -- */
/* Queries whose build keys are not a dense unique range (scattered 64-bit
 * keys, repeated keys) join through radix partitions of both sides, as the
 * reference's radix join does (exec.c:7431-7950), every query of the batch
 * in the same launches:
 *
 *   hist     rows per (query, partition) of each side: the top `hbits` bits
 *            of the key's mix; probe rows only when they pass the predicate;
 *   scan     per (query, side) exclusive sums -> record positions;
 *   scatter  build rows -> (key, group slot) records, passing probe rows ->
 *            (key, aggregate inputs) records, each in partition order;
 *   join     one workgroup per (query, partition): the build records go into
 *            an LDS hash table (linear probing, a repeated key takes a slot
 *            of its own), each probe record walks its cluster and folds into
 *            the group of every build row with its key -- LDS group table
 *            when the query's table fits, flushed with one atomic per touched
 *            cell, else global atomics.  A build partition larger than one
 *            table is joined in table-sized chunks, the probe records
 *            streamed once per chunk, so every (probe, build) pair is met
 *            exactly once.
 *
 * The radix bits of a query follow from a target of build rows per
 * partition that the autotuner picks per batch size class. */
#define BH_MAXB  10                 /* radix bits of a hash query */
#define BH_CAP   2048               /* build rows per LDS table */
#define BH_TS    4096               /* LDS hash slots (load <= 0.5) */
#define BH_ST    2048               /* rows per partition-kernel tile */
#define BH_T     256                /* threads of the partition kernels */
#define BH_JT    512                /* threads of the join kernel */
#define BH_LDS_TAB (40 * 1024)      /* group table bytes kept in LDS by the join */
#define BH_EMPTY 0xffffffffu

struct BHRec {
    uint64_t* bkey;                 /* build records: key, group slot */
    int32_t*  bslot;
    uint64_t* pkey;                 /* probe records: key, aggregate input words */
    uint64_t* pval[TL_NC - 1];
};

__device__ __forceinline__ uint32_t bh_part(uint64_t k, int bits) {
    return bits ? (uint32_t)(jmix(k) >> (64 - bits)) : 0u;
}

__device__ __forceinline__ uint32_t bh_slot(uint64_t k) {
    return (uint32_t)jmix(k) & (BH_TS - 1);
}

/* Probe row r of q passes the star predicate. */
__device__ __forceinline__ bool bh_pass(const BQuery& q, int64_t r) {
    for (int i = 0; i < q.SP.npred; i++) {
        const int64_t v[1] = {ld_raw(q.T.col[1 + i], r)};
        if (!(range_mask<1>(q.SP, i, q.T.col[1 + i].type == TD_F64, v) & 1u)) return false;
    }
    return true;
}

/* Row r's key (side 0: build, 1: probe) and whether it takes part. */
__device__ __forceinline__ bool bh_row(const BQuery& q, int side, int64_t r, uint64_t& k) {
    if (side == 0) {
        k = (uint64_t)ld_i64(q.rk, r);
        return true;
    }
    k = (uint64_t)ld_i64(q.T.col[0], r);
    return bh_pass(q, r);
}

__global__ __launch_bounds__(BH_T) void k_bh_hist(const BQuery* Q, const BTile* tiles,
                                                  int64_t n_tiles, int side, uint32_t* cnt) {
    __shared__ uint32_t h[1 << BH_MAXB];
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        const int nb = 1 << q.hbits;
        for (int i = threadIdx.x; i < nb; i += BH_T) h[i] = 0;
        __syncthreads();
        for (int64_t r = W.r0 + threadIdx.x; r < W.r1; r += BH_T) {
            uint64_t k;
            if (bh_row(q, side, r, k)) atomicAdd(&h[bh_part(k, q.hbits)], 1u);
        }
        __syncthreads();
        const int64_t off = side ? q.hp_off : q.hb_off;
        for (int i = threadIdx.x; i < nb; i += BH_T)
            if (h[i]) atomicAdd(&cnt[off + i], h[i]);
        __syncthreads();
    }
}

/* Block (query, side): positions and cursors of its partitions. */
__global__ __launch_bounds__(256) void k_bh_scan(const BQuery* Q, const uint32_t* cnt,
                                                 int64_t* pos, unsigned long long* cur) {
    __shared__ int64_t ws[256 / WAVE];
    const BQuery& q = Q[blockIdx.x >> 1];
    const int side = blockIdx.x & 1;
    if (!q.active || !q.hash) return;
    const int nb = 1 << q.hbits;                 /* <= 1024: four entries a thread */
    const int64_t off = side ? q.hp_off : q.hb_off;
    const int64_t base = side ? q.hpr_off : q.hbr_off;
    const int i0 = threadIdx.x * 4;
    uint32_t c[4];
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        c[j] = i0 + j < nb ? cnt[off + i0 + j] : 0u;
        s += c[j];
    }
    int64_t incl = s;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t y = __shfl_up(incl, d, WAVE);
        if (lane_id() >= d) incl += y;
    }
    if (lane_id() == WAVE - 1) ws[threadIdx.x / WAVE] = incl;
    __syncthreads();
    int64_t a = base + incl - s;
    for (int w = 0; w < (int)threadIdx.x / WAVE; w++) a += ws[w];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (i0 + j < nb) {
            pos[off + i0 + j] = a;
            cur[off + i0 + j] = (unsigned long long)a;
        }
        a += c[j];
    }
    if (threadIdx.x == 255) pos[off + nb] = a;   /* base + every count */
}

/* Rows of a tile to their partitions' records: LDS ranks, one reservation
 * per (tile, partition). */
/* Exclusive scan of a[0, n) in place by the NT threads of a block; returns
 * the total to every thread. */
template <int NT>
__device__ __forceinline__ uint32_t bh_block_scan(uint32_t* a, int n, uint32_t* wsum) {
    const int per = (n + NT - 1) / NT;
    const int i0 = threadIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
    uint32_t t = 0;
    for (int i = i0; i < i1; i++) t += a[i];
    const int lane = lane_id(), w = threadIdx.x / WAVE;
    uint32_t inc = t;
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
    }
    if (lane == WAVE - 1) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < NT / WAVE; k++) {
        if (k < w) before += wsum[k];
        total += wsum[k];
    }
    uint32_t run = before + inc - t;
    for (int i = i0; i < i1; i++) {
        const uint32_t c = a[i];
        a[i] = run;
        run += c;
    }
    __syncthreads();
    return total;
}

/* Records of one side in partition order.  Per tile of BH_ST rows the
 * records are counting-sorted by partition in LDS, each partition's part
 * of its run is reserved with one atomic, and the records go out one array
 * at a time through an LDS stage, so consecutive lanes write consecutive
 * addresses of a partition's run. */
__global__ __launch_bounds__(BH_T) void k_bh_scatter(const BQuery* Q, const BTile* tiles,
                                                     int64_t n_tiles, int side,
                                                     unsigned long long* cur, BHRec R) {
    __shared__ uint32_t h[1 << BH_MAXB];             /* counts, then tile offsets */
    __shared__ unsigned long long hb[1 << BH_MAXB];
    __shared__ uint16_t spart[BH_ST];
    __shared__ uint64_t stage[BH_ST];
    __shared__ uint32_t wsum[BH_T / WAVE];
    constexpr int K = BH_ST / BH_T;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BTile W = tiles[t];
        const BQuery& q = Q[W.q];
        const int nb = 1 << q.hbits;
        const int64_t off = side ? q.hp_off : q.hb_off;
        for (int i = threadIdx.x; i < nb; i += BH_T) h[i] = 0;
        __syncthreads();
        uint64_t key[K];
        uint32_t part[K], rank[K], live = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int64_t r = W.r0 + threadIdx.x + (int64_t)j * BH_T;
            part[j] = rank[j] = 0;
            key[j] = 0;
            if (r < W.r1 && bh_row(q, side, r, key[j])) {
                live |= 1u << j;
                part[j] = bh_part(key[j], q.hbits);
                rank[j] = atomicAdd(&h[part[j]], 1u);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nb; i += BH_T)
            hb[i] = h[i] ? atomicAdd(&cur[off + i], (unsigned long long)h[i]) : 0ull;
        __syncthreads();
        const int n = (int)bh_block_scan<BH_T>(h, nb, wsum);
#pragma unroll
        for (int j = 0; j < K; j++)
            if ((live >> j) & 1) {
                rank[j] += h[part[j]];                      /* sorted place in the tile */
                spart[rank[j]] = (uint16_t)part[j];
                stage[rank[j]] = key[j];
            }
        __syncthreads();
        uint64_t* okey = side ? R.pkey : R.bkey;
        for (int i = threadIdx.x; i < n; i += BH_T) {
            const int p = spart[i];
            okey[hb[p] + (uint32_t)i - h[p]] = stage[i];
        }
        const int nw = side ? q.nrw : 1;
        for (int w = 0; w < nw; w++) {
            uint64_t v[K];
#pragma unroll
            for (int j = 0; j < K; j++) {                   /* the word's loads first */
                v[j] = 0;
                if (!((live >> j) & 1)) continue;
                const int64_t r = W.r0 + threadIdx.x + (int64_t)j * BH_T;
                if (side == 0) {
                    int64_t slot = 0;
                    for (int k = 0; k < q.P.n_keys; k++)
                        slot += (ld_i64(q.P.keys[k], r) - q.P.kmin[k]) * q.P.kstride[k];
                    v[j] = (uint64_t)slot;
                } else {
                    v[j] = (uint64_t)ld_raw(q.T.col[q.rw_slot[w]], r);
                }
            }
            __syncthreads();                                /* the previous array is out */
#pragma unroll
            for (int j = 0; j < K; j++)
                if ((live >> j) & 1) stage[rank[j]] = v[j];
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += BH_T) {
                const int p = spart[i];
                const int64_t at = (int64_t)hb[p] + (uint32_t)i - h[p];
                if (side == 0) R.bslot[at] = (int32_t)stage[i];
                else R.pval[w][at] = stage[i];
            }
        }
        __syncthreads();
    }
}

/* Every cell of group slot gs for one probe record (input words v0..v2). */
__device__ __forceinline__ void bh_fold(uint64_t* tab, int nc, const int8_t* ckind,
                                        const int8_t* cword, const int8_t* cf64, int32_t gs,
                                        uint64_t v0, uint64_t v1, uint64_t v2) {
    uint64_t* c = tab + (int64_t)gs * nc;
    atomicAdd((unsigned long long*)&c[0], 1ull);
    for (int k = 1; k < nc; k++) {
        const int w = cword[k];
        const uint64_t x = w == 0 ? v0 : (w == 1 ? v1 : v2);
        const double d = cf64[k] ? __longlong_as_double((long long)x) : (double)(int64_t)x;
        switch (ckind[k]) {
        case C_ADD_F64: atomicAdd((double*)&c[k], d); break;
        case C_SQ_F64: atomicAdd((double*)&c[k], d * d); break;
        case C_ADD_I64: atomicAdd((unsigned long long*)&c[k], (unsigned long long)x); break;
        case C_MIN_I64: cell_min(&c[k], (int64_t)x); break;
        case C_MAX_I64: cell_max(&c[k], (int64_t)x); break;
        case C_MIN_F64: cell_min(&c[k], f64_ord(d)); break;
        case C_MAX_F64: cell_max(&c[k], f64_ord(d)); break;
        default: break;
        }
    }
}

#define BH_PU 4                     /* probe records per thread in flight */

__global__ __launch_bounds__(BH_JT) void k_bh_join(const BQuery* Q, const BTile* items,
                                                   int64_t n_items, const int64_t* pos, BHRec R,
                                                   uint64_t* tabs) {
    extern __shared__ uint64_t lt[];             /* the query's group table (q.lds) */
    __shared__ uint32_t hs[BH_TS];               /* slot -> build-local row */
    __shared__ uint64_t hk[BH_CAP];
    __shared__ int32_t hg[BH_CAP];
    __shared__ int8_t ckind[GP_MAX_CELLS], cword[GP_MAX_CELLS], cf64[GP_MAX_CELLS];
    for (int64_t t = blockIdx.x; t < n_items; t += gridDim.x) {
        const BTile W = items[t];
        const BQuery& q = Q[W.q];
        const GroupPlan& P = q.P;
        const int nc = P.ncells, nrw = q.nrw;
        uint64_t* gtab = tabs + q.tab_off;
        const bool in_lds = q.lds != 0;
        if ((int)threadIdx.x < nc) {
            const int k = threadIdx.x;
            ckind[k] = (int8_t)P.kind[k];
            cword[k] = (int8_t)q.cell_rw[k];
            cf64[k] = k > 0 && q.T.col[q.rw_slot[q.cell_rw[k]]].type == TD_F64;
        }
        const int64_t ncell_total = q.n_slots * nc;
        if (in_lds)
            for (int64_t i = threadIdx.x; i < ncell_total; i += BH_JT)
                lt[i] = cell_identity(P.kind[i % nc]);
        const int64_t bs = pos[q.hb_off + W.r0], be = pos[q.hb_off + W.r0 + 1];
        const int64_t ps = pos[q.hp_off + W.r0], pe = pos[q.hp_off + W.r0 + 1];
        for (int64_t c0 = bs; c0 < be; c0 += BH_CAP) {       /* uniform */
            const int cn = (int)(be - c0 < BH_CAP ? be - c0 : BH_CAP);
            __syncthreads();
            for (int i = threadIdx.x; i < BH_TS; i += BH_JT) hs[i] = BH_EMPTY;
            __syncthreads();
            for (int i = threadIdx.x; i < cn; i += BH_JT) {
                const uint64_t k = R.bkey[c0 + i];
                hk[i] = k;
                hg[i] = R.bslot[c0 + i];
                uint32_t s = bh_slot(k);
                while (atomicCAS(&hs[s], BH_EMPTY, (uint32_t)i) != BH_EMPTY) s = (s + 1) & (BH_TS - 1);
            }
            __syncthreads();
            for (int64_t i0 = ps + threadIdx.x; i0 < pe; i0 += (int64_t)BH_JT * BH_PU) {
                uint64_t k[BH_PU], v[TL_NC - 1][BH_PU];
                uint32_t s[BH_PU], live = 0;
#pragma unroll
                for (int u = 0; u < BH_PU; u++) {            /* every load of the step first */
                    const int64_t i = i0 + (int64_t)u * BH_JT;
                    const bool in = i < pe;
                    k[u] = in ? R.pkey[i] : 0ull;
#pragma unroll
                    for (int w = 0; w < TL_NC - 1; w++) v[w][u] = in && w < nrw ? R.pval[w][i] : 0ull;
                    live |= (uint32_t)in << u;
                }
#pragma unroll
                for (int u = 0; u < BH_PU; u++) s[u] = bh_slot(k[u]);
                while (live) {
#pragma unroll
                    for (int u = 0; u < BH_PU; u++) {
                        if (!((live >> u) & 1)) continue;
                        const uint32_t li = hs[s[u]];
                        if (li == BH_EMPTY) { live &= ~(1u << u); continue; }
                        if (hk[li] == k[u]) {
                            if (in_lds) bh_fold(lt, nc, ckind, cword, cf64, hg[li], v[0][u], v[1][u], v[2][u]);
                            else bh_fold(gtab, nc, ckind, cword, cf64, hg[li], v[0][u], v[1][u], v[2][u]);
                        }
                        s[u] = (s[u] + 1) & (BH_TS - 1);
                    }
                }
            }
        }
        __syncthreads();
        if (in_lds) {
            bq_flush(P, lt, gtab, q.n_slots);
            __syncthreads();
        }
    }
}

/* ---------------------------------------------------------------- emit -- */

/* One workgroup per query: non-empty slots (count cell > 0). */
__global__ __launch_bounds__(256) void k_bq_ngroups(const BQuery* Q, const uint64_t* tabs,
                                                    int64_t* ngroups) {
    __shared__ int64_t ws[256 / WAVE];
    const BQuery& q = Q[blockIdx.x];
    const uint64_t* tab = tabs + q.tab_off;
    int64_t c = 0;
    if (q.active)
        for (int64_t s = threadIdx.x; s < q.n_slots; s += 256) c += tab[s * q.P.ncells] != 0;
    for (int d = WAVE / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, WAVE);
    if (lane_id() == 0) ws[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0;
        for (int w = 0; w < 256 / WAVE; w++) s += ws[w];
        ngroups[blockIdx.x] = s;
    }
}

/* One workgroup per query: groups in ascending slot (= key) order at the
 * query's output offset. */
__global__ __launch_bounds__(256) void k_bq_emit(const BQuery* Q, const uint64_t* tabs,
                                                 const int64_t* offs, tdd_group_result_t out) {
    __shared__ int64_t ws[256 / WAVE];
    __shared__ int64_t carry;
    const BQuery& q = Q[blockIdx.x];
    if (!q.active) return;
    const GroupPlan& P = q.P;
    const uint64_t* tab = tabs + q.tab_off;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t s0 = 0; s0 < q.n_slots; s0 += 256) {
        const int64_t s = s0 + threadIdx.x;
        const bool live = s < q.n_slots && tab[s * P.ncells] != 0;
        const uint64_t bal = __ballot(live);
        const int64_t below = lanes_below(bal);
        if (lane_id() == 0) ws[threadIdx.x / WAVE] = __popcll(bal);
        __syncthreads();
        int64_t off = carry + below;
        for (int w = 0; w < (int)threadIdx.x / WAVE; w++) off += ws[w];
        if (live) {
            const int64_t g = offs[blockIdx.x] + off;
            const uint64_t* c = tab + s * P.ncells;
            for (int k = 0; k < P.n_keys; k++) {
                const int64_t v = P.kmin[k] + (s / P.kstride[k]) % P.krange[k];
                st_int(out.keys[k], g, P.keys[k].esz, v);
            }
            for (int a = 0; a < P.n_aggs; a++) cells_emit(c, P, a, out.vals[a], g);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t tot = 0;
            for (int w = 0; w < 256 / WAVE; w++) tot += ws[w];
            carry += tot;
        }
        __syncthreads();
    }
}

/* -------------------------------------------------------------- driver -- */

struct tdd_jg_batch {
    int n;
    std::vector<BQuery> hq;
    std::vector<int> status;                /* per query: 0 batched, else TD_ERR_NYI */
    void* dq = nullptr;                     /* BQuery[n] */
    void* maps = nullptr;
    void* tabs = nullptr;
    void* occ = nullptr;                    /* u64[n] */
    void* ng = nullptr;                     /* i64[n] */
    void* offs = nullptr;                   /* i64[n + 1] */
    void* tl_build = nullptr; int64_t n_build = 0;
    void* tl_map = nullptr;   int64_t n_map = 0;
    void* tl_tab = nullptr;   int64_t n_tab = 0;
    void* tl_probe = nullptr; int64_t n_probe = 0;
    void* tl_prec = nullptr;  int64_t n_prec = 0;     /* probe tiles of bucketed queries */
    void* tl_bsc = nullptr;   int64_t n_bsc = 0;      /* bucket scatter: (query, records) */
    void* tl_bfold = nullptr; int64_t n_bfold = 0;    /* bucket fold: (query, slice) */
    void* rec_mem = nullptr;                          /* the two record buffers */
    BRec recA, recB;
    void* qcur = nullptr;                             /* u64[n]: records per query */
    void* sl_cnt = nullptr;                           /* u32: records per (query, slice) */
    void* sl_pos = nullptr;                           /* i64: slice starts (+ end) */
    void* sl_cur = nullptr;                           /* u64: slice cursors */
    int64_t sl_total = 0;
    int nsl_max = 0;
    size_t lds_fold = 0;
    size_t map_bytes = 0, tab_cells = 0;
    int grid_probe = 0, grid_prec = 0, block_probe = 512;
    size_t lds_probe = 0;
    int64_t tile_rows = 16384;
    /* hash queries (build keys not a dense unique range) */
    int n_hash = 0;
    void* tl_hb = nullptr;    int64_t n_hb = 0;       /* build tiles */
    void* tl_hp = nullptr;    int64_t n_hp = 0;       /* probe tiles */
    void* tl_hj = nullptr;    int64_t n_hj = 0;       /* join items: (query, partition) */
    void* h_mem = nullptr;                            /* build and probe records */
    BHRec hrec;
    void* h_cnt = nullptr;                            /* u32 rows per (query, side, partition) */
    void* h_pos = nullptr;                            /* i64 record positions (+ end) */
    void* h_cur = nullptr;                            /* u64 cursors */
    int64_t h_ent = 0;
    int64_t h_target = 0;                             /* build rows per partition */
    size_t lds_join = 0;
    int grid_join = 0;
    GroupPlan shape;                        /* output types of the shape */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void bq_release(tdd_jg_batch* b) {
    void* ps[] = {b->dq, b->maps, b->tabs, b->occ, b->ng, b->offs, b->tl_build, b->tl_map,
                  b->tl_tab, b->tl_probe, b->tl_prec, b->tl_bsc, b->tl_bfold, b->rec_mem, b->qcur,
                  b->sl_cnt, b->sl_pos, b->sl_cur, b->tl_hb, b->tl_hp, b->tl_hj, b->h_mem,
                  b->h_cnt, b->h_pos, b->h_cur};
    for (void* p : ps) if (p) tdd::release(p);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
}

static int upload(void** dst, const void* src, size_t bytes) {
    *dst = tdd::scratch(bytes ? bytes : 8);
    if (!*dst) return TD_ERR_OOM;
    if (bytes) HC(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, tdd::g_stream));
    return TD_OK;
}

/* Tiles of `chunk` rows over each query's range [0, n(q)). */
template <class F>
static std::vector<BTile> make_tiles(int n, int64_t chunk, F count) {
    std::vector<BTile> v;
    for (int q = 0; q < n; q++) {
        const int64_t m = count(q);
        for (int64_t r = 0; r < m; r += chunk) v.push_back(BTile{q, 0, r, r + chunk < m ? r + chunk : m});
    }
    return v;
}

/* Launch sequence of one run (capturable: kernels only). */
static int bq_launch(tdd_jg_batch* b) {
    int rc = tdd_fill(b->maps, ~0ull, b->map_bytes);
    if (rc) return rc;
    const BQuery* Q = (const BQuery*)b->dq;
    if (b->n_build)
        LAUNCH("batch_map", k_bq_scatter, stream_grid(b->n_build, 8), 256, 0, Q,
               (const BTile*)b->tl_build, b->n_build, (uint8_t*)b->maps);
    if (b->n_tab)
        LAUNCH("batch_init", k_bq_init, stream_grid(b->n_tab, 8), 256, 0, Q, (const BTile*)b->tl_tab,
               b->n_tab, (uint64_t*)b->tabs);
    if (b->n_bfold) {
        rc = tdd_fill(b->qcur, 0, (size_t)b->n * 8);
        if (!rc) rc = tdd_fill(b->sl_cnt, 0, (size_t)b->sl_total * 4);
        if (rc) return rc;
    }
    if (b->n_probe)
        LAUNCH("batch_probe", k_bq_probe, b->grid_probe, b->block_probe, b->lds_probe, Q,
               (const BTile*)b->tl_probe, b->n_probe, (const uint8_t*)b->maps, (uint64_t*)b->tabs);
    if (b->n_prec)
        LAUNCH("batch_records", k_bq_precord, b->grid_prec, b->block_probe,
               (size_t)b->nsl_max * 4, Q, (const BTile*)b->tl_prec, b->n_prec,
               (const uint8_t*)b->maps, b->recA, (unsigned long long*)b->qcur,
               (uint32_t*)b->sl_cnt);
    if (b->n_bfold) {
        LAUNCH("batch_bucket", k_bq_bscan, b->n, 64, 0, Q, (const uint32_t*)b->sl_cnt,
               (int64_t*)b->sl_pos, (unsigned long long*)b->sl_cur);
        if (b->n_bsc)
            LAUNCH("batch_bucket", k_bq_bscatter, stream_grid(b->n_bsc, 8), BQ_SCT,
                   (size_t)((b->nsl_max + 1) & ~1) * 4 + (size_t)b->nsl_max * 8, Q,
                   (const BTile*)b->tl_bsc, b->n_bsc, (const unsigned long long*)b->qcur, b->recA,
                   b->recB, (unsigned long long*)b->sl_cur);
        LAUNCH("batch_fold", k_bq_bfold, stream_grid(b->n_bfold, 4), 512, b->lds_fold, Q,
               (const BTile*)b->tl_bfold, b->n_bfold, (const int64_t*)b->sl_pos, b->recB,
               (uint64_t*)b->tabs);
    }
    if (b->n_hj) {
        uint32_t* cnt = (uint32_t*)b->h_cnt;
        rc = tdd_fill(b->h_cnt, 0, (size_t)b->h_ent * 4);
        if (rc) return rc;
        if (b->n_hb)
            LAUNCH("batch_hash_part", k_bh_hist, stream_grid(b->n_hb, 8), BH_T, 0, Q,
                   (const BTile*)b->tl_hb, b->n_hb, 0, cnt);
        if (b->n_hp)
            LAUNCH("batch_hash_part", k_bh_hist, stream_grid(b->n_hp, 8), BH_T, 0, Q,
                   (const BTile*)b->tl_hp, b->n_hp, 1, cnt);
        LAUNCH("batch_hash_part", k_bh_scan, 2 * b->n, 256, 0, Q, (const uint32_t*)cnt,
               (int64_t*)b->h_pos, (unsigned long long*)b->h_cur);
        if (b->n_hb)
            LAUNCH("batch_hash_part", k_bh_scatter, stream_grid(b->n_hb, 8), BH_T, 0, Q,
                   (const BTile*)b->tl_hb, b->n_hb, 0, (unsigned long long*)b->h_cur, b->hrec);
        if (b->n_hp)
            LAUNCH("batch_hash_part", k_bh_scatter, stream_grid(b->n_hp, 8), BH_T, 0, Q,
                   (const BTile*)b->tl_hp, b->n_hp, 1, (unsigned long long*)b->h_cur, b->hrec);
        LAUNCH("batch_hash_join", k_bh_join, b->grid_join, BH_JT, b->lds_join, Q,
               (const BTile*)b->tl_hj, b->n_hj, (const int64_t*)b->h_pos, b->hrec,
               (uint64_t*)b->tabs);
    }
    return TD_OK;
}

/* Record words of query q: the distinct tile slots its cells read; false
 * when they exceed the record's TL_NC - 1 words. */
static bool bq_record_words(BQuery& q) {
    const GroupPlan& P = q.P;
    q.nrw = 0;
    for (int c = 1; c < P.ncells; c++) {
        const int sl = q.cell_slot[c];
        int w = -1;
        for (int k = 0; k < q.nrw; k++) if (q.rw_slot[k] == sl) w = k;
        if (w < 0) {
            if (q.nrw >= TL_NC - 1) { q.nrw = 0; return false; }
            w = q.nrw++;
            q.rw_slot[w] = (uint8_t)sl;
        }
        q.cell_rw[c] = (uint8_t)w;
    }
    return true;
}

/* Radix plan of the hash queries for a target of build rows per partition:
 * each query's bits, partition entries, record regions, the work lists;
 * the query descriptors are uploaded again. */
static int bh_plan(tdd_jg_batch* b, int64_t target) {
    void* old[] = {b->tl_hb, b->tl_hp, b->tl_hj, b->h_mem, b->h_cnt, b->h_pos, b->h_cur};
    for (void* p : old) if (p) tdd::release(p);
    b->tl_hb = b->tl_hp = b->tl_hj = b->h_mem = b->h_cnt = b->h_pos = b->h_cur = nullptr;
    b->n_hb = b->n_hp = b->n_hj = 0;
    b->h_target = target;
    b->lds_join = 0;
    int64_t ent = 0, nbr = 0, npr = 0;
    int nrw_max = 0;
    std::vector<BTile> hb, hp, hj;
    for (int i = 0; i < b->n; i++) {
        BQuery& q = b->hq[i];
        if (b->status[i] || !q.hash) continue;
        int bits = 0;
        while (bits < BH_MAXB && (target << bits) < q.nr) bits++;
        q.hbits = bits;
        q.hb_off = ent;
        ent += ((int64_t)1 << bits) + 1;
        q.hp_off = ent;
        ent += ((int64_t)1 << bits) + 1;
        q.hbr_off = nbr;
        nbr += q.nr;
        q.hpr_off = npr;
        npr += q.nl;
        if (q.nrw > nrw_max) nrw_max = q.nrw;
        const size_t tb = (size_t)q.n_slots * q.P.ncells * 8;
        q.lds = tb <= BH_LDS_TAB;
        if (q.lds && tb > b->lds_join) b->lds_join = tb;
        for (int64_t r = 0; r < q.nr; r += BH_ST)
            hb.push_back(BTile{i, 0, r, r + BH_ST < q.nr ? r + BH_ST : q.nr});
        for (int64_t r = 0; r < q.nl; r += BH_ST)
            hp.push_back(BTile{i, 0, r, r + BH_ST < q.nl ? r + BH_ST : q.nl});
        for (int p = 0; p < (1 << bits); p++) hj.push_back(BTile{i, 0, p, p + 1});
    }
    b->h_ent = (ent + 1) & ~(int64_t)1;
    const size_t per_p = 8 + 8 * (size_t)nrw_max;
    b->h_mem = tdd::scratch((size_t)nbr * 12 + (size_t)npr * per_p + 2048);
    b->h_cnt = tdd::scratch((size_t)b->h_ent * 4);
    b->h_pos = tdd::scratch((size_t)b->h_ent * 8);
    b->h_cur = tdd::scratch((size_t)b->h_ent * 8);
    if (!b->h_mem || !b->h_cnt || !b->h_pos || !b->h_cur) return TD_ERR_OOM;
    uint8_t* m = (uint8_t*)b->h_mem;
    auto carve = [&](size_t bytes) { uint8_t* r = m; m += (bytes + 255) & ~(size_t)255; return r; };
    memset(&b->hrec, 0, sizeof(BHRec));
    b->hrec.bkey = (uint64_t*)carve((size_t)nbr * 8);
    b->hrec.bslot = (int32_t*)carve((size_t)nbr * 4);
    b->hrec.pkey = (uint64_t*)carve((size_t)npr * 8);
    for (int w = 0; w < nrw_max; w++) b->hrec.pval[w] = (uint64_t*)carve((size_t)npr * 8);
    b->n_hb = (int64_t)hb.size();
    b->n_hp = (int64_t)hp.size();
    b->n_hj = (int64_t)hj.size();
    int rc = upload(&b->tl_hb, hb.data(), hb.size() * sizeof(BTile));
    if (!rc) rc = upload(&b->tl_hp, hp.data(), hp.size() * sizeof(BTile));
    if (!rc) rc = upload(&b->tl_hj, hj.data(), hj.size() * sizeof(BTile));
    if (rc) return rc;
    HC(hipMemcpyAsync(b->dq, b->hq.data(), sizeof(BQuery) * b->n, hipMemcpyHostToDevice,
                      tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    b->grid_join = resident_grid(k_bh_join, BH_JT, b->lds_join, b->n_hj);
    return TD_OK;
}

/* Bucketed fold of the queries whose group tables exceed the per-tile LDS
 * budget: slices whose tables fit it, record buffers, work lists.  Queries
 * with more slices than BQ_MAX_SL keep the global-table fold. */
static int bq_bucket_plan(tdd_jg_batch* b) {
    int64_t rec_total = 0, sl_total = 0;
    int nrw_max = 0;
    b->nsl_max = 0;
    b->lds_fold = 0;
    for (int i = 0; i < b->n; i++) {
        BQuery& q = b->hq[i];
        q.nsl = q.sg = 0;
        const GroupPlan& P = q.P;
        const int64_t tbytes = q.n_slots * P.ncells * 8;
        if (b->status[i] || q.hash ||
            tbytes <= (int64_t)TUNE_INT("TEIDE_BQ_BUCKET_KB", 48) * 1024)
            continue;
        const int64_t sg = BQ_LDS_MAX / (8 * (int64_t)P.ncells);
        const int64_t nsl = ceil_div(q.n_slots, sg);
        if (sg < 1 || nsl > BQ_MAX_SL) continue;
        if (!bq_record_words(q)) continue;            /* record words: the tile slots the cells read */
        q.nsl = (int32_t)nsl;
        q.sg = (int32_t)sg;
        q.rec_off = rec_total;
        q.sl_off = sl_total;
        rec_total += q.nl;
        sl_total += nsl + 1;
        if (q.nrw > nrw_max) nrw_max = q.nrw;
        if ((int)nsl > b->nsl_max) b->nsl_max = (int)nsl;
        const size_t lf = (size_t)sg * P.ncells * 8;
        if (lf > b->lds_fold) b->lds_fold = lf;
    }
    b->sl_total = sl_total;
    if (!sl_total) return TD_OK;
    /* two record buffers: slot u32 + nrw_max words each */
    const size_t per = 4 + 8 * (size_t)nrw_max;
    b->rec_mem = tdd::scratch((size_t)rec_total * per * 2 + 1024);
    b->qcur = tdd::scratch((size_t)b->n * 8);
    b->sl_cnt = tdd::scratch((size_t)sl_total * 4);
    b->sl_pos = tdd::scratch((size_t)sl_total * 8);
    b->sl_cur = tdd::scratch((size_t)sl_total * 8);
    if (!b->rec_mem || !b->qcur || !b->sl_cnt || !b->sl_pos || !b->sl_cur) return TD_ERR_OOM;
    uint8_t* m = (uint8_t*)b->rec_mem;
    auto carve = [&](size_t bytes) { uint8_t* r = m; m += (bytes + 255) & ~(size_t)255; return r; };
    memset(&b->recA, 0, sizeof(BRec));
    memset(&b->recB, 0, sizeof(BRec));
    b->recA.slot = (uint32_t*)carve((size_t)rec_total * 4);
    b->recB.slot = (uint32_t*)carve((size_t)rec_total * 4);
    for (int w = 0; w < nrw_max; w++) {
        b->recA.val[w] = (uint64_t*)carve((size_t)rec_total * 8);
        b->recB.val[w] = (uint64_t*)carve((size_t)rec_total * 8);
    }
    std::vector<BTile> sc, fo;
    for (int i = 0; i < b->n; i++) {
        const BQuery& q = b->hq[i];
        if (!q.nsl) continue;
        for (int64_t r = 0; r < q.nl; r += BQ_SC)
            sc.push_back(BTile{i, 0, r, r + BQ_SC < q.nl ? r + BQ_SC : q.nl});
        for (int k = 0; k < q.nsl; k++) fo.push_back(BTile{i, 0, k, k + 1});
    }
    b->n_bsc = (int64_t)sc.size();
    b->n_bfold = (int64_t)fo.size();
    int rc = upload(&b->tl_bsc, sc.data(), sc.size() * sizeof(BTile));
    if (!rc) rc = upload(&b->tl_bfold, fo.data(), fo.size() * sizeof(BTile));
    return rc;
}

/* Grids of the two probe kernels for the current block size and tiles. */
static void bq_grids(tdd_jg_batch* b) {
    b->grid_probe = resident_grid(k_bq_probe, b->block_probe, b->lds_probe, b->n_probe);
    b->grid_prec = resident_grid(k_bq_precord, b->block_probe, (size_t)b->nsl_max * 4, b->n_prec);
}

/* Autotuned launch shape per size class (log2 of the probe rows, and
 * whether the batch has hash queries): probe tile rows, block size, and
 * the hash queries' build rows per partition (their radix bits). */
struct BQTuned {
    int64_t chunk;
    int     block;
    int64_t target;
};
static std::map<int, BQTuned> g_bq_tuned;

#define BH_TARGET0 ((int64_t)BH_CAP * 5 / 8)   /* default: 0.3 of an LDS table's slots */

static int bq_tiles_for(tdd_jg_batch* b, int64_t chunk) {
    if (b->tl_probe) { tdd::release(b->tl_probe); b->tl_probe = nullptr; }
    if (b->tl_prec) { tdd::release(b->tl_prec); b->tl_prec = nullptr; }
    std::vector<BTile> tp = make_tiles(b->n, chunk, [&](int q) {
        return b->status[q] || b->hq[q].nsl || b->hq[q].hash ? (int64_t)0 : b->hq[q].nl;
    });
    std::vector<BTile> tr = make_tiles(b->n, chunk, [&](int q) {
        return b->status[q] || !b->hq[q].nsl ? (int64_t)0 : b->hq[q].nl;
    });
    b->n_probe = (int64_t)tp.size();
    b->n_prec = (int64_t)tr.size();
    b->tile_rows = chunk;
    int rc = upload(&b->tl_probe, tp.data(), tp.size() * sizeof(BTile));
    if (!rc) rc = upload(&b->tl_prec, tr.data(), tr.size() * sizeof(BTile));
    return rc;
}

static int bq_autotune(tdd_jg_batch* b, int64_t total_rows) {
    int cls = 0;
    while ((int64_t)1 << (cls + 1) <= total_rows) cls++;
    cls = cls * 2 + (b->n_hash > 0);
    auto it = g_bq_tuned.find(cls);
    if (it != g_bq_tuned.end()) {
        b->block_probe = it->second.block;
        int rc = b->n_hash && it->second.target != b->h_target ? bh_plan(b, it->second.target)
                                                                : TD_OK;
        return rc ? rc : bq_tiles_for(b, it->second.chunk);
    }
    const int64_t chunks[] = {4096, 16384, 65536};
    const int blocks[] = {256, 512};
    /* radix bits of the hash queries: partitions of 0.3, 0.15 or 0.075 of a
     * table (fewer, larger builds against more, smaller ones) */
    const int64_t targets[] = {BH_TARGET0, BH_TARGET0 / 2, BH_TARGET0 / 4};
    const int n_targets = b->n_hash ? 3 : 1;
    float best = 1e30f;
    BQTuned bt = {16384, 512, b->h_target};
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    for (int ti = 0; ti < n_targets; ti++) {
        if (b->n_hash && b->h_target != targets[ti]) {
            int rc = bh_plan(b, targets[ti]);
            if (rc) return rc;
        }
        for (int64_t c : chunks) {
            int rc = bq_tiles_for(b, c);
            if (rc) return rc;
            for (int bs : blocks) {
                b->block_probe = bs;
                bq_grids(b);
                rc = bq_launch(b);                     /* warm + reset tables */
                if (rc) return rc;
                HC(hipEventRecord(e0, tdd::g_stream));
                rc = bq_launch(b);
                if (rc) return rc;
                HC(hipEventRecord(e1, tdd::g_stream));
                HC(hipEventSynchronize(e1));
                float ms = 0;
                HC(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; bt = {c, bs, b->h_target}; }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g_bq_tuned[cls] = bt;
    b->block_probe = bt.block;
    if (b->n_hash && bt.target != b->h_target) {
        int rc = bh_plan(b, bt.target);
        if (rc) return rc;
    }
    return bq_tiles_for(b, bt.chunk);
}

extern "C" int tdd_jg_batch_prepare(const tdd_join_group_spec_t* const* specs, int n,
                                    int tune, tdd_jg_batch_t** out) {
    *out = nullptr;
    if (n <= 0) return TD_ERR_RANGE;
    tdd_jg_batch* b = new tdd_jg_batch();
    b->n = n;
    b->hq.resize(n);
    b->status.assign(n, 0);
    size_t map_bytes = 0, tab_cells = 0;
    size_t lds_max = 0;
    int64_t total_rows = 0;
    int shape_q = -1;
    for (int i = 0; i < n; i++) {
        const tdd_join_group_spec_t* S = specs[i];
        const tdd_join_spec_t& J = S->join;
        BQuery& q = b->hq[i];
        memset(&q, 0, sizeof(q));
        int st = TD_OK;
        bool ok = J.join_type == 0 && J.n_keys == 1 && !J.lsel && J.lkeys[0].type != TD_F64 &&
                  J.rkeys[0].type != TD_F64 && S->n_keys >= 1 && S->n_aggs <= TDD_MAX_AGGS;
        GroupPlan& P = q.P;
        if (ok) {
            P.n_keys = S->n_keys;
            P.n_aggs = S->n_aggs;
            for (int k = 0; k < P.n_keys; k++) {
                P.keys[k] = S->keys[k];
                P.key_side[k] = S->key_side[k];
                ok = ok && S->key_side[k] == 1 && S->keys[k].ranged && S->keys[k].type != TD_F64;
            }
            for (int a = 0; a < P.n_aggs; a++) {
                P.aggs[a] = S->aggs[a];
                P.ops[a] = S->ops[a];
                P.agg_side[a] = S->agg_side[a];
                ok = ok && S->agg_side[a] == 0 && S->ops[a] != OP_FIRST && S->ops[a] != OP_LAST;
            }
        }
        if (ok) ok = plan_cells(P) == TD_OK;
        if (ok) {
            double slots = 1;
            for (int k = 0; k < P.n_keys; k++) {
                P.kmin[k] = P.keys[k].lo;
                P.krange[k] = P.keys[k].hi - P.keys[k].lo + 1;
                slots *= (double)P.krange[k];
            }
            int64_t st1 = 1;
            for (int k = P.n_keys - 1; k >= 0; k--) { P.kstride[k] = st1; st1 *= P.krange[k]; }
            q.n_slots = (int64_t)slots;
            ok = slots <= (double)(1 << 24);
            /* a dense build key range takes the slot map, any other the hash join */
            q.kmin = J.rkeys[0].ranged ? J.rkeys[0].lo : 0;
            q.range = J.rkeys[0].ranged ? J.rkeys[0].hi - J.rkeys[0].lo + 1 : 0;
            if (!J.rkeys[0].ranged || q.range <= 0 || q.range > 4 * J.nr + 4096 ||
                q.range >= ((int64_t)1 << 31)) {
                q.hash = 1;
                q.kmin = q.range = 0;
            }
        }
        {
            if (ok) ok = star_plan(J.lkeys[0], S->lpred, q.T, q.SP);
            int agg_slot[TDD_MAX_AGGS] = {};
            for (int a = 0; ok && a < P.n_aggs; a++) {
                if (!P.aggs[a].data) continue;
                int c = tile_slot(q.T, P.aggs[a]);
                if (c < 0) ok = false;
                else agg_slot[a] = c;
            }
            for (int c = 1; ok && c < P.ncells; c++) q.cell_slot[c] = (uint8_t)agg_slot[P.col[c]];
            if (ok && q.hash) ok = bq_record_words(q);
        }
        if (ok && shape_q >= 0) {                   /* one output shape per batch */
            const GroupPlan& P0 = b->hq[shape_q].P;
            ok = P0.n_keys == P.n_keys && P0.n_aggs == P.n_aggs;
            for (int k = 0; ok && k < P.n_keys; k++) ok = P0.keys[k].esz == P.keys[k].esz;
            for (int a = 0; ok && a < P.n_aggs; a++)
                ok = P0.ops[a] == P.ops[a] && P0.out_types[a] == P.out_types[a] &&
                     P0.aggs[a].esz == P.aggs[a].esz;
        }
        if (!ok) st = TD_ERR_NYI;
        b->status[i] = st;
        if (st) continue;
        if (shape_q < 0) shape_q = i;
        q.active = 1;
        q.nl = J.nl;
        q.nr = J.nr;
        q.rk = J.rkeys[0];
        q.map_u16 = q.n_slots < 0xffff;
        q.map_off = (int64_t)map_bytes;
        map_bytes += (((size_t)q.range * (q.map_u16 ? 2 : 4)) + 255) & ~(size_t)255;
        q.tab_off = (int64_t)tab_cells;
        tab_cells += (size_t)q.n_slots * P.ncells;
        const size_t tb = (size_t)q.n_slots * P.ncells * 8;
        q.lds = tb <= BQ_LDS_MAX;
        if (q.lds && !q.hash && tb > lds_max) lds_max = tb;
        b->n_hash += q.hash;
        total_rows += q.nl;
    }
    if (shape_q < 0) { delete b; return TD_ERR_NYI; }   /* nothing to batch */
    b->shape = b->hq[shape_q].P;
    b->map_bytes = map_bytes ? map_bytes : 256;
    b->tab_cells = tab_cells ? tab_cells : 1;
    b->lds_probe = lds_max;
    int rc = TD_OK;
    b->maps = tdd::scratch(b->map_bytes);
    b->tabs = tdd::scratch(b->tab_cells * 8);
    b->occ = tdd::scratch((size_t)n * 8);
    b->ng = tdd::scratch((size_t)n * 8);
    b->offs = tdd::scratch((size_t)(n + 1) * 8);
    if (!b->maps || !b->tabs || !b->occ || !b->ng || !b->offs) rc = TD_ERR_OOM;
    std::vector<BTile> tb_ = make_tiles(n, 16384, [&](int q) {
        return b->status[q] || b->hq[q].hash ? (int64_t)0 : b->hq[q].nr;
    });
    std::vector<BTile> tm_ = make_tiles(n, 65536, [&](int q) { return b->status[q] ? (int64_t)0 : b->hq[q].range; });
    std::vector<BTile> tt_ = make_tiles(n, 65536, [&](int q) {
        return b->status[q] ? (int64_t)0 : b->hq[q].n_slots * b->hq[q].P.ncells;
    });
    b->n_build = (int64_t)tb_.size();
    b->n_map = (int64_t)tm_.size();
    b->n_tab = (int64_t)tt_.size();
    if (!rc) rc = upload(&b->dq, b->hq.data(), sizeof(BQuery) * n);
    if (!rc) rc = upload(&b->tl_build, tb_.data(), tb_.size() * sizeof(BTile));
    if (!rc) rc = upload(&b->tl_map, tm_.data(), tm_.size() * sizeof(BTile));
    if (!rc) rc = upload(&b->tl_tab, tt_.data(), tt_.size() * sizeof(BTile));
    if (!rc) rc = bq_tiles_for(b, 16384);
    if (!rc) bq_grids(b);
    /* unique build keys: one map pass and a count, checked on the host */
    if (!rc) rc = tdd_fill(b->maps, ~0ull, b->map_bytes);
    if (!rc) rc = tdd_fill(b->occ, 0, (size_t)n * 8);
    if (!rc && b->n_build)
        LAUNCH("batch_map", k_bq_scatter, stream_grid(b->n_build, 8), 256, 0, (const BQuery*)b->dq,
               (const BTile*)b->tl_build, b->n_build, (uint8_t*)b->maps);
    if (!rc && b->n_map)
        LAUNCH("batch_map", k_bq_count, stream_grid(b->n_map, 8), 256, 0, (const BQuery*)b->dq,
               (const BTile*)b->tl_map, b->n_map, (const uint8_t*)b->maps,
               (unsigned long long*)b->occ);
    if (!rc) {
        std::vector<uint64_t> occ(n);
        HC(hipMemcpyAsync(occ.data(), b->occ, (size_t)n * 8, hipMemcpyDeviceToHost, tdd::g_stream));
        HC(hipStreamSynchronize(tdd::g_stream));
        bool any_dup = false;
        for (int i = 0; i < n; i++) {
            BQuery& q = b->hq[i];
            if (b->status[i] || q.hash || (int64_t)occ[i] == q.nr) continue;
            /* repeated build keys: the hash join */
            any_dup = true;
            q.hash = 1;
            q.kmin = q.range = 0;
            if (bq_record_words(q)) b->n_hash++;
            else b->status[i] = TD_ERR_NYI;
        }
        if (any_dup) {                               /* drop them from the map work lists */
            b->lds_probe = 0;
            for (int i = 0; i < n; i++) {
                const BQuery& q = b->hq[i];
                const size_t tb = (size_t)q.n_slots * q.P.ncells * 8;
                if (!b->status[i] && !q.hash && q.lds && tb > b->lds_probe) b->lds_probe = tb;
            }
            std::vector<BTile> a2 = make_tiles(n, 16384, [&](int q) {
                return b->status[q] || b->hq[q].hash ? (int64_t)0 : b->hq[q].nr;
            });
            std::vector<BTile> c2 = make_tiles(n, 65536, [&](int q) {
                return b->status[q] ? (int64_t)0 : b->hq[q].n_slots * b->hq[q].P.ncells;
            });
            tdd::release(b->tl_build);
            tdd::release(b->tl_tab);
            b->tl_build = b->tl_tab = nullptr;
            b->n_build = (int64_t)a2.size();
            b->n_tab = (int64_t)c2.size();
            rc = upload(&b->tl_build, a2.data(), a2.size() * sizeof(BTile));
            if (!rc) rc = upload(&b->tl_tab, c2.data(), c2.size() * sizeof(BTile));
            if (!rc) rc = bq_tiles_for(b, 16384);
        }
    }
    if (!rc) rc = bq_bucket_plan(b);
    if (!rc && b->sl_total) {
        HC(hipMemcpyAsync(b->dq, b->hq.data(), sizeof(BQuery) * n, hipMemcpyHostToDevice,
                          tdd::g_stream));
        HC(hipStreamSynchronize(tdd::g_stream));
        rc = bq_tiles_for(b, b->tile_rows);          /* bucketed queries' own probe tiles */
        /* the slice folds store whole tables: no identity fill for them */
        std::vector<BTile> c3 = make_tiles(n, 65536, [&](int q) {
            return b->status[q] || b->hq[q].nsl ? (int64_t)0
                                                : b->hq[q].n_slots * b->hq[q].P.ncells;
        });
        tdd::release(b->tl_tab);
        b->tl_tab = nullptr;
        b->n_tab = (int64_t)c3.size();
        if (!rc) rc = upload(&b->tl_tab, c3.data(), c3.size() * sizeof(BTile));
    }
    if (!rc && b->n_hash) rc = bh_plan(b, BH_TARGET0);
    if (!rc && tune) rc = bq_autotune(b, total_rows);
    if (!rc) bq_grids(b);
    if (rc) { bq_release(b); delete b; return rc; }
    *out = b;
    return TD_OK;
}

extern "C" int tdd_jg_batch_status(const tdd_jg_batch_t* b, int i) { return b->status[i]; }

extern "C" int tdd_jg_batch_shape(const tdd_jg_batch_t* b, int64_t* tile_rows, int* block) {
    *tile_rows = b->tile_rows;
    *block = b->block_probe;
    return TD_OK;
}

extern "C" int tdd_jg_batch_radix(const tdd_jg_batch_t* b, int64_t* part_rows, int* max_bits) {
    int mb = 0;
    for (int i = 0; i < b->n; i++)
        if (!b->status[i] && b->hq[i].hash && b->hq[i].hbits > mb) mb = b->hq[i].hbits;
    *part_rows = b->n_hash ? b->h_target : 0;
    *max_bits = mb;
    return b->n_hash;
}

/* One run: replays the captured graph after the first run.  Outputs: one
 * arena block per output column holding every batched query's groups;
 * offsets[q] / offsets[q + 1] bound query q's rows (n + 1 entries). */
extern "C" int tdd_jg_batch_run(tdd_jg_batch_t* b, int use_graph, tdd_group_result_t* out,
                                int64_t* offsets) {
    memset(out, 0, sizeof(*out));
    int rc;
    if (use_graph && b->exec) {
        HC(hipGraphLaunch(b->exec, tdd::g_stream));
    } else if (use_graph) {
        HC(hipStreamBeginCapture(tdd::g_stream, hipStreamCaptureModeThreadLocal));
        rc = bq_launch(b);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(tdd::g_stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        HC(ce);
        b->graph = g;
        HC(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        HC(hipGraphLaunch(b->exec, tdd::g_stream));
    } else {
        rc = bq_launch(b);
        if (rc) return rc;
    }
    const int n = b->n;
    LAUNCH("batch_emit", k_bq_ngroups, n, 256, 0, (const BQuery*)b->dq, (const uint64_t*)b->tabs,
           (int64_t*)b->ng);
    rc = tdd::scan_i64((const int64_t*)b->ng, n, (int64_t*)b->offs);
    if (rc) return rc;
    HC(hipMemcpyAsync(offsets, b->offs, (size_t)(n + 1) * 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    rc = alloc_group_outputs(b->shape, offsets[n], out);
    if (rc) return rc;
    LAUNCH("batch_emit", k_bq_emit, n, 256, 0, (const BQuery*)b->dq, (const uint64_t*)b->tabs,
           (const int64_t*)b->offs, *out);
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}

extern "C" void tdd_jg_batch_free(tdd_jg_batch_t* b) {
    if (!b) return;
    bq_release(b);
    delete b;
}
