/*
 * k_rjoin.hip -- the materialising radix join (exec.c:7431-7950, 8607-8728):
 * an equi-join whose output columns are written directly, with no pair
 * arrays and no random gathers.
 *
 *  1. Partition passes (both sides): a key word per row (the value of a
 *     single integer key, else a hash of all keys), the partition is the top
 *     `bits` of its mix.  Each partition's region is cut into one run per
 *     XCD (workgroup b runs on XCD b mod 8): a histogram pass counts rows per
 *     (partition, XCD) -- LDS counts per workgroup, added into the global
 *     counts -- an exclusive scan gives the runs' offsets, and the scatter
 *     pass appends each row to its run through that run's cursor, carrying
 *     every carried column along.  All workgroups of an XCD append to the
 *     same run of a partition, so the open end of every run (a cache line
 *     per partition and column, ~1 MB for 4,096 partitions) stays in that
 *     XCD's L2 until it is full and is written back whole.  The probe side
 *     carries every output left column, so its partitioned buffers *are* the
 *     join's left output columns; the build side carries the right output
 *     columns.
 *  2. Build + probe, one workgroup per partition: the build partition's key
 *     words go into LDS with a table of 32-bit slots (load <= 0.5, sized to
 *     leave room for two workgroups per CU), then the probe partition streams
 *     through it; a hit copies the build row's payload into the right output
 *     columns at the probe record's own index.
 *
 * This form needs every build key to be unique (each probe row then has at
 * most one partner, so outputs are slotted by probe record) and every build
 * partition to fit one LDS table; the kernel reports either violation and
 * the caller runs the pair-producing join instead.  LEFT joins keep every
 * probe row (zeros for a missing partner); INNER joins compact the slotted
 * output when some probe rows found no partner.
 */
#include "pred_eval.h"
#include "sel_map.h"
#include "join_keys.h"

#include <vector>

#define RJ_MAXC  16              /* carried columns per side */
#define RJ_T     512             /* threads per workgroup */
#define RJ_RPT   8               /* rows per thread per step */
#define RJ_STEP  (RJ_T * RJ_RPT)
#define RJ_SLOTS 8192            /* LDS hash slots per partition table */
#define RJ_CAP   4000            /* build rows per partition table: 16 KB of slots + 16 B a row,
                                    so two build + probe workgroups fit one CU's 160 KB of LDS */
#define RJ_PART_ROWS 2457       /* expected build rows per partition: 0.3 of the slots */
#define RJ_EMPTY 0xffffffffu
#define RJ_MAXB  14              /* radix bits */

#define RJ_XCD   8               /* XCDs: workgroup b runs on XCD b mod 8 */
#define RJ_FMAX  256             /* ways of one scatter pass */

/* One scatter pass over one side.  Pass 1 reads the side's own columns
 * (rows [b * span, ...) per workgroup, through the selection); pass 2 reads
 * the pass-1 buffers, workgroup b taking the tiles of XCD b mod 8 from the
 * tile table.  A row's digit is `dbits` bits of its key word's mix below the
 * top `dshift` bits already used. */
struct RJPass {
    JKeys           K;           /* key columns (pass 2: the carried key column) */
    const uint64_t* sel;         /* pass 1: rows taking part; NULL: all */
    int64_t         n;           /* pass 1: rows */
    int64_t         span;        /* pass 1: rows per workgroup (a multiple of RJ_STEP) */
    const int64_t*  tiles;       /* pass 2: (start, end) per tile, by XCD */
    const int32_t*  xtiles;      /* pass 2: first tile of each XCD, RJ_XCD + 1 entries */
    int32_t         pass2;
    int32_t         plain;       /* key read by a plain 4- or 8-byte load, else 0 */
    int32_t         dshift, dbits;
    int32_t         xruns;       /* cursors per digit: RJ_XCD (pass 1) or 1 (pass 2) */
    int32_t         ncols;
    const void*     src[RJ_MAXC];
    void*           dst[RJ_MAXC];
    uint8_t         esz[RJ_MAXC];
    const uint64_t* kw_src;      /* pass 2: carried key words (key read from them) */
    uint64_t*       kw_dst;      /* carried key words out, NULL: not carried */
    const uint32_t* row_src;     /* pass 2: carried rows */
    uint32_t*       row_dst;     /* carried source rows out, NULL: not carried */
    unsigned long long* cur;     /* run cursors: digit * xruns + run */
};

/* Histogram pass: counts per (final partition, XCD) of each side's rows. */
struct RJHist {
    JKeys           K;
    const uint64_t* sel;
    int64_t         n, span;
    int32_t         bits, plain;
    uint32_t*       C;           /* C[q * RJ_XCD + x] */
};

__device__ __forceinline__ uint32_t rj_part(uint64_t kw, int bits) {
    return bits ? (uint32_t)(jmix(kw) >> (64 - bits)) : 0u;
}

__device__ __forceinline__ uint32_t rj_slot(uint64_t kw) {
    return (uint32_t)jmix(kw) & (RJ_SLOTS - 1);
}

/* Key word of row r of key set K read the fast way when `plain`. */
__device__ __forceinline__ uint64_t rj_kw(const JKeys& K, int plain, int64_t r) {
    const void* d = K.cols[0].data;
    if (plain == 8) return (uint64_t)((const int64_t*)d)[r];
    if (plain == 4)
        return K.cols[0].type == TD_SYM ? (uint64_t)((const uint32_t*)d)[r]
                                        : (uint64_t)(int64_t)((const int32_t*)d)[r];
    return key_word(K, r);
}

__global__ __launch_bounds__(RJ_T) void k_rj_hist(RJHist H) {
    extern __shared__ uint32_t hist[];
    const int Q = 1 << H.bits;
    for (int i = threadIdx.x; i < Q; i += RJ_T) hist[i] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * H.span;
    const int64_t r1 = r0 + H.span < H.n ? r0 + H.span : H.n;
    for (int64_t base = r0; base < r1; base += RJ_STEP) {
        uint64_t kw[RJ_RPT];
        uint32_t live = 0;
#pragma unroll
        for (int j = 0; j < RJ_RPT; j++) {
            const int64_t r = base + j * RJ_T + threadIdx.x;
            if (r < r1 && row_selected(H.sel, r)) live |= 1u << j;
        }
        if (H.plain) {
#pragma unroll
            for (int j = 0; j < RJ_RPT; j++) {
                const int64_t r = base + j * RJ_T + threadIdx.x;
                kw[j] = rj_kw(H.K, H.plain, r < r1 ? r : r0);
            }
        } else {
            for (int j = 0; j < RJ_RPT; j++)
                kw[j] = ((live >> j) & 1) ? key_word(H.K, base + j * RJ_T + threadIdx.x) : 0;
        }
#pragma unroll
        for (int j = 0; j < RJ_RPT; j++)
            if ((live >> j) & 1) atomicAdd(&hist[rj_part(kw[j], H.bits)], 1u);
    }
    __syncthreads();
    const int x = blockIdx.x % RJ_XCD;
    for (int q = threadIdx.x; q < Q; q += RJ_T)
        if (hist[q]) atomicAdd(&H.C[(int64_t)q * RJ_XCD + x], hist[q]);
}

/* One carried column of a tile through LDS: the thread's rows read in
 * natural order (coalesced), placed at their sorted positions, then each
 * thread's sorted positions stored to their runs (consecutive lanes,
 * consecutive addresses).  The caller's barrier separates the uses. */
template <class T>
__device__ __forceinline__ void rj_stage(const T* src, T* dst, int64_t tb, uint32_t live,
                                         const uint32_t (&at)[RJ_RPT], int cnt,
                                         const uint32_t (&pos)[RJ_RPT], uint64_t* stage) {
    T v[RJ_RPT];
#pragma unroll
    for (int j = 0; j < RJ_RPT; j++)
        if ((live >> j) & 1) v[j] = src[tb + j * RJ_T + threadIdx.x];
#pragma unroll
    for (int j = 0; j < RJ_RPT; j++)
        if ((live >> j) & 1) stage[at[j]] = (uint64_t)v[j];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RJ_RPT; k++) {
        const int i = threadIdx.x + k * RJ_T;
        if (i < cnt) dst[pos[k]] = (T)stage[i];
    }
    __syncthreads();
}

/* One scatter pass.  Per tile of up to RJ_STEP consecutive rows: each row's
 * digit and LDS rank, one global reservation per digit of the tile in its
 * run, then every carried column staged through LDS into digit order and
 * written so that consecutive lanes store consecutive addresses of a run. */
__global__ __launch_bounds__(RJ_T) void k_rj_scatter(RJPass S) {
    __shared__ uint32_t hist[RJ_FMAX], loff[RJ_FMAX], gbase[RJ_FMAX];
    __shared__ uint8_t spart[RJ_STEP];
    __shared__ uint64_t stage[RJ_STEP];
    __shared__ uint32_t s_hi;                    /* the tile's digits above this pass */
    const int F = 1 << S.dbits;
    const int x = blockIdx.x % RJ_XCD;
    const int run = S.xruns > 1 ? x : 0;
    if (threadIdx.x == 0) s_hi = 0;             /* pass 1: no digits above; published by
                                                   the first tile's barrier */
    int64_t tb = 0, te = 0;
    int64_t next = blockIdx.x;                  /* pass 2: tile of this XCD */
    int32_t t_end = 0, t_step = 1;
    int64_t r_end = 0;
    if (S.pass2) {
        next = S.xtiles[x] + blockIdx.x / RJ_XCD;
        t_end = S.xtiles[x + 1];
        t_step = (gridDim.x + RJ_XCD - 1 - x) / RJ_XCD;  /* workgroups of this XCD */
        if (t_step < 1) t_step = 1;
    } else {
        tb = (int64_t)blockIdx.x * S.span;
        r_end = tb + S.span < S.n ? tb + S.span : S.n;
    }
    for (;;) {
        if (S.pass2) {
            if (next >= t_end) break;
            tb = S.tiles[2 * next];
            te = S.tiles[2 * next + 1];
            next += t_step;
        } else {
            if (tb >= r_end) break;
            te = tb + RJ_STEP < r_end ? tb + RJ_STEP : r_end;
        }
        for (int i = threadIdx.x; i < F; i += RJ_T) hist[i] = 0;
        __syncthreads();
        uint64_t kw[RJ_RPT];
        uint32_t dig[RJ_RPT], at[RJ_RPT];
        uint32_t live = 0;
#pragma unroll
        for (int j = 0; j < RJ_RPT; j++) {
            const int64_t r = tb + j * RJ_T + threadIdx.x;
            if (r < te && (S.pass2 || row_selected(S.sel, r))) live |= 1u << j;
        }
        if (S.kw_src) {
#pragma unroll
            for (int j = 0; j < RJ_RPT; j++) {
                const int64_t r = tb + j * RJ_T + threadIdx.x;
                kw[j] = S.kw_src[r < te ? r : tb];
            }
        } else if (S.plain) {
#pragma unroll
            for (int j = 0; j < RJ_RPT; j++) {
                const int64_t r = tb + j * RJ_T + threadIdx.x;
                kw[j] = rj_kw(S.K, S.plain, r < te ? r : tb);
            }
        } else {
            for (int j = 0; j < RJ_RPT; j++)
                kw[j] = ((live >> j) & 1) ? key_word(S.K, tb + j * RJ_T + threadIdx.x) : 0;
        }
#pragma unroll
        for (int j = 0; j < RJ_RPT; j++) {
            if (!((live >> j) & 1)) continue;
            const uint64_t h = jmix(kw[j]);
            dig[j] = (uint32_t)(h >> (64 - S.dshift - S.dbits)) & (uint32_t)(F - 1);
            at[j] = atomicAdd(&hist[dig[j]], 1u);
            /* pass 2: every row of a tile lies in one pass-1 partition */
            if (S.dshift && j == 0 && threadIdx.x == 0) s_hi = (uint32_t)(h >> (64 - S.dshift));
        }
        __syncthreads();
        if (threadIdx.x < WAVE) {                /* scan of <= 256 digit counts by one wave */
            uint32_t carry = 0;
            for (int d0 = 0; d0 < F; d0 += WAVE) {
                const int d = d0 + threadIdx.x;
                const uint32_t c = d < F ? hist[d] : 0;
                uint32_t incl = c;
                for (int o = 1; o < WAVE; o <<= 1) {
                    const uint32_t y = __shfl_up(incl, o, WAVE);
                    if (threadIdx.x >= o) incl += y;
                }
                if (d < F) {
                    loff[d] = carry + incl - c;
                    const int64_t q = ((int64_t)s_hi << S.dbits) | d;
                    gbase[d] = c ? (uint32_t)atomicAdd(&S.cur[q * S.xruns + run],
                                                       (unsigned long long)c) : 0u;
                }
                carry += __shfl(incl, WAVE - 1, WAVE);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RJ_RPT; j++) {
            if (!((live >> j) & 1)) continue;
            at[j] += loff[dig[j]];               /* sorted position in the tile */
            spart[at[j]] = (uint8_t)dig[j];
        }
        __syncthreads();
        const int cnt = (int)(loff[F - 1] + hist[F - 1]);
        uint32_t pos[RJ_RPT];
#pragma unroll
        for (int k = 0; k < RJ_RPT; k++) {
            const int i = threadIdx.x + k * RJ_T;
            pos[k] = i < cnt ? gbase[spart[i]] + (uint32_t)(i - (int)loff[spart[i]]) : 0u;
        }
        for (int c = 0; c < S.ncols; c++) {                 /* uniform across the block */
            switch (S.esz[c]) {
            case 8: rj_stage<uint64_t>((const uint64_t*)S.src[c], (uint64_t*)S.dst[c], tb, live,
                                       at, cnt, pos, stage); break;
            case 4: rj_stage<uint32_t>((const uint32_t*)S.src[c], (uint32_t*)S.dst[c], tb, live,
                                       at, cnt, pos, stage); break;
            case 2: rj_stage<uint16_t>((const uint16_t*)S.src[c], (uint16_t*)S.dst[c], tb, live,
                                       at, cnt, pos, stage); break;
            default: rj_stage<uint8_t>((const uint8_t*)S.src[c], (uint8_t*)S.dst[c], tb, live,
                                       at, cnt, pos, stage); break;
            }
        }
        if (S.kw_dst) {
#pragma unroll
            for (int j = 0; j < RJ_RPT; j++)
                if ((live >> j) & 1) stage[at[j]] = kw[j];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RJ_RPT; k++) {
                const int i = threadIdx.x + k * RJ_T;
                if (i < cnt) S.kw_dst[pos[k]] = stage[i];
            }
            __syncthreads();
        }
        if (S.row_dst) {
            if (S.row_src)
                rj_stage<uint32_t>(S.row_src, S.row_dst, tb, live, at, cnt, pos, stage);
            else {
#pragma unroll
                for (int j = 0; j < RJ_RPT; j++)
                    if ((live >> j) & 1) stage[at[j]] = (uint64_t)(tb + j * RJ_T + threadIdx.x);
                __syncthreads();
#pragma unroll
                for (int k = 0; k < RJ_RPT; k++) {
                    const int i = threadIdx.x + k * RJ_T;
                    if (i < cnt) S.row_dst[pos[k]] = (uint32_t)stage[i];
                }
                __syncthreads();
            }
        }
        if (!S.pass2) tb = te;
    }
}

/* Build + probe arguments: partition p of a side spans
 * [off[p], off[p + 1]) of its partitioned buffers. */
struct RJBP {
    const int64_t*  loff;
    const int64_t*  roff;
    tdd_col_t       lkey;        /* partitioned probe key column (exact keys, when carried) */
    const uint64_t* lkw;         /* else partitioned probe key words */
    const uint64_t* rkw;         /* partitioned build key words */
    const uint32_t* lrow;        /* hashed keys: source rows, for verification */
    const uint32_t* rrow;
    JKeys           LK, RK;      /* original key columns (verification) */
    int32_t         exact, left_join, nrc, pad;
    const void*     rsrc[RJ_MAXC];   /* partitioned build payload */
    void*           rdst[RJ_MAXC];   /* right output columns, slotted by probe record */
    uint8_t         resz[RJ_MAXC];
    uint8_t*        lflag;       /* 1 when the probe record found its partner; NULL: not kept */
    unsigned long long* stat;    /* [0] matched records, [1] repeated build key, [2] oversized partition */
};

#define RJ_PB 4                  /* probe records per thread in flight */
#define RJ_BT 1024               /* threads per build + probe workgroup */

/* Payload of the build rows bs + hl[] into a right output column at the probe
 * records i0 + j * RJ_T (zero for a LEFT row without a partner). */
template <class T>
__device__ __forceinline__ void rj_fetch(const void* src, void* dst, int64_t i0, int64_t pe,
                                         int64_t bs, const int32_t (&hl)[RJ_PB], bool left_join) {
    T v[RJ_PB];
#pragma unroll
    for (int j = 0; j < RJ_PB; j++) v[j] = hl[j] >= 0 ? ((const T*)src)[bs + hl[j]] : (T)0;
#pragma unroll
    for (int j = 0; j < RJ_PB; j++) {
        const int64_t i = i0 + j * RJ_BT;
        if (i < pe && (hl[j] >= 0 || left_join)) ((T*)dst)[i] = v[j];
    }
}

/* The first right column, from its LDS copy: build-local rows hl[] (-1: no
 * partner) into the output at the probe records i0 + j * RJ_BT. */
template <class T>
__device__ __forceinline__ void rj_put_lds(void* dst, int64_t i0, int64_t pe,
                                           const int32_t (&hl)[RJ_PB], const uint64_t* pay,
                                           bool left_join) {
#pragma unroll
    for (int j = 0; j < RJ_PB; j++) {
        const int64_t i = i0 + j * RJ_BT;
        if (i < pe && (hl[j] >= 0 || left_join)) ((T*)dst)[i] = hl[j] >= 0 ? (T)pay[hl[j]] : (T)0;
    }
}

#define RJ_SLOT_EMPTY 0xffffu

/* 16-bit slot s of the packed table (two slots per word). */
__device__ __forceinline__ uint32_t rj_slot_get(const uint32_t* tab, uint32_t s) {
    return (tab[s >> 1] >> ((s & 1) * 16)) & 0xffffu;
}

/* Claim slot s for build row li when it is empty; returns the occupant
 * otherwise (RJ_SLOT_EMPTY when the claim succeeded). */
__device__ __forceinline__ uint32_t rj_slot_claim(uint32_t* tab, uint32_t s, uint32_t li) {
    const uint32_t sh = (s & 1) * 16;
    uint32_t old = tab[s >> 1];
    for (;;) {
        const uint32_t cur = (old >> sh) & 0xffffu;
        if (cur != RJ_SLOT_EMPTY) return cur;
        const uint32_t nw = (old & ~(0xffffu << sh)) | (li << sh);
        const uint32_t prev = atomicCAS(&tab[s >> 1], old, nw);
        if (prev == old) return RJ_SLOT_EMPTY;
        old = prev;                              /* the other half changed: retry */
    }
}

template <int PLAIN>
__global__ __launch_bounds__(RJ_BT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_rj_bp(RJBP A) {
    extern __shared__ uint8_t sm[];
    uint32_t* slot = (uint32_t*)sm;                                 /* RJ_SLOTS / 2 words */
    uint64_t* tkey = (uint64_t*)(sm + RJ_SLOTS * 2);                /* RJ_CAP */
    uint64_t* pay = (uint64_t*)(sm + RJ_SLOTS * 2 + RJ_CAP * 8);    /* RJ_CAP: right column 0 */
    const int p = blockIdx.x;
    const int64_t bs = A.roff[p], be = A.roff[p + 1];
    const int64_t ps = A.loff[p], pe = A.loff[p + 1];
    if (be - bs > RJ_CAP) {                      /* uniform: the caller joins another way */
        if (threadIdx.x == 0) atomicAdd(&A.stat[2], 1ull);
        return;
    }
    for (int i = threadIdx.x; i < RJ_SLOTS / 2; i += RJ_BT) slot[i] = 0xffffffffu;
    __syncthreads();
    const int e0 = A.nrc ? A.resz[0] : 0;
    for (int64_t i = bs + threadIdx.x; i < be; i += RJ_BT) {
        const uint64_t k = A.rkw[i];
        const uint32_t li = (uint32_t)(i - bs);
        tkey[li] = k;                            /* before the slot is published */
        if (e0 == 8) pay[li] = ((const uint64_t*)A.rsrc[0])[i];
        else if (e0 == 4) pay[li] = ((const uint32_t*)A.rsrc[0])[i];
        else if (e0 == 2) pay[li] = ((const uint16_t*)A.rsrc[0])[i];
        else if (e0 == 1) pay[li] = ((const uint8_t*)A.rsrc[0])[i];
        uint32_t s = rj_slot(k);
        for (;;) {
            const uint32_t prev = rj_slot_claim(slot, s, li);
            if (prev == RJ_SLOT_EMPTY) break;
            if (tkey[prev] == k &&
                (A.exact || keys_equal(A.RK, A.rrow[bs + prev], A.RK, A.rrow[i])))
                A.stat[1] = 1;                   /* a repeated build key */
            s = (s + 1) & (RJ_SLOTS - 1);
        }
    }
    __syncthreads();
    uint32_t hits = 0;
    for (int64_t i0 = ps + threadIdx.x; i0 < pe; i0 += RJ_PB * RJ_BT) {
        uint64_t k[RJ_PB];
        int32_t hl[RJ_PB];
#pragma unroll
        for (int j = 0; j < RJ_PB; j++) {           /* every key load first */
            const int64_t i = i0 + j * RJ_BT;
            const int64_t ii = i < pe ? i : ps;
            if (PLAIN == 8) k[j] = (uint64_t)((const int64_t*)A.lkey.data)[ii];
            else if (PLAIN == 4)
                k[j] = A.lkey.type == TD_SYM ? (uint64_t)((const uint32_t*)A.lkey.data)[ii]
                                             : (uint64_t)(int64_t)((const int32_t*)A.lkey.data)[ii];
            else k[j] = A.lkw ? A.lkw[ii] : (uint64_t)ld_i64(A.lkey, ii);
        }
        /* the RJ_PB probes advance together, so their LDS reads overlap */
        uint32_t sl[RJ_PB], live = 0;
#pragma unroll
        for (int j = 0; j < RJ_PB; j++) {
            hl[j] = -1;
            sl[j] = rj_slot(k[j]);
            if (i0 + j * RJ_BT < pe) live |= 1u << j;
        }
        while (live) {
            uint32_t li[RJ_PB];
#pragma unroll
            for (int j = 0; j < RJ_PB; j++) li[j] = (live >> j) & 1 ? rj_slot_get(slot, sl[j]) : 0u;
#pragma unroll
            for (int j = 0; j < RJ_PB; j++) {
                if (!((live >> j) & 1)) continue;
                if (li[j] == RJ_SLOT_EMPTY) { live &= ~(1u << j); continue; }
                if (tkey[li[j]] == k[j] &&
                    (A.exact || keys_equal(A.LK, A.lrow[i0 + j * RJ_BT], A.RK,
                                           A.rrow[bs + li[j]]))) {
                    hl[j] = (int32_t)li[j];
                    live &= ~(1u << j);
                    continue;
                }
                sl[j] = (sl[j] + 1) & (RJ_SLOTS - 1);
            }
        }
#pragma unroll
        for (int j = 0; j < RJ_PB; j++) hits += hl[j] >= 0;
        switch (e0) {                                       /* uniform across the block */
        case 8: rj_put_lds<uint64_t>(A.rdst[0], i0, pe, hl, pay, A.left_join); break;
        case 4: rj_put_lds<uint32_t>(A.rdst[0], i0, pe, hl, pay, A.left_join); break;
        case 2: rj_put_lds<uint16_t>(A.rdst[0], i0, pe, hl, pay, A.left_join); break;
        case 1: rj_put_lds<uint8_t>(A.rdst[0], i0, pe, hl, pay, A.left_join); break;
        default: break;
        }
        for (int c = 1; c < A.nrc; c++) {                   /* uniform across the block */
            switch (A.resz[c]) {
            case 8: rj_fetch<uint64_t>(A.rsrc[c], A.rdst[c], i0, pe, bs, hl, A.left_join); break;
            case 4: rj_fetch<uint32_t>(A.rsrc[c], A.rdst[c], i0, pe, bs, hl, A.left_join); break;
            case 2: rj_fetch<uint16_t>(A.rsrc[c], A.rdst[c], i0, pe, bs, hl, A.left_join); break;
            default: rj_fetch<uint8_t>(A.rsrc[c], A.rdst[c], i0, pe, bs, hl, A.left_join); break;
            }
        }
        if (A.lflag) {
#pragma unroll
            for (int j = 0; j < RJ_PB; j++)
                if (i0 + j * RJ_BT < pe) A.lflag[i0 + j * RJ_BT] = hl[j] >= 0;
        }
    }
    for (int d = WAVE / 2; d >= 1; d >>= 1) hits += __shfl_xor(hits, d);
    /* one atomic per workgroup (thousands of workgroups share the counter):
     * the wave counts go through the slot table's words once every wave is
     * past its probes */
    __syncthreads();
    if (lane_id() == 0) slot[threadIdx.x / WAVE] = hits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < RJ_BT / WAVE; w++) t += slot[w];
        if (t) atomicAdd(&A.stat[0], t);
    }
}

/* Plan of one side's scatter passes from its counts per (final partition,
 * XCD), on the device so that no count travels to the host between the
 * passes: the final partitions' offsets, the pass-1 run cursors (ordered
 * pass-1 partition, then XCD), the pass-2 cursors and its tile list (the
 * pass-1 partitions of XCD x are those p1 = x mod 8, each cut into tiles of
 * RJ_STEP rows), and the side's row total.  One workgroup per side. */
#define RJP_T 1024

struct RJPlan {
    const uint32_t*     C;       /* C[q * RJ_XCD + x] */
    int32_t             bits, b1, b2, pad;
    int64_t*            off;     /* Q + 1 */
    unsigned long long* cur1;    /* P1 * RJ_XCD */
    unsigned long long* cur2;    /* Q (b2 > 0) */
    int64_t*            tiles;   /* (start, end) pairs (b2 > 0) */
    int32_t*            xtiles;  /* RJ_XCD + 1 (b2 > 0) */
    unsigned long long* total;   /* the side's rows */
};

/* Exclusive scan of one value per thread over the workgroup; *tot gets
 * the sum.  Every thread calls it. */
__device__ __forceinline__ int64_t rj_block_scan(int64_t v, int64_t* ws, int64_t* tot) {
    int64_t incl = v;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t y = __shfl_up(incl, d, WAVE);
        if (lane_id() >= d) incl += y;
    }
    if (lane_id() == WAVE - 1) ws[threadIdx.x / WAVE] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t a = 0;
        for (int w = 0; w < RJP_T / WAVE; w++) {
            const int64_t t = ws[w];
            ws[w] = a;
            a += t;
        }
        ws[RJP_T / WAVE] = a;
    }
    __syncthreads();
    const int64_t ex = ws[threadIdx.x / WAVE] + incl - v;
    *tot = ws[RJP_T / WAVE];
    __syncthreads();
    return ex;
}

struct RJPlans {
    RJPlan side[2];
};

__global__ __launch_bounds__(RJP_T) void k_rj_plan(RJPlans PS) {
    const RJPlan& A = PS.side[blockIdx.x];
    __shared__ int64_t ws[RJP_T / WAVE + 1];
    __shared__ int64_t p1s[RJ_FMAX + 1];         /* pass-1 partition starts */
    __shared__ int64_t tb[RJ_FMAX];              /* first pass-2 tile of each pass-1 partition */
    const int Q = 1 << A.bits, P1 = 1 << A.b1, sub = 1 << A.b2;
    /* final partitions: per consecutive q of this thread */
    const int per = (Q + RJP_T - 1) / RJP_T, q0 = (int)threadIdx.x * per;
    const int q1 = q0 + per < Q ? q0 + per : Q;
    int64_t s = 0;
    for (int q = q0; q < q1; q++)
        for (int x = 0; x < RJ_XCD; x++) s += A.C[(int64_t)q * RJ_XCD + x];
    int64_t tot;
    int64_t a = rj_block_scan(s, ws, &tot);
    for (int q = q0; q < q1; q++) {
        A.off[q] = a;
        if (A.b2) A.cur2[q] = (unsigned long long)a;
        for (int x = 0; x < RJ_XCD; x++) a += A.C[(int64_t)q * RJ_XCD + x];
    }
    if (threadIdx.x == 0) {
        A.off[Q] = tot;
        *A.total = (unsigned long long)tot;
    }
    /* pass-1 runs e = p1 * RJ_XCD + x, in that order */
    const int E = P1 * RJ_XCD;
    const int pe = (E + RJP_T - 1) / RJP_T, e0 = (int)threadIdx.x * pe;
    const int e1 = e0 + pe < E ? e0 + pe : E;
    s = 0;
    for (int e = e0; e < e1; e++)
        for (int q = (e / RJ_XCD) * sub; q < (e / RJ_XCD + 1) * sub; q++)
            s += A.C[(int64_t)q * RJ_XCD + e % RJ_XCD];
    a = rj_block_scan(s, ws, &tot);
    for (int e = e0; e < e1; e++) {
        A.cur1[e] = (unsigned long long)a;
        if (e % RJ_XCD == 0) p1s[e / RJ_XCD] = a;
        for (int q = (e / RJ_XCD) * sub; q < (e / RJ_XCD + 1) * sub; q++)
            a += A.C[(int64_t)q * RJ_XCD + e % RJ_XCD];
    }
    if (threadIdx.x == 0) p1s[P1] = tot;
    __syncthreads();
    if (!A.b2) return;
    /* pass-2 tiles: XCD by XCD, its pass-1 partitions in order, each in
     * RJ_STEP-row tiles */
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int x = 0; x < RJ_XCD; x++) {
            A.xtiles[x] = (int32_t)t;
            for (int p1 = x; p1 < P1; p1 += RJ_XCD) {
                tb[p1] = t;
                t += (p1s[p1 + 1] - p1s[p1] + RJ_STEP - 1) / RJ_STEP;
            }
        }
        A.xtiles[RJ_XCD] = (int32_t)t;
    }
    __syncthreads();
    for (int p1 = threadIdx.x; p1 < P1; p1 += RJP_T) {
        const int64_t b = p1s[p1], e = p1s[p1 + 1];
        int64_t t = tb[p1];
        for (int64_t r = b; r < e; r += RJ_STEP, t++) {
            A.tiles[2 * t] = r;
            A.tiles[2 * t + 1] = r + RJ_STEP < e ? r + RJ_STEP : e;
        }
    }
}

/* ----------------------------------------------------------- host side -- */

static bool rj_is_int(int8_t t) {
    return t != TD_F64 && t != TD_GUID && t != TD_LIST && t != TD_TABLE && t != TD_CHAR;
}

static int rj_plain(const JKeys& K) {
    if (!K.exact || K.n != 1) return 0;
    const tdd_col_t& c = K.cols[0];
    switch (c.type) {
    case TD_I64: case TD_TIMESTAMP: return 8;
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

/* Owned device blocks freed on scope exit unless released. */
struct RJBlocks {
    void* p[4 * RJ_MAXC + 8] = {};
    int n = 0;
    void* take(size_t bytes) {
        void* q = td__arena_alloc(bytes ? bytes : 8);
        if (q) p[n++] = q;
        return q;
    }
    void keep(void* q) { for (int i = 0; i < n; i++) if (p[i] == q) p[i] = nullptr; }
    ~RJBlocks() { for (int i = 0; i < n; i++) if (p[i]) td__arena_free(p[i]); }
};

/* One side through the partition passes, host view. */
struct RJHostSide {
    JKeys           K;
    const uint64_t* sel = nullptr;
    int64_t         n = 0;
    int             plain = 0, ncols = 0;
    const void*     src[RJ_MAXC] = {};
    uint8_t         esz[RJ_MAXC] = {};
    int             key_col = -1;        /* carried column holding the exact key */
    bool            carry_kw = false, carry_row = false;
    int64_t         span = 0;
    int32_t         nspan = 0;
    /* partitioned results */
    void*           dst[RJ_MAXC] = {};
    uint64_t*       kw = nullptr;
    uint32_t*       row = nullptr;
};

/* Workgroup spans of pass 1: enough workgroups to fill the chip twice. */
static void rj_spans(RJHostSide& S) {
    const int64_t steps = ceil_div(S.n > 0 ? S.n : 1, RJ_STEP);
    int64_t want = (int64_t)tdd::g_num_cu * TUNE_INT("TEIDE_RJ_WG_PER_CU", 2);
    if (want > steps) want = steps;
    if (want < 1) want = 1;
    S.span = ceil_div(steps, want) * RJ_STEP;
    S.nspan = (int32_t)ceil_div(S.n > 0 ? S.n : 1, S.span);
}

static int rj_hist(RJHostSide& S, int bits, Scratch& C) {
    const int Q = 1 << bits;
    int rc = tdd_fill(C.p, 0, (size_t)Q * RJ_XCD * 4);
    if (rc) return rc;
    RJHist H;
    memset(&H, 0, sizeof(H));
    H.K = S.K;
    H.sel = S.sel;
    H.n = S.n;
    H.span = S.span;
    H.bits = bits;
    H.plain = S.plain;
    H.C = C.as<uint32_t>();
    LAUNCH("rjoin_hist", k_rj_hist, S.nspan, RJ_T, (size_t)Q * 4, H);
    return TD_OK;
}

/* Device plan of one side's passes from its counts C; the side's row
 * total lands in *total on the device. */
static RJPlan rj_plan_of(int bits, const Scratch& C, Scratch& offs, Scratch& cur1, Scratch& cur2,
                         Scratch& tiles, Scratch& xt, unsigned long long* total) {
    RJPlan PL;
    memset(&PL, 0, sizeof(PL));
    PL.C = C.as<uint32_t>();
    PL.bits = bits;
    PL.b1 = bits <= 8 ? bits : (bits + 1) / 2;
    PL.b2 = bits - PL.b1;
    PL.off = offs.as<int64_t>();
    PL.cur1 = cur1.as<unsigned long long>();
    PL.cur2 = cur2.as<unsigned long long>();
    PL.tiles = tiles.as<int64_t>();
    PL.xtiles = xt.as<int32_t>();
    PL.total = total;
    return PL;
}

/* The scatter passes of one side after its plan: one pass (bits <= 8) or
 * two (b1 top bits, then b2).  Buffers hold S.n rows (every selected row
 * at most). */
static int rj_partition(RJHostSide& S, int bits, void* const* final_dst, Scratch& cur1,
                        Scratch& cur2, Scratch& tiles, Scratch& xt, Scratch& mid, Scratch& kw,
                        Scratch& kwm, Scratch& row, Scratch& rowm) {
    const int b1 = bits <= 8 ? bits : (bits + 1) / 2, b2 = bits - b1;
    const int64_t n = S.n > 0 ? S.n : 1;
    /* buffers: final, and the pass-1 copies when there is a second pass */
    size_t mid_off[RJ_MAXC] = {};
    if (b2) {
        size_t need = 0;
        for (int c = 0; c < S.ncols; c++) {
            mid_off[c] = need;
            need += ((size_t)n * S.esz[c] + 255) & ~(size_t)255;
        }
        if (!mid.reset(need)) return TD_ERR_OOM;
    }
    if (S.carry_kw) {
        if (!kw.reset((size_t)n * 8) || (b2 && !kwm.reset((size_t)n * 8))) return TD_ERR_OOM;
        S.kw = kw.as<uint64_t>();
    }
    if (S.carry_row) {
        if (!row.reset((size_t)n * 4) || (b2 && !rowm.reset((size_t)n * 4))) return TD_ERR_OOM;
        S.row = row.as<uint32_t>();
    }
    for (int c = 0; c < S.ncols; c++) S.dst[c] = final_dst[c];

    RJPass A;
    memset(&A, 0, sizeof(A));
    A.K = S.K;
    A.sel = S.sel;
    A.n = S.n;
    A.span = S.span;
    A.plain = S.plain;
    A.dshift = 0;
    A.dbits = b1;
    A.xruns = RJ_XCD;
    A.ncols = S.ncols;
    for (int c = 0; c < S.ncols; c++) {
        A.src[c] = S.src[c];
        A.esz[c] = S.esz[c];
        A.dst[c] = b2 ? (void*)(mid.as<uint8_t>() + mid_off[c]) : S.dst[c];
    }
    A.kw_dst = S.carry_kw ? (b2 ? kwm.as<uint64_t>() : S.kw) : nullptr;
    A.row_dst = S.carry_row ? (b2 ? rowm.as<uint32_t>() : S.row) : nullptr;
    A.cur = cur1.as<unsigned long long>();
    LAUNCH("rjoin_scatter", k_rj_scatter, S.nspan, RJ_T, 0, A);
    if (!b2) return TD_OK;
    POLL_CANCEL();

    RJPass B = A;
    B.pass2 = 1;
    B.sel = nullptr;
    B.tiles = tiles.as<int64_t>();
    B.xtiles = xt.as<int32_t>();
    B.dshift = b1;
    B.dbits = b2;
    B.xruns = 1;
    for (int c = 0; c < S.ncols; c++) {
        B.src[c] = A.dst[c];
        B.dst[c] = S.dst[c];
    }
    if (S.carry_kw) {
        B.kw_src = kwm.as<uint64_t>();
        B.kw_dst = S.kw;
    } else {
        B.K.cols[0].data = A.dst[S.key_col];     /* the carried exact key */
    }
    B.row_src = S.carry_row ? rowm.as<uint32_t>() : nullptr;
    B.row_dst = S.carry_row ? S.row : nullptr;
    B.cur = cur2.as<unsigned long long>();
    /* the tile list is on the device: workgroups to fill each XCD, capped
     * by the most tiles an XCD can have (every row of the side on it) */
    int64_t per_xcd = tdd::g_num_cu / RJ_XCD * TUNE_INT("TEIDE_RJ_WG_PER_CU", 2);
    const int64_t maxt = ceil_div(n, RJ_STEP) + ((1 << b1) + RJ_XCD - 1) / RJ_XCD;
    if (per_xcd > maxt) per_xcd = maxt;
    if (per_xcd < 1) per_xcd = 1;
    LAUNCH("rjoin_scatter", k_rj_scatter, (int)per_xcd * RJ_XCD, RJ_T, 0, B);
    return TD_OK;
}

/* ---------------------------------------------- dense build keys: a map -- */
/* When the build keys are unique integers in a dense range, a map replaces
 * the partitions: each right column is scattered to its key's place
 * (dense[c][key - kmin]) beside a present bit per key, and the output rows
 * are the selected probe rows (INNER: those whose key is present) in row
 * order, left columns compacted and right columns read at their keys in
 * the same pass -- the direct-map counterpart of the reference's chained
 * join for this case (exec.c:7974-8128). */
struct RJDBuild {
    tdd_col_t   key;
    int64_t     nr, kmin;
    uint8_t*    present;     /* a byte per key of the range: plain stores, no atomics */
    int32_t     n, pad;
    const void* src[RJ_MAXC];
    void*       dst[RJ_MAXC];
    uint8_t     esz[RJ_MAXC];
};

template <class T>
__device__ __forceinline__ void rjd_put(const RJDBuild& B, int c, int64_t r, uint64_t i) {
    ((T*)B.dst[c])[i] = ((const T*)B.src[c])[r];
}

#define RJD_R 4                  /* build rows per thread per step */

__global__ void k_rjd_build(RJDBuild B) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * RJD_R;
    for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x * RJD_R + threadIdx.x; r0 < B.nr;
         r0 += stride) {
        uint64_t i[RJD_R];
#pragma unroll
        for (int j = 0; j < RJD_R; j++) {           /* the step's key loads first */
            const int64_t r = r0 + (int64_t)j * blockDim.x;
            i[j] = r < B.nr ? (uint64_t)ld_i64(B.key, r) - (uint64_t)B.kmin : 0;
        }
#pragma unroll
        for (int j = 0; j < RJD_R; j++) {
            const int64_t r = r0 + (int64_t)j * blockDim.x;
            if (r >= B.nr) continue;
            B.present[i[j]] = 1;
            for (int c = 0; c < B.n; c++) {
                switch (B.esz[c]) {
                case 8: rjd_put<uint64_t>(B, c, r, i[j]); break;
                case 4: rjd_put<uint32_t>(B, c, r, i[j]); break;
                case 2: rjd_put<uint16_t>(B, c, r, i[j]); break;
                default: rjd_put<uint8_t>(B, c, r, i[j]); break;
                }
            }
        }
    }
}

/* Keys present (the build side repeats none exactly when this equals nr),
 * and, when `bits` is given, the present bytes packed as bits: lane l of a
 * wave reads key w * 64 + l and the wave's ballot is bits word w. */
__global__ void k_rjd_count(const uint8_t* present, int64_t range, uint64_t* bits,
                            unsigned long long* count) {
    unsigned long long c = 0;
    const int64_t nw = (range + 63) >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t nwave = (int64_t)gridDim.x * blockDim.x / WAVE;
    for (int64_t w = wave; w < nw; w += nwave) {
        const int64_t k = w * 64 + lane_id();
        const bool on = k < range && present[k] != 0;
        const uint64_t m = __ballot(on);
        if (bits && lane_id() == 0) bits[w] = m;
        c += on;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane_id() == 0 && c) atomicAdd(count, c);
}

/* The output rows as a TD_SEL: selected probe rows (inner: with a present
 * key; present == nullptr: every key of the range is), one 1024-row
 * segment per workgroup step as k_sel_eval lays it out, with the segment
 * popcounts, flags and total.  Only selected rows look their key up. */
__global__ __launch_bounds__(256) void k_rjd_sel(tdd_col_t lk, const uint64_t* lsel, int64_t nl,
                                                 int64_t kmin, int64_t range,
                                                 const uint32_t* present, int inner,
                                                 uint32_t n_segs, uint8_t* flags, uint16_t* pop,
                                                 uint64_t* bits, unsigned long long* total) {
    __shared__ uint32_t wcnt[4];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t n_words = (nl + 63) >> 6;
    unsigned long long my_total = 0;
    for (int64_t seg = blockIdx.x; seg < n_segs; seg += gridDim.x) {
        const int64_t r0 = (seg * 16 + w * 4) * 64 + lane;
        int64_t key[4] = {0, 0, 0, 0};
        if (inner) ld4_i64(lk, r0, nl, key);
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t word = seg * 16 + w * 4 + k, row = r0 + 64 * k;
            bool on = row < nl && (!lsel || ((lsel[word] >> lane) & 1));
            if (inner && on) {
                const uint64_t i = (uint64_t)key[k] - (uint64_t)kmin;
                on = i < (uint64_t)range && (!present || ((present[i >> 5] >> (i & 31)) & 1));
            }
            const uint64_t m = __ballot(on);
            if (word < n_words && lane == 0) bits[word] = m;
            cnt += __popcll(m);
        }
        if (lane == 0) wcnt[w] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t c = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
            int64_t seg_rows = nl - seg * 1024;
            if (seg_rows > 1024) seg_rows = 1024;
            pop[seg] = (uint16_t)c;
            flags[seg] = c == 0 ? TD_SEL_NONE : (c == (uint32_t)seg_rows ? TD_SEL_ALL : TD_SEL_MIX);
            my_total += c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && my_total) atomicAdd(total, my_total);
}

/* A device TD_SEL over n rows carved out of one block of `mem`. */
static void rj_sel_view(void* mem, int64_t n, tdd_sel_t* sv) {
    const int64_t nseg = ceil_div(n, 1024);
    sv->base = mem;
    sv->flags = (uint8_t*)mem + 16;
    sv->popcnt = (uint16_t*)((uint8_t*)sv->flags + ((nseg + 7) & ~7));
    sv->bits = (uint64_t*)((uint8_t*)sv->popcnt + ((2 * nseg + 7) & ~7));
    sv->nrows = n;
    sv->n_segs = (uint32_t)nseg;
    sv->pad = 0;
}

static size_t rj_sel_bytes(int64_t n) {
    const int64_t nseg = ceil_div(n, 1024), nw = ceil_div(n, 64);
    return 16 + (size_t)((nseg + 7) & ~7) + (size_t)((2 * nseg + 7) & ~7) + (size_t)nw * 8 + 8;
}

static int rj_direct(const tdd_join_mat_spec_t* S, int64_t mn, int64_t range,
                     tdd_join_mat_result_t* out) {
    const tdd_join_spec_t* J = &S->join;
    if (S->n_lcols > TDD_MAP_MAXC || S->n_rcols > TDD_MAP_MAXC) return TD_ERR_NYI;
    for (int c = 0; c < S->n_rcols; c++) {
        const int e = S->rcols[c].esz;
        if (e != 1 && e != 2 && e != 4 && e != 8) return TD_ERR_NYI;
    }
    const int64_t pw = ceil_div(range, 64);
    SCRATCH(pbytes, (size_t)range);
    /* bits for the probe only when some key of the range is absent (a full
     * permutation of the range is answered by the range check) */
    const bool full = J->nr == range;
    Scratch present(full ? 8 : (size_t)pw * 8);
    if (!present.ok()) return TD_ERR_OOM;
    SCRATCH(count, 8);
    RJBlocks dense;
    RJDBuild B;
    memset(&B, 0, sizeof(B));
    B.key = J->rkeys[0];
    B.nr = J->nr;
    B.kmin = mn;
    B.present = pbytes.as<uint8_t>();
    B.n = S->n_rcols;
    for (int c = 0; c < S->n_rcols; c++) {
        B.src[c] = S->rcols[c].data;
        B.esz[c] = S->rcols[c].esz;
        if (!(B.dst[c] = dense.take((size_t)range * B.esz[c]))) return TD_ERR_OOM;
    }
    int rc = tdd_fill(pbytes.p, 0, (size_t)range);
    if (!rc) rc = tdd_fill(count.p, 0, 8);
    if (rc) return rc;
    if (J->nr > 0)
        LAUNCH("rjoin_map", k_rjd_build, stream_grid(ceil_div(J->nr, 256 * RJD_R)), 256, 0, B);
    LAUNCH("rjoin_map_count", k_rjd_count, stream_grid(ceil_div(pw, 256 / WAVE * 8), 4), 256, 0,
           (const uint8_t*)pbytes.as<uint8_t>(), range,
           full ? nullptr : present.as<uint64_t>(), count.as<unsigned long long>());
    int64_t distinct = 0;
    HC(hipMemcpyAsync(&distinct, count.p, 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    if (distinct != J->nr) return TD_ERR_NYI;         /* repeated build keys: pairs */
    /* every key of the range present: a range check answers for the map */
    const uint32_t* pres = full ? nullptr : present.as<uint32_t>();
    const int64_t nl = J->nl;
    const bool inner = J->join_type == 0;
    tdd_sel_t sv;
    Scratch selmem(inner || !J->lsel ? rj_sel_bytes(nl) : 8);
    if (!selmem.ok()) return TD_ERR_OOM;
    if (inner || !J->lsel) {
        rj_sel_view(selmem.p, nl, &sv);
        rc = tdd_fill(selmem.p, 0, 16);
        if (rc) return rc;
        if (sv.n_segs > 0)
            LAUNCH("rjoin_map_sel", k_rjd_sel, stream_grid(sv.n_segs, 8), 256, 0, J->lkeys[0],
                   J->lsel ? (const uint64_t*)J->lsel->bits : nullptr, nl, mn, range,
                   pres, (int)inner, sv.n_segs, sv.flags,
                   sv.popcnt, sv.bits, (unsigned long long*)sv.base);
    } else {
        sv = *J->lsel;                                /* LEFT: every selected row */
    }
    int64_t nout = 0;
    if (nl > 0) {
        rc = tdd_sel_total(&sv, &nout);
        if (rc) return rc;
    }
    POLL_CANCEL();
    RJBlocks own;
    void* ldst[TDD_MAP_MAXC];
    for (int c = 0; c < S->n_lcols; c++)
        if (!(ldst[c] = own.take((size_t)nout * S->lcols[c].esz))) return TD_ERR_OOM;
    TddMapGather M;
    memset(&M, 0, sizeof(M));
    M.key = J->lkeys[0];
    M.kmin = mn;
    M.range = range;
    M.present = pres;
    M.n = S->n_rcols;
    for (int c = 0; c < S->n_rcols; c++) {
        M.src[c] = B.dst[c];
        M.esz[c] = B.esz[c];
        if (!(M.dst[c] = own.take((size_t)nout * M.esz[c]))) return TD_ERR_OOM;
    }
    if (nout > 0) {
        rc = tdd__sel_compact_map(&sv, S->n_lcols, S->lcols, ldst, &M);
        if (rc) return rc;
    }
    for (int c = 0; c < S->n_lcols; c++) { own.keep(ldst[c]); out->lcols[c] = ldst[c]; }
    for (int c = 0; c < S->n_rcols; c++) { own.keep(M.dst[c]); out->rcols[c] = M.dst[c]; }
    out->n_rows = nout;
    out->path = 4;
    out->radix_bits = 0;
    return TD_OK;
}

extern "C" int tdd_join_mat(const tdd_join_mat_spec_t* S, tdd_join_mat_result_t* out) {
    memset(out, 0, sizeof(*out));
    const tdd_join_spec_t* J = &S->join;
    if (J->join_type > 1 || J->n_keys < 1 || J->n_keys > TDD_MAX_KEYS) return TD_ERR_NYI;
    if (J->force_path >= 3) return TD_ERR_NYI;
    if (S->n_lcols < 0 || S->n_lcols > RJ_MAXC || S->n_rcols < 0 || S->n_rcols > RJ_MAXC)
        return TD_ERR_NYI;
    if (J->nl >= ((int64_t)1 << 31) || J->nr >= ((int64_t)1 << 31)) return TD_ERR_NYI;
    for (int k = 0; k < J->n_keys; k++)
        if ((J->lkeys[k].type == TD_F64) != (J->rkeys[k].type == TD_F64)) return TD_ERR_NYI;
    const bool exact = J->n_keys == 1 && rj_is_int(J->lkeys[0].type) && rj_is_int(J->rkeys[0].type);
    if (J->force_path == 1 && !exact) return TD_ERR_NYI;
    if (J->force_path != 2 && exact) {
        /* dense build keys: the map (forced: any range a map can span) */
        int64_t mn = 0, mx = -1;
        if (J->nr > 0 && J->rkeys[0].ranged) { mn = J->rkeys[0].lo; mx = J->rkeys[0].hi; }
        else if (J->nr > 0) {
            int rc = tdd_minmax(&J->rkeys[0], J->nr, nullptr, &mn, &mx);
            if (rc) return rc;
        }
        if (mn > mx) { mn = 0; mx = 0; }                  /* no build rows */
        const uint64_t range = (uint64_t)mx - (uint64_t)mn + 1;
        if (range < ((uint64_t)1 << 31) &&
            (J->force_path == 1 || range <= (uint64_t)(4 * J->nr + 4096)))
            return rj_direct(S, mn, (int64_t)range, out);
        if (J->force_path == 1) return TD_ERR_NYI;
    }

    /* radix bits: build partitions of about 0.6 of an LDS table, and enough
     * partitions to spread the probe side over the chip */
    int bits = J->radix_bits;
    if (bits <= 0) {
        int bb = 0, bp = 0;
        while (bb < RJ_MAXB && ((int64_t)RJ_PART_ROWS << bb) < J->nr) bb++;
        while (bp < 10 && ((int64_t)32768 << bp) < J->nl) bp++;
        bits = bb > bp ? bb : bp;
    }
    if (bits > RJ_MAXB) bits = RJ_MAXB;
    if (bits < 1) bits = 1;
    const int P = 1 << bits;

    RJBlocks own;                 /* output blocks until handed over */
    /* probe side: every output left column, plus the key word when the key
     * is hashed or is not among the carried columns, plus rows to verify */
    RJHostSide L, R;
    L.K.n = R.K.n = J->n_keys;
    L.K.exact = R.K.exact = exact;
    for (int k = 0; k < J->n_keys; k++) { L.K.cols[k] = J->lkeys[k]; R.K.cols[k] = J->rkeys[k]; }
    L.sel = J->lsel ? J->lsel->bits : nullptr;
    L.n = J->nl;
    R.n = J->nr;
    L.plain = rj_plain(L.K);
    R.plain = rj_plain(R.K);
    L.ncols = S->n_lcols;
    for (int c = 0; c < S->n_lcols; c++) {
        L.src[c] = S->lcols[c].data;
        L.esz[c] = S->lcols[c].esz;
        if (exact && L.key_col < 0 && S->lcols[c].data == J->lkeys[0].data &&
            S->lcols[c].type == J->lkeys[0].type && S->lcols[c].attrs == J->lkeys[0].attrs)
            L.key_col = c;
    }
    L.carry_kw = L.key_col < 0;
    L.carry_row = R.carry_row = !exact;
    R.carry_kw = true;
    R.ncols = S->n_rcols;
    for (int c = 0; c < S->n_rcols; c++) { R.src[c] = S->rcols[c].data; R.esz[c] = S->rcols[c].esz; }
    rj_spans(L);
    rj_spans(R);

    /* counts of both sides; the plans, passes and build + probe follow on
     * the device, and the host reads the row totals with the match count at
     * the end.  Buffers are sized for every selected row. */
    SCRATCH(lc, (size_t)P * RJ_XCD * 4);
    SCRATCH(rcn, (size_t)P * RJ_XCD * 4);
    SCRATCH(stat, 64);          /* matched, repeated key, oversized, left rows, right rows */
    int rc = tdd_fill(stat.p, 0, 64);
    if (!rc) rc = rj_hist(L, bits, lc);
    if (!rc) rc = rj_hist(R, bits, rcn);
    if (rc) return rc;
    unsigned long long* st_dev = stat.as<unsigned long long>();
    const int64_t nlp = L.n, nrp = R.n;         /* bounds; the totals come back with the stat */
    POLL_CANCEL();
    void* ldst[RJ_MAXC] = {};
    void* rbuf[RJ_MAXC] = {};
    struct Frees { void** p; int n; ~Frees() { for (int i = 0; i < n; i++) td__arena_free(p[i]); } }
        rfree{rbuf, RJ_MAXC};
    for (int c = 0; c < L.ncols; c++)
        if (!(ldst[c] = own.take((size_t)nlp * L.esz[c]))) return TD_ERR_OOM;
    for (int c = 0; c < R.ncols; c++)
        if (!(rbuf[c] = td__arena_alloc((size_t)(nrp > 0 ? nrp : 1) * R.esz[c])))
            return TD_ERR_OOM;
    const size_t tiles_cap = (size_t)(ceil_div(nlp > nrp ? nlp : nrp, RJ_STEP) + P + 8) * 16;
    SCRATCH(loff, ((size_t)P + 1) * 8);
    SCRATCH(roff, ((size_t)P + 1) * 8);
    SCRATCH(lcur1, (size_t)P * RJ_XCD * 8);
    SCRATCH(rcur1, (size_t)P * RJ_XCD * 8);
    SCRATCH(lcur2, (size_t)P * 8);
    SCRATCH(rcur2, (size_t)P * 8);
    SCRATCH(ltiles, tiles_cap);
    SCRATCH(rtiles, tiles_cap);
    SCRATCH(lxt, (RJ_XCD + 1) * 4);
    SCRATCH(rxt, (RJ_XCD + 1) * 4);
    Scratch lmid(8), rmid(8), lkw(8), lkwm(8), rkw(8), rkwm(8), lrow(8), lrowm(8), rrow(8), rrowm(8);
    RJPlans PS;
    PS.side[0] = rj_plan_of(bits, lc, loff, lcur1, lcur2, ltiles, lxt, st_dev + 3);
    PS.side[1] = rj_plan_of(bits, rcn, roff, rcur1, rcur2, rtiles, rxt, st_dev + 4);
    LAUNCH("rjoin_plan", k_rj_plan, 2, RJP_T, 0, PS);
    rc = rj_partition(L, bits, ldst, lcur1, lcur2, ltiles, lxt, lmid, lkw, lkwm, lrow, lrowm);
    if (!rc) rc = rj_partition(R, bits, rbuf, rcur1, rcur2, rtiles, rxt, rmid, rkw, rkwm, rrow, rrowm);
    if (rc) return rc;
    POLL_CANCEL();

    RJBP A;
    memset(&A, 0, sizeof(A));
    A.loff = loff.as<int64_t>();
    A.roff = roff.as<int64_t>();
    int plain = 0;
    if (L.key_col >= 0) {
        A.lkey = S->lcols[L.key_col];
        A.lkey.data = L.dst[L.key_col];
        plain = L.plain;
    } else {
        A.lkw = L.kw;
    }
    A.rkw = R.kw;
    A.lrow = L.row;
    A.rrow = R.row;
    A.LK = L.K;
    A.RK = R.K;
    A.exact = exact;
    A.left_join = J->join_type == 1;
    A.nrc = R.ncols;
    for (int c = 0; c < R.ncols; c++) {
        A.rsrc[c] = rbuf[c];
        A.resz[c] = R.esz[c];
        if (!(A.rdst[c] = own.take((size_t)nlp * R.esz[c]))) return TD_ERR_OOM;
    }
    SCRATCH(flag, (size_t)(J->join_type == 0 ? nlp : 1));
    if (J->join_type == 0) A.lflag = flag.as<uint8_t>();
    A.stat = st_dev;
    const size_t lds = (size_t)RJ_SLOTS * 2 + (size_t)RJ_CAP * 16;
    if (plain == 8) LAUNCH("rjoin_build_probe", k_rj_bp<8>, P, RJ_BT, lds, A);
    else if (plain == 4) LAUNCH("rjoin_build_probe", k_rj_bp<4>, P, RJ_BT, lds, A);
    else LAUNCH("rjoin_build_probe", k_rj_bp<0>, P, RJ_BT, lds, A);
    unsigned long long st[5];
    HC(hipMemcpyAsync(st, stat.p, sizeof(st), hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    if (st[1] || st[2]) return TD_ERR_NYI;      /* repeated build keys / oversized partition */
    const int64_t nl_rows = (int64_t)st[3];     /* probe rows taking part */

    int64_t nout = J->join_type == 1 ? nl_rows : (int64_t)st[0];
    void* lo[RJ_MAXC] = {};
    void* ro[RJ_MAXC] = {};
    for (int c = 0; c < L.ncols; c++) lo[c] = L.dst[c];
    for (int c = 0; c < R.ncols; c++) ro[c] = A.rdst[c];
    if (J->join_type == 0 && nout < nl_rows) {
        /* inner join with probe rows that found no partner: compact the
         * slotted outputs through the match flags */
        const int64_t nseg = ceil_div(nl_rows, 1024), nw = ceil_div(nl_rows, 64);
        SCRATCH(selmem, 16 + (size_t)((nseg + 7) & ~7) + (size_t)((2 * nseg + 7) & ~7) +
                            (size_t)nw * 8);
        tdd_sel_t sv;
        sv.base = selmem.p;
        sv.flags = (uint8_t*)selmem.p + 16;
        sv.popcnt = (uint16_t*)((uint8_t*)sv.flags + ((nseg + 7) & ~7));
        sv.bits = (uint64_t*)((uint8_t*)sv.popcnt + ((2 * nseg + 7) & ~7));
        sv.nrows = nl_rows;
        sv.n_segs = (uint32_t)nseg;
        sv.pad = 0;
        rc = tdd_sel_from_bool(flag.as<uint8_t>(), &sv);
        if (rc) return rc;
        tdd_col_t src[2 * RJ_MAXC];
        void* dst[2 * RJ_MAXC];
        int nc = 0;
        memset(src, 0, sizeof(src));
        for (int c = 0; c < L.ncols; c++, nc++) {
            src[nc].data = L.dst[c];
            src[nc].esz = L.esz[c];
            if (!(dst[nc] = own.take((size_t)nout * L.esz[c]))) return TD_ERR_OOM;
            lo[c] = dst[nc];
        }
        for (int c = 0; c < R.ncols; c++, nc++) {
            src[nc].data = A.rdst[c];
            src[nc].esz = R.esz[c];
            if (!(dst[nc] = own.take((size_t)nout * R.esz[c]))) return TD_ERR_OOM;
            ro[c] = dst[nc];
        }
        for (int c0 = 0; c0 < nc; c0 += RJ_MAXC) {
            rc = tdd_sel_compact(&sv, nc - c0 < RJ_MAXC ? nc - c0 : RJ_MAXC, src + c0, dst + c0);
            if (rc) return rc;
        }
    }
    for (int c = 0; c < L.ncols; c++) { own.keep(lo[c]); out->lcols[c] = lo[c]; }
    for (int c = 0; c < R.ncols; c++) { own.keep(ro[c]); out->rcols[c] = ro[c]; }
    out->n_rows = nout;
    out->path = 3;
    out->radix_bits = bits;
    return TD_OK;
}
