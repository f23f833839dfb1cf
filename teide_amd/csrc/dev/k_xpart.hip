/*
 * k_xpart.hip -- XCD-partitioned fused probe for the star shape of
 * filter -> join -> group-by with a large dense build side.
 *
 * The single-pass star probe gathers one slot-map entry per passing fact
 * row.  With a 16 M-key build side the u16 map is 32 MB: every gather
 * misses the 4 MB L2 of its XCD and moves a whole cache line across the
 * fabric, so the fabric, not HBM, bounds the probe.  Here the probe side is
 * split by key range instead, so that each XCD's L2 holds the map slice its
 * rows gather from:
 *
 *   k_xp_fact   one pass over the fact rows: star predicate, then every
 *               passing row is written as a packed record -- a 32-bit word
 *               (key offset within its partition, plus integer aggregate
 *               inputs narrowed by their column statistics) and one 64-bit
 *               word per remaining aggregate input -- into one of <= 16
 *               key-range partitions.  Ranks within a wave come from
 *               ballots; a workgroup reserves each partition's run once per
 *               round with one global atomic, so each partition's records
 *               leave as contiguous runs.
 *   k_xp_probe  partition p is read by the workgroups of XCD p mod 8
 *               (workgroups are dispatched to the 8 XCDs round-robin by
 *               workgroup id), which gather from the 2 MB slice of the map
 *               that partition covers and fold into an LDS group table;
 *               per-workgroup partials are then folded in workgroup order.
 *
 * Measured on MI355X (fact 200 M x dim 16 M, 1,000 groups): the fact pass
 * moves 4.8 GB read + 1.2 GB written in 1.43 ms and the probe 1.2 GB in
 * 0.61 ms, where the gathers now hit L2 but still move one 128-byte line
 * per 2-byte entry between L2 and L1 -- level with the single-pass probe.
 *
 * Records of a partition carry no row order, so f64 sums add in an order
 * that differs between runs, as the LDS atomics of the single-pass probe
 * already do.  A partition whose records outgrow its region (key skew)
 * raises a flag and the caller falls back to the single-pass probe.
 */
#include "dev_common.h"
#include "group_plan.h"
#include "star.h"
#include "tile.h"

#define XP_T     512            /* threads per workgroup (probe) */
#define XP_MAXP  16             /* partitions */
#define XP_MAXV  3              /* 64-bit value words per record */

/* How an aggregate input (tile slot q) travels in a record. */
struct XPCarry {
    int8_t  kind[TL_NC];        /* 0 not carried, 1 packed in the word, 2 value word */
    int8_t  idx[TL_NC];         /* value word index (kind 2) */
    uint8_t shift[TL_NC];       /* bit offset in the word (kind 1) */
    uint8_t width[TL_NC];
    int64_t base[TL_NC];        /* packed value = column value - base */
};

struct XPFact {
    TileCols        T;
    StarPred        SP;
    XPCarry         C;
    const uint64_t* lsel;
    int64_t         nl, kmin, range;
    int32_t         s, P;       /* partition = key offset >> s */
    int64_t         cap;        /* records per partition region */
    uint32_t*       words;      /* [P][cap] */
    uint64_t*       vals[XP_MAXV];
    uint32_t*       gcount;     /* [P] records reserved */
    int32_t*        overflow;
};

/* PF: prefetch the next round's tile into registers (one workgroup per CU)
 * or rely on two resident workgroups per CU instead. */
template <int NPRED, int NV, bool PF, int R, int T>
__global__ __launch_bounds__(T, PF ? 2 : 4) void k_xp_fact(XPFact A) {
    constexpr int ROWS = T * R;
    constexpr int WAVES = T / WAVE;
    __shared__ uint32_t wcnt[WAVES][XP_MAXP];
    __shared__ uint32_t woff[WAVES][XP_MAXP];
    __shared__ int64_t  gbase[XP_MAXP];
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int64_t n_rounds = (A.nl + ROWS - 1) / ROWS;
    const int64_t lofs = (int64_t)wv * 64 * R + (int64_t)lane * R;
    const uint32_t lmask = A.s >= 32 ? 0xffffffffu : (1u << A.s) - 1u;
    bool over = false;
    int64_t cv[TL_NC][R], nx[TL_NC][R];
    int64_t r = blockIdx.x;
    if (PF && r < n_rounds) load_tile<R>(A.T, r * ROWS + lofs, A.nl, cv);
    for (; r < n_rounds; r += gridDim.x) {
        const int64_t base = r * ROWS + lofs;
        const int64_t rn = r + gridDim.x;
        if (PF) {
            if (rn < n_rounds) load_tile<R>(A.T, rn * ROWS + lofs, A.nl, nx);   /* prefetch */
        } else {
            load_tile<R>(A.T, base, A.nl, cv);
        }
        const uint32_t m = tile_live<R>(A.lsel, base, A.nl) & star_pred<R, NPRED>(A.SP, A.T, cv);
        int32_t p[R];
        uint32_t w[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            const uint64_t k = (uint64_t)(cv[0][j] - A.kmin);
            const bool live = ((m >> j) & 1) && k < (uint64_t)A.range;
            p[j] = live ? (int32_t)(k >> A.s) : -1;
            w[j] = (uint32_t)k & lmask;
        }
#pragma unroll
        for (int q = 1; q < TL_NC; q++) {
            if (A.C.kind[q] != 1) continue;
#pragma unroll
            for (int j = 0; j < R; j++)
                w[j] |= (uint32_t)(cv[q][j] - A.C.base[q]) << A.C.shift[q];
        }
        /* rank of each row among the wave's rows of its partition */
        int32_t rk[R];
        uint32_t wc[XP_MAXP];
#pragma unroll
        for (int q = 0; q < XP_MAXP; q++) wc[q] = 0;
#pragma unroll
        for (int j = 0; j < R; j++) {
            rk[j] = 0;
#pragma unroll
            for (int q = 0; q < XP_MAXP; q++) {
                const uint64_t b = __ballot(p[j] == q);
                if (p[j] == q) rk[j] = (int32_t)(lanes_below(b) + wc[q]);
                wc[q] += (uint32_t)__popcll(b);
            }
        }
#pragma unroll
        for (int q = 0; q < XP_MAXP; q++)
            if (lane == q) wcnt[wv][q] = wc[q];
        __syncthreads();
        if (threadIdx.x < XP_MAXP) {
            const int q = threadIdx.x;
            uint32_t run = 0;
#pragma unroll
            for (int v = 0; v < WAVES; v++) {
                woff[v][q] = run;
                run += wcnt[v][q];
            }
            int64_t g = 0;
            if (run && q < A.P) {
                g = (int64_t)atomicAdd(&A.gcount[q], run);
                if (g + run > A.cap) over = true;
            }
            gbase[q] = (int64_t)q * A.cap + g;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < R; j++) {
            if (p[j] < 0) continue;
            const int64_t off = gbase[p[j]] - (int64_t)p[j] * A.cap + woff[wv][p[j]] + rk[j];
            if (off >= A.cap) continue;                       /* overflow: flagged */
            const int64_t at = (int64_t)p[j] * A.cap + off;
            A.words[at] = w[j];
#pragma unroll
            for (int q = 1; q < TL_NC; q++) {
                if (A.C.kind[q] != 2) continue;
                switch (A.C.idx[q]) {                          /* compile-time value word */
                case 0: if (NV > 0) A.vals[0][at] = (uint64_t)cv[q][j]; break;
                case 1: if (NV > 1) A.vals[1][at] = (uint64_t)cv[q][j]; break;
                default: if (NV > 2) A.vals[2][at] = (uint64_t)cv[q][j]; break;
                }
            }
        }
        if (PF && rn < n_rounds) {
#pragma unroll
            for (int c = 0; c < TL_NC; c++)
#pragma unroll
                for (int j = 0; j < R; j++) cv[c][j] = nx[c][j];
        }
    }
    if (over) *A.overflow = 1;
}

struct XPProbe {
    TileCols        T;          /* types of the aggregate inputs by tile slot */
    XPCarry         C;
    const uint32_t* words;
    const uint64_t* vals[XP_MAXV];
    const uint32_t* gcount;
    int64_t         cap;
    int32_t         s, P;
    const uint16_t* map;        /* slot of each build key offset, 0xffff: none */
    int64_t         n_slots;
    uint8_t         cell_slot[GP_MAX_CELLS];
};

/* One tile of records: R per lane (words, then NV value words). */
template <int R, int NV>
struct XPRec {
    uint32_t w[R];
    int64_t  v[NV > 0 ? NV : 1][R];
};

template <int R, int NV>
__device__ __forceinline__ void xp_load(const XPProbe& A, int64_t reg, int64_t i0, int64_t n,
                                        XPRec<R, NV>& r) {
    if (i0 + R <= n) {
#pragma unroll
        for (int k = 0; k < R / 4; k++) {
            const tl_v4i32 x = __builtin_nontemporal_load((const tl_v4i32*)(A.words + reg + i0) + k);
            r.w[4 * k] = (uint32_t)x.x; r.w[4 * k + 1] = (uint32_t)x.y;
            r.w[4 * k + 2] = (uint32_t)x.z; r.w[4 * k + 3] = (uint32_t)x.w;
        }
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int k = 0; k < R / 2; k++) {
                const tl_v2i64 x = __builtin_nontemporal_load((const tl_v2i64*)(A.vals[v] + reg + i0) + k);
                r.v[v][2 * k] = x.x;
                r.v[v][2 * k + 1] = x.y;
            }
    } else {
#pragma unroll
        for (int j = 0; j < R; j++) r.w[j] = i0 + j < n ? A.words[reg + i0 + j] : 0u;
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int j = 0; j < R; j++) r.v[v][j] = i0 + j < n ? (int64_t)A.vals[v][reg + i0 + j] : 0;
    }
}

/* Next (partition, tile) of workgroup wi of wn on its XCD, or p = -1. */
__device__ __forceinline__ void xp_next(const XPProbe& A, int R, int& p, int64_t& t, int64_t wi,
                                        int64_t wn) {
    t += wn;
    for (;;) {
        int64_t n = A.gcount[p];
        if (n > A.cap) n = A.cap;
        if (t * XP_T * R < n) return;
        p += 8;
        t = wi;
        if (p >= A.P) { p = -1; return; }
    }
}

template <int R, int NV, bool PF>
__global__ __launch_bounds__(XP_T) void k_xp_probe(XPProbe A, GroupPlan P, uint64_t* partials) {
    extern __shared__ uint64_t tab[];
    const int64_t ncell_total = A.n_slots * P.ncells;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x)
        tab[i] = cell_identity(P.kind[i % P.ncells]);
    __syncthreads();
    const int64_t wi = blockIdx.x >> 3, wn = gridDim.x >> 3;
    const uint32_t lmask = A.s >= 32 ? 0xffffffffu : (1u << A.s) - 1u;
    int p = blockIdx.x & 7;
    int64_t t = wi - wn;
    if (p < A.P) xp_next(A, R, p, t, wi, wn);
    else p = -1;
    XPRec<R, NV> cur, nxt;
    if (PF && p >= 0) {
        int64_t n = A.gcount[p];
        if (n > A.cap) n = A.cap;
        xp_load<R, NV>(A, (int64_t)p * A.cap, t * XP_T * R + (int64_t)threadIdx.x * R, n, cur);
    }
    while (p >= 0) {
        int64_t n = A.gcount[p];
        if (n > A.cap) n = A.cap;
        const int64_t reg = (int64_t)p * A.cap;
        const int64_t i0 = t * XP_T * R + (int64_t)threadIdx.x * R;
        int pn = p;
        int64_t tn = t;
        xp_next(A, R, pn, tn, wi, wn);
        if (PF) {
            if (pn >= 0) {
                int64_t nn = A.gcount[pn];
                if (nn > A.cap) nn = A.cap;
                xp_load<R, NV>(A, (int64_t)pn * A.cap, tn * XP_T * R + (int64_t)threadIdx.x * R, nn, nxt);
            }
        } else {
            xp_load<R, NV>(A, reg, i0, n, cur);
        }
        int64_t cv[TL_NC][R];
#pragma unroll
        for (int q = 0; q < TL_NC; q++) {
            if (A.C.kind[q] == 2) {
                switch (A.C.idx[q]) {                          /* compile-time value word */
                case 0:
#pragma unroll
                    for (int j = 0; j < R; j++) cv[q][j] = cur.v[0][j];
                    break;
                case 1:
#pragma unroll
                    for (int j = 0; j < R; j++) cv[q][j] = cur.v[NV > 1 ? 1 : 0][j];
                    break;
                default:
#pragma unroll
                    for (int j = 0; j < R; j++) cv[q][j] = cur.v[NV > 2 ? 2 : 0][j];
                    break;
                }
            } else if (A.C.kind[q] == 1) {
                const uint32_t msk = A.C.width[q] >= 32 ? 0xffffffffu : (1u << A.C.width[q]) - 1u;
#pragma unroll
                for (int j = 0; j < R; j++)
                    cv[q][j] = (int64_t)((cur.w[j] >> A.C.shift[q]) & msk) + A.C.base[q];
            } else {
#pragma unroll
                for (int j = 0; j < R; j++) cv[q][j] = 0;
            }
        }
        const uint16_t* map = A.map + ((int64_t)p << A.s);
        int32_t sl[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            sl[j] = -1;
            if (i0 + j < n) {
                const uint16_t v = map[cur.w[j] & lmask];
                sl[j] = v == 0xffff ? -1 : (int32_t)v;
            }
        }
        acc_tile<R>(tab, P, A.cell_slot, A.T, cv, sl);
        if (PF) cur = nxt;
        p = pn;
        t = tn;
    }
    __syncthreads();
    uint64_t* dst = partials + (int64_t)blockIdx.x * ncell_total;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x) dst[i] = tab[i];
}

static int bits_for(uint64_t span) {      /* bits to hold 0 .. span */
    int b = 0;
    while (b < 64 && (span >> b)) b++;
    return b;
}

/* The XCD-partitioned probe: TD_ERR_NYI when the shape does not fit (the
 * caller then runs the single-pass probe); on key skew that overflows a
 * partition, TD_ERR_NYI after the passes (results discarded). */
int jg_xpart(const tdd_join_group_spec_t* S, const GroupPlan& GP, const TileCols& T0,
             const StarPred& SP, const uint8_t* cell_slot, const uint16_t* map, int64_t kmin,
             int64_t range, int64_t n_slots, tdd_group_result_t* out) {
    const tdd_join_spec_t* J = &S->join;
    const size_t table_bytes = (size_t)n_slots * GP.ncells * 8;
    if (table_bytes > 64 * 1024 || J->nl <= 0) return TD_ERR_NYI;
    int s = 0;
    while (s < 40 && ((range - 1) >> s) >= XP_MAXP) s++;
    const int P = (int)(((range - 1) >> s) + 1);
    if (s > 24) return TD_ERR_NYI;               /* map slices beyond 32 MB */

    /* carry plan: which tile slots the cells read, packed or as value words */
    XPCarry C;
    memset(&C, 0, sizeof(C));
    bool used[TL_NC] = {false, false, false, false};
    for (int i = 1; i < GP.ncells; i++) used[cell_slot[i]] = true;
    if (used[0]) return TD_ERR_NYI;              /* aggregate over the probe key */
    int bits = s, nv = 0;
    for (int q = 1; q < TL_NC; q++) {
        if (!used[q]) continue;
        const tdd_col_t& c = T0.col[q];
        const bool is_i = c.type != TD_F64;
        if (is_i && c.ranged && c.hi >= c.lo) {
            const int wb = bits_for((uint64_t)c.hi - (uint64_t)c.lo);
            if (bits + wb <= 32) {
                C.kind[q] = 1;
                C.shift[q] = (uint8_t)bits;
                C.width[q] = (uint8_t)wb;
                C.base[q] = c.lo;
                bits += wb;
                continue;
            }
        }
        if (nv >= XP_MAXV) return TD_ERR_NYI;
        C.kind[q] = 2;
        C.idx[q] = (int8_t)nv++;
    }

    /* regions: every partition can take its share of all rows + 15 % + a round */
    const int64_t per = ceil_div(J->nl, P);
    int64_t cap = per + per / 7 + 8192 * 4;
    cap = ceil_div(cap, 4) * 4;
    SCRATCH(words, (size_t)P * cap * 4);
    Scratch vbuf((size_t)(nv > 0 ? nv : 1) * P * cap * 8);
    if (!vbuf.ok()) return TD_ERR_OOM;
    SCRATCH(ctr, 256);
    int rc = tdd_fill(ctr.p, 0, 256);
    if (rc) return rc;

    XPFact F;
    memset(&F, 0, sizeof(F));
    F.T = T0;
    F.SP = SP;
    F.C = C;
    F.lsel = J->lsel ? J->lsel->bits : nullptr;
    F.nl = J->nl;
    F.kmin = kmin;
    F.range = range;
    F.s = s;
    F.P = P;
    F.cap = cap;
    F.words = words.as<uint32_t>();
    for (int v = 0; v < nv; v++) F.vals[v] = vbuf.as<uint64_t>() + (int64_t)v * P * cap;
    F.gcount = ctr.as<uint32_t>();
    F.overflow = (int32_t*)(ctr.as<uint8_t>() + 128);

    const int fv = TUNE_INT("TEIDE_XP_FACT", 0);      /* tile shape of the fact pass */
#define XPF1(NP, NV_, PF_, R_, T_)                                                          \
    LAUNCH("jg_xp_fact", (k_xp_fact<NP, NV_, PF_, R_, T_>),                                 \
           resident_grid(k_xp_fact<NP, NV_, PF_, R_, T_>, T_, 0,                            \
                         ceil_div(J->nl, (int64_t)(R_) * (T_))), T_, 0, F)
#define XPF(NP, NV_)                                                                        \
    do {                                                                                    \
        if (fv == 1) XPF1(NP, NV_, false, 8, 256);                                          \
        else if (fv == 2) XPF1(NP, NV_, true, 4, 512);                                      \
        else XPF1(NP, NV_, false, 4, 512);                                                  \
    } while (0)
#define XPF_NV(NP)                        \
    switch (nv) {                         \
    case 0: XPF(NP, 0); break;            \
    case 1: XPF(NP, 1); break;            \
    case 2: XPF(NP, 2); break;            \
    default: XPF(NP, 3); break;           \
    }
    switch (SP.npred) {
    case 0: XPF_NV(0); break;
    case 1: XPF_NV(1); break;
    default: XPF_NV(2); break;
    }
#undef XPF_NV
#undef XPF
#undef XPF1

    XPProbe B;
    memset(&B, 0, sizeof(B));
    B.T = T0;
    B.C = C;
    B.words = words.as<uint32_t>();
    for (int v = 0; v < nv; v++) B.vals[v] = F.vals[v];
    B.gcount = F.gcount;
    B.cap = cap;
    B.s = s;
    B.P = P;
    B.map = map;
    B.n_slots = n_slots;
    memcpy(B.cell_slot, cell_slot, sizeof(B.cell_slot));
    int per_cu = (int)(160 * 1024 / (table_bytes + 1024));
    if (per_cu > 2) per_cu = 2;
    if (per_cu < 1) per_cu = 1;
    const int grid = (tdd::g_num_cu * per_cu) & ~7;
    SCRATCH(parts, (size_t)grid * table_bytes);
    const int pr = TUNE_INT("TEIDE_XP_PROBE_R", 4) == 8 ? 8 : 4;
    const bool ppf = TUNE_INT("TEIDE_XP_PROBE_PF", 1) != 0;
    uint64_t* pp = parts.as<uint64_t>();
#define XPP(R_, NV_)                                                                         \
    do {                                                                                     \
        if (ppf) LAUNCH("jg_xp_probe", (k_xp_probe<R_, NV_, true>), grid, XP_T, table_bytes, B, GP, pp); \
        else LAUNCH("jg_xp_probe", (k_xp_probe<R_, NV_, false>), grid, XP_T, table_bytes, B, GP, pp);    \
    } while (0)
#define XPP_NV(R_)                        \
    switch (nv) {                         \
    case 0: XPP(R_, 0); break;            \
    case 1: XPP(R_, 1); break;            \
    case 2: XPP(R_, 2); break;            \
    default: XPP(R_, 3); break;           \
    }
    if (pr == 8) XPP_NV(8) else XPP_NV(4)
#undef XPP_NV
#undef XPP
    int32_t flag = 0;
    HC(hipMemcpyAsync(&flag, F.overflow, 4, hipMemcpyDeviceToHost, tdd::g_stream));
    rc = da_finish(GP, parts.as<uint64_t>(), grid, n_slots, out);
    if (rc) return rc;
    HC(hipStreamSynchronize(tdd::g_stream));
    if (flag) {                                   /* key skew: results incomplete */
        for (int k = 0; k < TDD_MAX_KEYS; k++) td__arena_free(out->keys[k]);
        for (int a = 0; a < TDD_MAX_AGGS; a++) td__arena_free(out->vals[a]);
        memset(out, 0, sizeof(*out));
        return TD_ERR_NYI;
    }
    out->path = 4;
    return TD_OK;
}
