/*
 * k_part.hip -- radix-partitioned fused join -> group-by.
 *
 * For a large build side the direct slot map (key -> group slot) is far
 * bigger than an XCD's 4 MB L2, so every probe gather pulls a 128-byte line
 * across the fabric for 2 useful bytes.  This path instead partitions both
 * sides on the high bits of the join key, with the slice size chosen so one
 * slice of the slot map (u16 per key) plus the group table fits the 160 KB
 * of LDS of one CU:
 *
 *   A  build side   per 16K-row chunk: partition histogram, then a counting
 *                   sort in LDS and one contiguous copy per partition of
 *                   (key offset | group slot) words;
 *   B  probe side   one streaming pass over the probe columns: predicate,
 *                   key -> partition, and a packed row (key offset, small
 *                   integers bit-packed from column bounds, other aggregate
 *                   inputs as 8-byte values) staged in LDS per partition and
 *                   flushed each round as one contiguous run into the
 *                   workgroup's region of that partition (write-combined:
 *                   a few lines per partition instead of a line per row).
 *                   Cursors live in LDS; no global atomics.  Rounds are
 *                   interleaved over workgroups so every workgroup sees the
 *                   whole key distribution and its regions fill evenly;
 *   C  per partition one workgroup builds the slice of the slot map in LDS
 *                   from the partition's build words, checks that the build
 *                   keys are unique, then streams the partition's rows and
 *                   folds them into its LDS group table.
 *
 * HBM traffic: the probe columns once, plus the packed passing rows written
 * and read once.  Skewed keys that overflow a region, and duplicate build
 * keys, report TD_ERR_NYI and the caller runs the direct-map path.
 */
#include "group_plan.h"
#include "tile.h"
#include "star.h"

#define PB_T 512          /* pass B workgroup */
#define PB_R 4            /* rows per lane per round, pass B */
#define PC_T 1024         /* pass C workgroup */
#define PC_U 4            /* rows per thread per step, pass C */
#define PA_T 256          /* pass A workgroup */
#define PA_CHUNK 16384    /* build rows per pass A workgroup */
#define PART_MAX_VALS 3
#define PART_MAX_PACK 4
#define PART_MAX_W 1024

struct PackField {
    int32_t word;         /* 0: bits above the key offset in w0, 1: w1 */
    int32_t shift, bits;
    int32_t col;          /* tile slot of the source column */
    int64_t base;         /* stored value = v - base */
};

struct PartFact {
    TileCols   T;                      /* slot 0 key, 1..npred predicate columns */
    StarPred   SP;
    const uint64_t* lsel;
    int64_t    nl;
    int64_t    kmin, range;
    int32_t    s;                      /* slice bits */
    int32_t    P;                      /* partitions */
    int32_t    nv;                     /* 8-byte values per row */
    int32_t    npack;
    int32_t    val_col[PART_MAX_VALS]; /* tile slot of each value */
    PackField  pack[PART_MAX_PACK];
    int32_t    pad[2];
    int64_t    cap;                    /* rows per (partition, workgroup) region */
    uint64_t*  rows;                   /* [P][W][cap][1 + nv] u64 */
    int32_t*   counts;                 /* [P][W] */
    int*       overflow;
};

/* ------------------------------------------------------------- pass A -- */

__device__ __forceinline__ uint32_t part_slot(const GroupPlan& P, int64_t r) {
    int64_t s = 0;
    for (int k = 0; k < P.n_keys; k++) s += (ld_i64(P.keys[k], r) - P.kmin[k]) * P.kstride[k];
    return (uint32_t)s;
}

/* Histogram of build rows per partition for each PA_CHUNK-row chunk,
 * written partition-major (hist[p * G + g]) so one scan gives every
 * (partition, chunk) its output offset. */
__global__ __launch_bounds__(PA_T) void k_pa_hist(tdd_col_t rk, int vec, int64_t nr, int64_t kmin,
                                                  int s, int P, uint32_t* hist) {
    extern __shared__ uint32_t h[];
    for (int i = threadIdx.x; i < P; i += PA_T) h[i] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * PA_CHUNK;
    const int64_t r1 = r0 + PA_CHUNK < nr ? r0 + PA_CHUNK : nr;
    for (int64_t b = r0 + (int64_t)threadIdx.x * 8; b < r1; b += PA_T * 8) {
        int64_t v[8];
        load_rows<8>(rk, b, r1, vec != 0, v);
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (b + j < r1) atomicAdd(&h[(v[j] - kmin) >> s], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += PA_T) hist[(int64_t)i * gridDim.x + blockIdx.x] = h[i];
}

/* Build words (key offset within the slice, low 16 bits | group slot) of a
 * chunk, counting-sorted by partition in LDS, then copied out as one
 * contiguous run per partition. */
__global__ __launch_bounds__(PA_T) void k_pa_scatter(tdd_col_t rk, int64_t nr, int64_t kmin,
                                                     int s, int P, GroupPlan GP,
                                                     const uint32_t* hist, const int64_t* pos,
                                                     uint32_t* words) {
    extern __shared__ uint32_t lds[];
    uint32_t* buf = lds;                       /* PA_CHUNK */
    uint32_t* lstart = buf + PA_CHUNK;         /* P + 1 */
    uint32_t* cur = lstart + P + 1;            /* P */
    __shared__ uint32_t wsum[PA_T / WAVE];
    const int64_t r0 = (int64_t)blockIdx.x * PA_CHUNK;
    const int64_t r1 = r0 + PA_CHUNK < nr ? r0 + PA_CHUNK : nr;
    /* exclusive scan of this chunk's histogram: P entries, PA_T threads */
    const int per = (P + PA_T - 1) / PA_T;
    uint32_t run = 0;
    for (int i = 0; i < per; i++) {
        int q = threadIdx.x * per + i;
        if (q < P) run += hist[(int64_t)q * gridDim.x + blockIdx.x];
    }
    uint32_t incl = run;
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t y = __shfl_up(incl, d, WAVE);
        if (lane_id() >= d) incl += y;
    }
    if (lane_id() == WAVE - 1) wsum[threadIdx.x / WAVE] = incl;
    __syncthreads();
    uint32_t off = incl - run;
    for (int w = 0; w < (int)threadIdx.x / WAVE; w++) off += wsum[w];
    for (int i = 0; i < per; i++) {
        int q = threadIdx.x * per + i;
        if (q < P) {
            lstart[q] = off;
            cur[q] = 0;
            off += hist[(int64_t)q * gridDim.x + blockIdx.x];
        }
    }
    if (threadIdx.x == PA_T - 1) lstart[P] = off;
    __syncthreads();
    const int64_t mask = ((int64_t)1 << s) - 1;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += PA_T) {
        int64_t k = ld_i64(rk, r) - kmin;
        int q = (int)(k >> s);
        uint32_t at = lstart[q] + atomicAdd(&cur[q], 1u);
        buf[at] = (uint32_t)(k & mask) | (part_slot(GP, r) << 16);
    }
    __syncthreads();
    const int wv = threadIdx.x / WAVE, lane = lane_id();
    for (int q = wv; q < P; q += PA_T / WAVE) {
        const uint32_t a = lstart[q], n = lstart[q + 1] - a;
        uint32_t* dst = words + pos[(int64_t)q * gridDim.x + blockIdx.x];
        for (uint32_t i = lane; i < n; i += WAVE) dst[i] = buf[a + i];
    }
}

__global__ void k_part_starts(const int64_t* hpos, int P, int G, int64_t* wstart) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= P) wstart[i] = hpos[(int64_t)i * G];      /* hpos[P * G] is the total */
}

/* ------------------------------------------------------------- pass B -- */

/* Region of partition p written by workgroup w. */
__device__ __forceinline__ uint64_t* part_region(uint64_t* rows, int p, int W, int w,
                                                 int64_t cap, int rw) {
    return rows + (((int64_t)p * W + w) * cap) * rw;
}

/* Tile slot q of a tile, with a compile-time index picked by a uniform
 * switch. */
template <int R>
__device__ __forceinline__ void pick_slot(const int64_t (&cv)[TL_NC][R], int q, int64_t (&out)[R]) {
    switch (q) {
    case 0:
#pragma unroll
        for (int j = 0; j < R; j++) out[j] = cv[0][j];
        break;
    case 1:
#pragma unroll
        for (int j = 0; j < R; j++) out[j] = cv[1][j];
        break;
    case 2:
#pragma unroll
        for (int j = 0; j < R; j++) out[j] = cv[2][j];
        break;
    default:
#pragma unroll
        for (int j = 0; j < R; j++) out[j] = cv[3][j];
        break;
    }
}

/* One round = PB_T * PB_R consecutive probe rows per workgroup.  The
 * round's passing rows are counting-sorted by partition in LDS (rank from
 * an LDS atomic, partition starts from a scan of the round's counts), then
 * copied out: consecutive threads copy consecutive sorted rows, so each
 * partition's rows of the round leave as one contiguous, coalesced run into
 * the workgroup's region of that partition. */
template <int NPRED, int NV>
__global__ __launch_bounds__(PB_T, 4) void k_pb_fact(PartFact A) {
    constexpr int R = PB_R;
    constexpr int RW = 1 + NV;
    constexpr int ROWS = PB_T * R;
    extern __shared__ uint64_t sm[];
    uint64_t* buf = sm;                                      /* [ROWS][RW] sorted rows */
    uint16_t* bufp = (uint16_t*)(buf + (int64_t)ROWS * RW);   /* partition of each */
    int32_t* cnt = (int32_t*)(bufp + ROWS);                  /* [P] round counts */
    int32_t* lstart = cnt + A.P;                             /* [P + 1] */
    int32_t* gbase = lstart + A.P + 1;                       /* [P] region offset */
    int32_t* nwr = gbase + A.P;                              /* [P] rows kept */
    int32_t* gcur = nwr + A.P;                               /* [P] rows in region */
    for (int i = threadIdx.x; i < A.P; i += PB_T) { cnt[i] = 0; gcur[i] = 0; }
    __syncthreads();
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int W = gridDim.x;
    const int64_t n_rounds = (A.nl + ROWS - 1) / ROWS;
    const int64_t smask = ((int64_t)1 << A.s) - 1;
    const int64_t lofs = (int64_t)wv * 64 * R + (int64_t)lane * R;
    bool over = false;
    int64_t cv[TL_NC][R], nx[TL_NC][R];
    int64_t r = blockIdx.x;
    if (r < n_rounds) load_tile<R>(A.T, r * ROWS + lofs, A.nl, cv);
    for (; r < n_rounds; r += W) {
        const int64_t base = r * ROWS + lofs;
        const int64_t rn = r + W;
        if (rn < n_rounds) load_tile<R>(A.T, rn * ROWS + lofs, A.nl, nx);   /* prefetch */
        const uint32_t m = tile_live<R>(A.lsel, base, A.nl) & star_pred<R, NPRED>(A.SP, A.T, cv);
        /* rows: packed word + values, partition and rank within the round */
        uint64_t w0[R], w1[R];
#pragma unroll
        for (int j = 0; j < R; j++) { w0[j] = (uint64_t)((cv[0][j] - A.kmin) & smask); w1[j] = 0; }
        for (int f = 0; f < A.npack; f++) {
            int64_t v[R];
            pick_slot<R>(cv, A.pack[f].col, v);
            const int sh = A.pack[f].shift;
            const int64_t b = A.pack[f].base;
            if (A.pack[f].word == 0) {
#pragma unroll
                for (int j = 0; j < R; j++) w0[j] |= (uint64_t)(v[j] - b) << sh;
            } else {
#pragma unroll
                for (int j = 0; j < R; j++) w1[j] |= (uint64_t)(v[j] - b) << sh;
            }
        }
        int64_t vv[NV > 0 ? NV : 1][R];
#pragma unroll
        for (int v = 0; v < NV; v++) pick_slot<R>(cv, A.val_col[v], vv[v]);
        int32_t pj[R], rk[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            uint64_t k = (uint64_t)(cv[0][j] - A.kmin);
            pj[j] = (((m >> j) & 1) && k < (uint64_t)A.range) ? (int32_t)(k >> A.s) : -1;
        }
#pragma unroll
        for (int j = 0; j < R; j++) rk[j] = pj[j] >= 0 ? atomicAdd(&cnt[pj[j]], 1) : 0;
        __syncthreads();
        /* round counts -> partition starts in the sorted buffer, and each
         * partition's place in its region (capacity-limited) */
        if (wv == 0) {
            const int per = (A.P + WAVE - 1) / WAVE;
            int32_t run = 0;
            for (int i = 0; i < per; i++) {
                int q = lane * per + i;
                if (q < A.P) run += cnt[q];
            }
            int32_t incl = run;
            for (int d = 1; d < WAVE; d <<= 1) {
                int32_t y = __shfl_up(incl, d, WAVE);
                if (lane >= d) incl += y;
            }
            int32_t off = incl - run;
            for (int i = 0; i < per; i++) {
                int q = lane * per + i;
                if (q < A.P) {
                    const int32_t n = cnt[q], g0 = gcur[q];
                    int32_t nw = n;
                    if ((int64_t)g0 + nw > A.cap) { nw = A.cap > g0 ? (int32_t)(A.cap - g0) : 0; over = true; }
                    lstart[q] = off;
                    gbase[q] = g0;
                    nwr[q] = nw;
                    gcur[q] = g0 + nw;
                    cnt[q] = 0;
                    off += n;
                }
            }
            if (lane == WAVE - 1) lstart[A.P] = off;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < R; j++) {
            if (pj[j] < 0) continue;
            const int32_t at = lstart[pj[j]] + rk[j];
            uint64_t* dst = buf + (int64_t)at * RW;
            dst[0] = (w1[j] << 32) | (w0[j] & 0xffffffffull);
#pragma unroll
            for (int v = 0; v < NV; v++) dst[1 + v] = (uint64_t)vv[v][j];
            bufp[at] = (uint16_t)pj[j];
        }
        __syncthreads();
        const int32_t T = lstart[A.P];
        for (int32_t i = threadIdx.x; i < T; i += PB_T) {
            const int p = bufp[i];
            const int32_t rr = i - lstart[p];
            if (rr >= nwr[p]) continue;
            uint64_t* dst = part_region(A.rows, p, W, blockIdx.x, A.cap, RW) +
                            (int64_t)(gbase[p] + rr) * RW;
            const uint64_t* src = buf + (int64_t)i * RW;
#pragma unroll
            for (int v = 0; v < RW; v++) dst[v] = src[v];
        }
        __syncthreads();
        if (rn < n_rounds) {
#pragma unroll
            for (int c = 0; c < TL_NC; c++)
#pragma unroll
                for (int j = 0; j < R; j++) cv[c][j] = nx[c][j];
        }
    }
    if (over) *A.overflow = 1;
    for (int i = threadIdx.x; i < A.P; i += PB_T) {
        int32_t c = gcur[i];
        A.counts[(int64_t)i * W + blockIdx.x] = c < A.cap ? c : (int32_t)A.cap;
    }
}

/* ------------------------------------------------------------- pass C -- */

struct PartProbe {
    const uint32_t* words;     /* build words, partition-major */
    const int64_t*  wstart;    /* [P + 1] */
    const uint64_t* rows;
    const int32_t*  counts;    /* [P][W] */
    int32_t W, P, s, nv;
    int64_t cap;
    int64_t n_slots;
    PackField pack[PART_MAX_PACK];
    /* aggregate sources (distinct input columns), decoded per row */
    int32_t nsrc;
    int8_t  src_kind[TL_NC];   /* 1 value slot, 2 packed field */
    int8_t  src_idx[TL_NC];
    uint8_t cell_src[GP_MAX_CELLS];
    TileCols srcT;             /* column types of the sources (for f64 inputs) */
    int*    dup;
};

/* Source q of a packed row: value slot or bit field (uniform branches). */
template <int RW>
__device__ __forceinline__ int64_t row_src(const PartProbe& B, int q, const uint64_t (&row)[RW + 1]) {
    const int idx = B.src_idx[q];
    if (B.src_kind[q] == 1) {
        switch (idx) {
        case 0:  return (int64_t)row[RW > 1 ? 1 : RW];
        case 1:  return (int64_t)row[RW > 2 ? 2 : RW];
        default: return (int64_t)row[RW > 3 ? 3 : RW];
        }
    }
    const uint64_t w = row[0];
    const PackField* F;
    switch (idx) {
    case 0:  F = &B.pack[0]; break;
    case 1:  F = &B.pack[1]; break;
    case 2:  F = &B.pack[2]; break;
    default: F = &B.pack[3]; break;
    }
    const uint32_t word = F->word == 0 ? (uint32_t)w : (uint32_t)(w >> 32);
    const uint64_t x = (word >> F->shift) & (F->bits >= 32 ? 0xffffffffull : ((1ull << F->bits) - 1));
    return (int64_t)x + F->base;
}

template <int RW>
__global__ __launch_bounds__(PC_T) void k_pc_probe(PartProbe B, GroupPlan P, uint64_t* partials) {
    extern __shared__ uint64_t smem[];
    const int64_t ncell_total = B.n_slots * P.ncells;
    const int64_t slice = (int64_t)1 << B.s;
    uint64_t* tab = smem;
    uint16_t* map = (uint16_t*)(smem + ncell_total);
    int32_t* pre = (int32_t*)(map + slice);                  /* W + 1 */
    __shared__ int32_t nonempty;
    __shared__ int32_t wsum[PC_T / WAVE];
    const uint32_t smask = (uint32_t)slice - 1;
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    for (int64_t i = threadIdx.x; i < ncell_total; i += PC_T)
        tab[i] = cell_identity(P.kind[i % P.ncells]);
    for (int p = blockIdx.x; p < B.P; p += gridDim.x) {
        for (int64_t i = threadIdx.x; i < slice; i += PC_T) map[i] = 0xffff;
        if (threadIdx.x == 0) nonempty = 0;
        /* region row counts -> exclusive prefix (W <= PC_T: one per thread) */
        int32_t c = (int)threadIdx.x < B.W ? B.counts[(int64_t)p * B.W + threadIdx.x] : 0;
        int32_t incl = c;
        for (int d = 1; d < WAVE; d <<= 1) {
            int32_t y = __shfl_up(incl, d, WAVE);
            if (lane >= d) incl += y;
        }
        if (lane == WAVE - 1) wsum[wv] = incl;
        __syncthreads();
        int32_t off = incl - c;
        for (int w = 0; w < wv; w++) off += wsum[w];
        if ((int)threadIdx.x < B.W) pre[threadIdx.x] = off;
        if ((int)threadIdx.x == B.W - 1) pre[B.W] = off + c;
        const int64_t e0 = B.wstart[p], e1 = B.wstart[p + 1];
        for (int64_t e = e0 + threadIdx.x; e < e1; e += PC_T) {
            uint32_t x = B.words[e];
            map[x & smask] = (uint16_t)(x >> 16);
        }
        __syncthreads();
        /* unique build keys <=> as many occupied entries as build words */
        int32_t ne = 0;
        for (int64_t i = threadIdx.x; i < slice; i += PC_T) ne += map[i] != 0xffff;
        for (int d = WAVE / 2; d >= 1; d >>= 1) ne += __shfl_xor(ne, d, WAVE);
        if (lane == 0 && ne) atomicAdd(&nonempty, ne);
        const int32_t T = pre[B.W];
        const uint64_t* prow = B.rows + (int64_t)p * B.W * B.cap * RW;
        for (int32_t i0 = (int32_t)threadIdx.x * PC_U; i0 < T; i0 += PC_T * PC_U) {
            /* region of row i0: last w with pre[w] <= i0 */
            int lo = 0, hi = B.W;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (pre[mid] <= i0) lo = mid; else hi = mid;
            }
            int w = lo;
            uint64_t rv[PC_U][RW + 1];
            int32_t s[PC_U];
#pragma unroll
            for (int u = 0; u < PC_U; u++) {
                const int32_t i = i0 + u;
                s[u] = -1;
                if (i < T) {
                    while (i >= pre[w + 1]) w++;
                    const uint64_t* row = prow + ((int64_t)w * B.cap + (i - pre[w])) * RW;
#pragma unroll
                    for (int v = 0; v < RW; v++) rv[u][v] = row[v];
                    s[u] = 0;
                } else {
#pragma unroll
                    for (int v = 0; v < RW; v++) rv[u][v] = 0;
                }
                rv[u][RW] = 0;
            }
#pragma unroll
            for (int u = 0; u < PC_U; u++) {
                if (s[u] < 0) continue;
                uint16_t sl = map[(uint32_t)rv[u][0] & smask];
                s[u] = sl == 0xffff ? -1 : (int32_t)sl;
            }
            int64_t sv[TL_NC][PC_U];
#pragma unroll
            for (int q = 0; q < TL_NC; q++) {
#pragma unroll
                for (int u = 0; u < PC_U; u++) sv[q][u] = q < B.nsrc ? row_src<RW>(B, q, rv[u]) : 0;
            }
            acc_tile<PC_U>(tab, P, B.cell_src, B.srcT, sv, s);
        }
        __syncthreads();
        if (threadIdx.x == 0 && nonempty != (int32_t)(e1 - e0)) *B.dup = 1;
        __syncthreads();
    }
    uint64_t* dst = partials + (int64_t)blockIdx.x * ncell_total;
    for (int64_t i = threadIdx.x; i < ncell_total; i += PC_T) dst[i] = tab[i];
}

/* --------------------------------------------------------------- driver -- */

static int bit_width_u64(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

/* Partitioned fused join -> group-by.  GP: the group plan with direct-array
 * slots (build-side keys, probe-side aggregates); kmin / range: build key
 * bounds.  TD_ERR_NYI: shape not taken, duplicate build keys or skew. */
int jg_partitioned(const tdd_join_group_spec_t* S, const GroupPlan& GP, int64_t kmin,
                   int64_t range, int64_t n_slots, bool forced, tdd_group_result_t* out) {
    const tdd_join_spec_t* J = &S->join;
    const size_t table_bytes = (size_t)n_slots * GP.ncells * 8;
    if (n_slots >= 0xffff || table_bytes > 48 * 1024) return TD_ERR_NYI;
    /* slice: map (u16 per key) + table + region prefix within 160 KB;
     * offsets fit 16 bits */
    const size_t lds_fixed = table_bytes + 4 * (PART_MAX_W + 1) + 1024;
    int s = 16;
    while (s > 8 && ((size_t)2 << s) + lds_fixed > 160 * 1024) s--;
    int64_t P64 = (range + ((int64_t)1 << s) - 1) >> s;
    if (P64 > 4096) return TD_ERR_NYI;
    const int P = (int)P64;
    /* Taken on request only: measured on MI355X for fact 200M x dim 16M it
     * trails the direct slot-map probe (the probe side's row writes and the
     * per-round flushes cost more than the gathers they save). */
    if (!forced) return TD_ERR_NYI;

    /* probe side: key at slot 0, star predicate at slots 1..npred */
    PartFact A;
    memset(&A, 0, sizeof(A));
    if (!star_plan(J->lkeys[0], S->lpred, A.T, A.SP)) return TD_ERR_NYI;

    /* row layout: distinct aggregate columns bit-packed when their bounds
     * are known and narrow, else 8-byte values */
    PartProbe B;
    memset(&B, 0, sizeof(B));
    int free0 = 32 - s, used0 = s, used1 = 0;
    int src_of_slot[TL_NC];
    for (int c = 0; c < TL_NC; c++) src_of_slot[c] = -1;
    int agg_src[TDD_MAX_AGGS];
    for (int a = 0; a < GP.n_aggs; a++) {
        agg_src[a] = 0;
        if (!GP.aggs[a].data) continue;
        if (GP.agg_side[a]) return TD_ERR_NYI;
        int c = tile_slot(A.T, GP.aggs[a]);
        if (c < 0) return TD_ERR_NYI;
        if (src_of_slot[c] < 0) {
            const tdd_col_t& col = GP.aggs[a];
            const int q = B.nsrc++;
            src_of_slot[c] = q;
            B.srcT.col[q] = col;
            int bits = col.ranged && col.type != TD_F64
                           ? bit_width_u64((uint64_t)col.hi - (uint64_t)col.lo) : 64;
            if (bits <= free0 && A.npack < PART_MAX_PACK) {
                A.pack[A.npack] = PackField{0, bits ? used0 : 0, bits, c, col.lo};
                used0 += bits;
                free0 -= bits;
                B.src_kind[q] = 2;
                B.src_idx[q] = (int8_t)A.npack++;
            } else if (bits <= 32 - used1 && A.npack < PART_MAX_PACK) {
                A.pack[A.npack] = PackField{1, bits ? used1 : 0, bits, c, col.lo};
                used1 += bits;
                B.src_kind[q] = 2;
                B.src_idx[q] = (int8_t)A.npack++;
            } else {
                if (A.nv >= PART_MAX_VALS) return TD_ERR_NYI;
                A.val_col[A.nv] = c;
                B.src_kind[q] = 1;
                B.src_idx[q] = (int8_t)A.nv++;
            }
        }
        agg_src[a] = src_of_slot[c];
    }
    B.srcT.n = B.nsrc;
    for (int i = 1; i < GP.ncells; i++) B.cell_src[i] = (uint8_t)agg_src[GP.col[i]];

    const int rw = 1 + A.nv;
    /* pass B: sorted-round buffer + per-partition counters in LDS */
    const size_t lds_b = (size_t)PB_T * PB_R * (rw * 8 + 2) + (size_t)4 * (5 * P + 1);
    if (lds_b > 150 * 1024) return TD_ERR_NYI;
    const int64_t n_rounds = ceil_div(J->nl > 0 ? J->nl : 1, (int64_t)PB_T * PB_R);
    int W = resident_grid(k_pb_fact<2, 2>, PB_T, lds_b, n_rounds);
    if (W > PART_MAX_W) W = PART_MAX_W;
    const int64_t rows_per_w = ceil_div(n_rounds, W) * PB_T * PB_R;
    A.cap = (int64_t)(1.3 * (double)rows_per_w / P) + 64;
    const size_t region_bytes = (size_t)W * P * A.cap * rw * 8;
    if (region_bytes > ((size_t)64 << 30)) return TD_ERR_NYI;

    /* ---- pass A: build side ---- */
    const int64_t nr = J->nr;
    const int64_t G64 = nr > 0 ? ceil_div(nr, PA_CHUNK) : 1;
    if (G64 * P > ((int64_t)1 << 26)) return TD_ERR_NYI;
    const int G = (int)G64;
    SCRATCH(hist, (size_t)P * G * 4);
    SCRATCH(hpos, ((size_t)P * G + 1) * 8);
    SCRATCH(words, (size_t)(nr > 0 ? nr : 1) * 4);
    SCRATCH(wstart, ((size_t)P + 1) * 8);
    int rc;
    if (nr > 0) {
        LAUNCH("jg_part_build", k_pa_hist, G, PA_T, (size_t)P * 4, J->rkeys[0],
               (int)tile_vec_ok(J->rkeys[0]), nr, kmin, s, P, hist.as<uint32_t>());
    } else if ((rc = tdd_fill(hist.p, 0, (size_t)P * G * 4))) {
        return rc;
    }
    rc = tdd::scan_u32_i64(hist.as<uint32_t>(), (int64_t)P * G, hpos.as<int64_t>());
    if (rc) return rc;
    if (nr > 0)
        LAUNCH("jg_part_build", k_pa_scatter, G, PA_T, ((size_t)PA_CHUNK + 2 * P + 1) * 4,
               J->rkeys[0], nr, kmin, s, P, GP, (const uint32_t*)hist.as<uint32_t>(),
               (const int64_t*)hpos.as<int64_t>(), words.as<uint32_t>());
    LAUNCH("jg_part_build", k_part_starts, ceil_div(P + 1, 256), 256, 0,
           (const int64_t*)hpos.as<int64_t>(), P, G, wstart.as<int64_t>());

    /* ---- pass B: probe side ---- */
    SCRATCH(rows, region_bytes);
    SCRATCH(counts, (size_t)W * P * 4);
    SCRATCH(flags, 16);
    rc = tdd_fill(flags.p, 0, 16);
    if (rc) return rc;
    A.lsel = J->lsel ? J->lsel->bits : nullptr;
    A.nl = J->nl;
    A.kmin = kmin;
    A.range = range;
    A.s = s;
    A.P = P;
    A.rows = rows.as<uint64_t>();
    A.counts = counts.as<int32_t>();
    A.overflow = flags.as<int>();
#define PB_LAUNCH(NP, NV) LAUNCH("jg_part_fact", (k_pb_fact<NP, NV>), W, PB_T, lds_b, A)
#define PB_LAUNCH_NV(NP)                  \
    switch (A.nv) {                       \
    case 0:  PB_LAUNCH(NP, 0); break;     \
    case 1:  PB_LAUNCH(NP, 1); break;     \
    case 2:  PB_LAUNCH(NP, 2); break;     \
    default: PB_LAUNCH(NP, 3); break;     \
    }
    switch (A.SP.npred) {
    case 0:  PB_LAUNCH_NV(0); break;
    case 1:  PB_LAUNCH_NV(1); break;
    default: PB_LAUNCH_NV(2); break;
    }
#undef PB_LAUNCH_NV
#undef PB_LAUNCH

    /* ---- pass C: per-partition probe ---- */
    B.words = words.as<uint32_t>();
    B.wstart = wstart.as<int64_t>();
    B.rows = rows.as<uint64_t>();
    B.counts = counts.as<int32_t>();
    B.W = W;
    B.P = P;
    B.s = s;
    B.nv = A.nv;
    B.cap = A.cap;
    B.n_slots = n_slots;
    for (int f = 0; f < A.npack; f++) B.pack[f] = A.pack[f];
    B.dup = flags.as<int>() + 1;
    const size_t lds = table_bytes + ((size_t)2 << s) + (size_t)4 * (W + 1);
    int C = resident_grid(k_pc_probe<2>, PC_T, lds, P);
    SCRATCH(parts, (size_t)C * table_bytes);
    uint64_t* pp = parts.as<uint64_t>();
    switch (rw) {
    case 1: LAUNCH("jg_part_probe", k_pc_probe<1>, C, PC_T, lds, B, GP, pp); break;
    case 2: LAUNCH("jg_part_probe", k_pc_probe<2>, C, PC_T, lds, B, GP, pp); break;
    case 3: LAUNCH("jg_part_probe", k_pc_probe<3>, C, PC_T, lds, B, GP, pp); break;
    default: LAUNCH("jg_part_probe", k_pc_probe<4>, C, PC_T, lds, B, GP, pp); break;
    }
    int fl[2] = {0, 0};
    HC(hipMemcpyAsync(fl, flags.p, 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    if (fl[0] || fl[1]) return TD_ERR_NYI;           /* skew overflow / duplicate keys */
    rc = da_finish(GP, pp, C, n_slots, out);
    if (rc) return rc;
    out->path = 3;
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}
