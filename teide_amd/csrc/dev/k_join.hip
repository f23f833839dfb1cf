/*
 * k_join.hip -- equi-join kernels and the fused filter -> join -> group-by.
 *
 * Direct map (single integer key on both sides, build keys dense in
 * [min, max]): head[key - min] holds the last build row inserted with that
 * key and next[row] chains the others, so a probe is one indexed load plus
 * a chain walk.  This is the device counterpart of the reference's chained
 * table for small builds (exec.c:7974-8128) without hashing.
 *
 * Radix partitioned (everything else, exec.c:7431-7950): both sides are
 * partitioned by the top bits of a 64-bit key hash in two passes (LDS
 * histogram per 2048-row tile; prefix sums; LDS counting sort of the tile and
 * coalesced run writes), then one workgroup per partition builds an LDS
 * open-addressing table from the build partition (in LDS-sized chunks when a
 * partition is large) and streams the probe partition through it.
 *
 * Output pairs are produced by a count pass and a fill pass, so the output
 * needs no atomics on a global cursor and the pairs of one probe row are
 * contiguous.
 */
#include "group_plan.h"
#include "pred_eval.h"
#include "tile.h"
#include "star.h"
#include "join_keys.h"

#include <stdio.h>

#define JT 256                  /* threads per join workgroup */
#define JROWS 8                 /* rows per thread per tile */
#define JTILE (JT * JROWS)      /* 2048 rows per tile */
#define PCAP 4096               /* LDS hash-table slots per partition chunk */
#define PCHUNK 2048             /* build rows per chunk (load <= 0.5) */
#define EMPTY_ROW 0xffffffffu

/* Block-wide exclusive scan of one int64 per thread (JT threads). */
__device__ int64_t jblock_scan(int64_t v, int64_t* total) {
    __shared__ int64_t ws[JT / WAVE];
    int lane = lane_id(), w = threadIdx.x / WAVE;
    int64_t x = v;
    for (int d = 1; d < WAVE; d <<= 1) {
        int64_t y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    if (lane == WAVE - 1) ws[w] = x;
    __syncthreads();
    int64_t base = 0, tot = 0;
    for (int i = 0; i < JT / WAVE; i++) {
        if (i < w) base += ws[i];
        tot += ws[i];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

/* ============================================================ direct map == */

__global__ void k_dm_build(tdd_col_t rk, int64_t nr, int64_t kmin, int32_t* head, int32_t* next,
                           int* dup) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nr;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = ld_i64(rk, r) - kmin;
        int32_t old = atomicExch(&head[i], (int32_t)r);
        next[r] = old;
        if (old >= 0) *dup = 1;
    }
}

/* Matches of left row l (probe side) through the direct map. */
struct DMap {
    const int32_t* head;
    const int32_t* next;
    int64_t kmin, range;
};

__device__ __forceinline__ int32_t dm_first(const DMap& D, const tdd_col_t& lk, int64_t l) {
    uint64_t i = (uint64_t)(ld_i64(lk, l) - D.kmin);
    return i < (uint64_t)D.range ? D.head[i] : -1;
}

__device__ __forceinline__ int64_t dm_count(const DMap& D, const tdd_col_t& lk, int64_t l) {
    int64_t c = 0;
    for (int32_t r = dm_first(D, lk, l); r >= 0; r = D.next[r]) c++;
    return c;
}

/* Count pass; a full join also marks the build rows that found a partner. */
__global__ __launch_bounds__(JT) void k_dm_count(DMap D, tdd_col_t lk, const uint64_t* lsel,
                                                 int64_t nl, int outer, int64_t* tile_cnt,
                                                 uint8_t* matched_r) {
    int64_t base = (int64_t)blockIdx.x * JTILE;
    int64_t c = 0;
    for (int j = 0; j < JROWS; j++) {
        int64_t l = base + j * JT + threadIdx.x;
        if (l < nl && row_selected(lsel, l)) {
            int64_t m = 0;
            for (int32_t r = dm_first(D, lk, l); r >= 0; r = D.next[r]) {
                m++;
                if (matched_r) matched_r[r] = 1;
            }
            c += (m == 0 && outer) ? 1 : m;
        }
    }
    int64_t tot;
    jblock_scan(c, &tot);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

__global__ __launch_bounds__(JT) void k_dm_fill(DMap D, tdd_col_t lk, const uint64_t* lsel,
                                                int64_t nl, int outer, const int64_t* tile_off,
                                                int64_t* l_idx, int64_t* r_idx) {
    int64_t base = (int64_t)blockIdx.x * JTILE;
    int64_t pos = tile_off[blockIdx.x];
    for (int j = 0; j < JROWS; j++) {
        int64_t l = base + j * JT + threadIdx.x;
        int64_t m = 0;
        bool on = l < nl && row_selected(lsel, l);
        if (on) m = dm_count(D, lk, l);
        int64_t c = (on && m == 0 && outer) ? 1 : m;
        int64_t tot;
        int64_t off = pos + jblock_scan(c, &tot);
        if (on) {
            if (m == 0 && outer) {
                l_idx[off] = l;
                r_idx[off] = -1;
            } else {
                for (int32_t r = dm_first(D, lk, l); r >= 0; r = D.next[r]) {
                    l_idx[off] = l;
                    r_idx[off] = r;
                    off++;
                }
            }
        }
        pos += tot;
    }
}

/* ================================================================ radix == */

struct PartIn {
    JKeys           K;
    const uint64_t* sel;      /* probe side only */
    int64_t         n;
    int32_t         bits;     /* radix bits */
    int32_t         n_tiles;
    int32_t         plain;    /* exact key read by a plain 8- or 4-byte load, else 0 */
    int32_t         span;     /* tiles per workgroup: workgroup b takes [b * span, ...) */
    int32_t         nwg;      /* workgroups */
};

/* Host: the element bytes of an exact single key that ld_i64 reads with a
 * plain signed load (4-byte SYM: unsigned), else 0. */
static int part_plain(const JKeys& K) {
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

/* Key words of a thread's JROWS rows of the tile at `base` (bit j of the
 * result: row j exists and is selected).  A plain key loads every row
 * first, with no width switch between the loads. */
__device__ __forceinline__ uint32_t part_keys(const PartIn& I, int64_t base,
                                              uint64_t (&kw)[JROWS]) {
    uint32_t live = 0;
#pragma unroll
    for (int j = 0; j < JROWS; j++) {
        const int64_t r = base + j * JT + threadIdx.x;
        if (r < I.n && row_selected(I.sel, r)) live |= 1u << j;
    }
    const void* d = I.K.cols[0].data;
    if (I.plain == 8) {
#pragma unroll
        for (int j = 0; j < JROWS; j++) {
            const int64_t r = base + j * JT + threadIdx.x;
            kw[j] = r < I.n ? (uint64_t)((const int64_t*)d)[r] : 0;
        }
    } else if (I.plain == 4) {
        const bool uns = I.K.cols[0].type == TD_SYM;
#pragma unroll
        for (int j = 0; j < JROWS; j++) {
            const int64_t r = base + j * JT + threadIdx.x;
            kw[j] = r >= I.n ? 0
                  : uns ? (uint64_t)((const uint32_t*)d)[r]
                        : (uint64_t)(int64_t)((const int32_t*)d)[r];
        }
    } else {
        for (int j = 0; j < JROWS; j++)
            kw[j] = ((live >> j) & 1) ? key_word(I.K, base + j * JT + threadIdx.x) : 0;
    }
    return live;
}

__device__ __forceinline__ uint32_t part_of(uint64_t kw, int bits) {
    return (uint32_t)(jmix(kw) >> (64 - bits));
}

/* Counts per (partition, workgroup) over the workgroup's span of tiles:
 * H[p * nwg + b], so the offsets take P x nwg entries, not P x tiles. */
__global__ __launch_bounds__(JT) void k_part_hist(PartIn I, uint32_t* H) {
    extern __shared__ uint32_t hist[];
    const int P = 1 << I.bits;
    for (int i = threadIdx.x; i < P; i += JT) hist[i] = 0;
    __syncthreads();
    const int t1 = (blockIdx.x + 1) * I.span < I.n_tiles ? (blockIdx.x + 1) * I.span : I.n_tiles;
    for (int t = blockIdx.x * I.span; t < t1; t++) {
        uint64_t kw[JROWS];
        const uint32_t live = part_keys(I, (int64_t)t * JTILE, kw);
#pragma unroll
        for (int j = 0; j < JROWS; j++)
            if ((live >> j) & 1) atomicAdd(&hist[part_of(kw[j], I.bits)], 1u);
    }
    __syncthreads();
    for (int p = threadIdx.x; p < P; p += JT) H[(int64_t)p * I.nwg + blockIdx.x] = hist[p];
}

/* Per tile of the workgroup's span: an LDS counting sort by partition, then
 * each run is written at its partition's running cursor (starting at
 * Off[p * nwg + b]) with consecutive lanes on consecutive addresses. */
__global__ __launch_bounds__(JT) void k_part_scatter(PartIn I, const int64_t* Off,
                                                     uint64_t* out_key, uint32_t* out_row) {
    extern __shared__ uint8_t smem[];
    const int P = 1 << I.bits;
    int64_t* cur = (int64_t*)smem;                          /* P */
    uint32_t* hist = (uint32_t*)(cur + P);                  /* P */
    uint32_t* loff = hist + P;                              /* P + 1 */
    uint64_t* skey = (uint64_t*)(loff + P + 2);             /* JTILE, 8-aligned (P even) */
    uint32_t* srow = (uint32_t*)(skey + JTILE);             /* JTILE */
    uint16_t* spart = (uint16_t*)(srow + JTILE);            /* JTILE */
    for (int i = threadIdx.x; i < P; i += JT) cur[i] = Off[(int64_t)i * I.nwg + blockIdx.x];
    const int t1 = (blockIdx.x + 1) * I.span < I.n_tiles ? (blockIdx.x + 1) * I.span : I.n_tiles;
    for (int t = blockIdx.x * I.span; t < t1; t++) {
    for (int i = threadIdx.x; i < P; i += JT) hist[i] = 0;
    __syncthreads();
    int64_t base = (int64_t)t * JTILE;
    uint64_t kw[JROWS];
    uint32_t pr[JROWS], rank[JROWS];
    const uint32_t live = part_keys(I, base, kw);
    for (int j = 0; j < JROWS; j++) {
        pr[j] = 0xffffffffu;
        if ((live >> j) & 1) {
            pr[j] = part_of(kw[j], I.bits);
            rank[j] = atomicAdd(&hist[pr[j]], 1u);
        }
    }
    __syncthreads();
    /* exclusive scan of the tile histogram (P <= 2048 bins, 8 per thread) */
    {
        const int per = (P + JT - 1) / JT;
        int64_t s = 0;
        for (int q = 0; q < per; q++) {
            int p = threadIdx.x * per + q;
            if (p < P) s += hist[p];
        }
        int64_t tot;
        int64_t pre = jblock_scan(s, &tot);
        for (int q = 0; q < per; q++) {
            int p = threadIdx.x * per + q;
            if (p < P) { loff[p] = (uint32_t)pre; pre += hist[p]; }
        }
        if (threadIdx.x == 0) loff[P] = (uint32_t)tot;
    }
    __syncthreads();
    for (int j = 0; j < JROWS; j++) {
        if (pr[j] == 0xffffffffu) continue;
        uint32_t at = loff[pr[j]] + rank[j];
        skey[at] = kw[j];
        srow[at] = (uint32_t)(base + j * JT + threadIdx.x);
        spart[at] = (uint16_t)pr[j];
    }
    __syncthreads();
    uint32_t cnt = loff[P];
    for (uint32_t i = threadIdx.x; i < cnt; i += JT) {
        uint32_t p = spart[i];
        int64_t dst = cur[p] + (i - loff[p]);
        out_key[dst] = skey[i];
        out_row[dst] = srow[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += JT) cur[i] += hist[i];
    __syncthreads();
    }
}

struct Parts {
    const uint64_t* key;
    const uint32_t* row;
    const int64_t*  off;      /* partition-major offsets: partition p starts at off[p * stride] */
    int32_t         stride;   /* workgroups of the partition passes */
};

__device__ __forceinline__ int64_t part_start(const Parts& S, int p) {
    return S.off[(int64_t)p * S.stride];
}

struct BPArgs {
    Parts    L, R;            /* probe (left) and build (right) partitions */
    JKeys    LK, RK;          /* for verification when not exact */
    int32_t  outer, full, exact, fill;
    int64_t* pcount;          /* count pass: matches per partition */
    const int64_t* poff;      /* fill pass: output offset per partition */
    int64_t* l_idx;
    int64_t* r_idx;
    uint8_t* lmatched;        /* per probe entry (outer joins) */
    uint8_t* rmatched;        /* per build row (full joins) */
};

__global__ __launch_bounds__(JT) void k_bp(BPArgs A) {
    __shared__ uint64_t tkey[PCAP];
    __shared__ uint32_t trow[PCAP];
    __shared__ unsigned long long cursor;
    const int p = blockIdx.x;
    const int64_t bs = part_start(A.R, p), be = part_start(A.R, p + 1);
    const int64_t ps = part_start(A.L, p), pe = part_start(A.L, p + 1);
    int64_t cnt = 0;
    if (threadIdx.x == 0) cursor = 0;
    __syncthreads();
    int64_t obase = A.fill ? A.poff[p] : 0;
    for (int64_t c0 = bs; c0 < be; c0 += PCHUNK) {          /* uniform across the block */
        int64_t c1 = c0 + PCHUNK < be ? c0 + PCHUNK : be;
        for (int i = threadIdx.x; i < PCAP; i += JT) trow[i] = EMPTY_ROW;
        __syncthreads();
        for (int64_t i = c0 + threadIdx.x; i < c1; i += JT) {
            uint64_t k = A.R.key[i];
            uint32_t s = (uint32_t)jmix(k) & (PCAP - 1);
            while (atomicCAS(&trow[s], EMPTY_ROW, A.R.row[i]) != EMPTY_ROW) s = (s + 1) & (PCAP - 1);
            tkey[s] = k;
        }
        __syncthreads();
        for (int64_t i = ps + threadIdx.x; i < pe; i += JT) {
            uint64_t k = A.L.key[i];
            uint32_t l = A.L.row[i];
            uint32_t s = (uint32_t)jmix(k) & (PCAP - 1);
            bool any = false;
            for (;;) {
                uint32_t r = trow[s];
                if (r == EMPTY_ROW) break;
                if (tkey[s] == k && (A.exact || keys_equal(A.LK, l, A.RK, r))) {
                    any = true;
                    if (A.fill) {
                        int64_t at = obase + (int64_t)atomicAdd(&cursor, 1ull);
                        A.l_idx[at] = l;
                        A.r_idx[at] = r;
                    } else {
                        cnt++;
                    }
                    if (A.full) A.rmatched[r] = 1;
                }
                s = (s + 1) & (PCAP - 1);
            }
            if (any && A.outer) A.lmatched[i] = 1;
        }
        __syncthreads();
    }
    if (A.outer) {
        for (int64_t i = ps + threadIdx.x; i < pe; i += JT) {
            if (A.lmatched[i]) continue;
            if (A.fill) {
                int64_t at = obase + (int64_t)atomicAdd(&cursor, 1ull);
                A.l_idx[at] = A.L.row[i];
                A.r_idx[at] = -1;
            } else {
                cnt++;
            }
        }
    }
    if (!A.fill) {
        int64_t tot;
        jblock_scan(cnt, &tot);
        if (threadIdx.x == 0) A.pcount[p] = tot;
    }
}

/* Unmatched build rows of a full join: flag, scan, append with l = -1. */
__global__ void k_unmatched_flag(const uint8_t* matched, int64_t n, uint32_t* flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = matched[i] == 0;
}

__global__ void k_unmatched_emit(const uint8_t* matched, int64_t n, const int64_t* pos,
                                 int64_t base, int64_t* l_idx, int64_t* r_idx) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !matched[i]) {
        l_idx[base + pos[i]] = -1;
        r_idx[base + pos[i]] = i;
    }
}

/* ----------------------------------------------------------- host side -- */

static bool is_int(int8_t t) {
    return t != TD_F64 && t != TD_GUID && t != TD_LIST && t != TD_TABLE && t != TD_CHAR;
}

static int64_t read_i64(const void* dptr) {
    int64_t v = 0;
    if (hipMemcpyAsync(&v, dptr, 8, hipMemcpyDeviceToHost, tdd::g_stream) != hipSuccess) return -1;
    if (hipStreamSynchronize(tdd::g_stream) != hipSuccess) return -1;
    return v;
}

static int alloc_pairs(int64_t n, tdd_join_result_t* out) {
    out->n_pairs = n;
    out->l_idx = (int64_t*)td__arena_alloc((size_t)(n > 0 ? n : 1) * 8);
    out->r_idx = (int64_t*)td__arena_alloc((size_t)(n > 0 ? n : 1) * 8);
    if (!out->l_idx || !out->r_idx) {
        td__arena_free(out->l_idx);
        td__arena_free(out->r_idx);
        out->l_idx = out->r_idx = nullptr;
        return TD_ERR_OOM;
    }
    return TD_OK;
}

/* Append the unmatched build rows of a full join after `base` pairs. */
static int append_unmatched(const uint8_t* matched, int64_t nr, int64_t* n_unmatched,
                            Scratch* flag_out, Scratch* pos_out) {
    int blocks = (int)ceil_div(nr > 0 ? nr : 1, 256);
    LAUNCH("join_unmatched", k_unmatched_flag, blocks, 256, 0, matched, nr, flag_out->as<uint32_t>());
    int rc = tdd::scan_u32_i64(flag_out->as<uint32_t>(), nr, pos_out->as<int64_t>());
    if (rc) return rc;
    *n_unmatched = read_i64(pos_out->as<int64_t>() + nr);
    return *n_unmatched < 0 ? TD_ERR_IO : TD_OK;
}

/* Dense unique-ish integer build keys: chained direct map.  Returns
 * TD_ERR_NYI when the build keys are not dense enough. */
static int join_direct(const tdd_join_spec_t* S, tdd_join_result_t* out, bool forced) {
    int64_t mn, mx;
    int rc = tdd_minmax(&S->rkeys[0], S->nr, nullptr, &mn, &mx);
    if (rc) return rc;
    int64_t range = mn > mx ? 1 : (int64_t)((uint64_t)mx - (uint64_t)mn + 1);
    if (mn > mx) mn = 0;
    if (range <= 0 || range > ((int64_t)1 << 31)) return TD_ERR_NYI;
    if (!forced && range > 4 * S->nr + 4096) return TD_ERR_NYI;
    if (S->nr >= ((int64_t)1 << 31)) return TD_ERR_NYI;
    SCRATCH(head, (size_t)range * 4);
    SCRATCH(next, (size_t)(S->nr > 0 ? S->nr : 1) * 4);
    SCRATCH(dup, 8);
    rc = tdd_fill(head.p, ~0ull, (size_t)range * 4);
    if (rc) return rc;
    rc = tdd_fill(dup.p, 0, 8);
    if (rc) return rc;
    if (S->nr > 0)
        LAUNCH("join_dm_build", k_dm_build, stream_grid(ceil_div(S->nr, 256 * 4)), 256, 0,
               S->rkeys[0], S->nr, mn, head.as<int32_t>(), next.as<int32_t>(), dup.as<int>());
    DMap D{head.as<int32_t>(), next.as<int32_t>(), mn, range};
    int outer = S->join_type >= 1, full = S->join_type == 2;
    int64_t n_tiles = ceil_div(S->nl > 0 ? S->nl : 1, JTILE);
    size_t nr1 = (size_t)(S->nr > 0 ? S->nr : 1);
    SCRATCH(tcnt, (size_t)n_tiles * 8);
    SCRATCH(toff, (size_t)(n_tiles + 1) * 8);
    SCRATCH(matched, nr1);
    SCRATCH(flag, nr1 * 4);
    SCRATCH(pos, (nr1 + 1) * 8);
    if (full) {
        rc = tdd_fill(matched.p, 0, nr1);
        if (rc) return rc;
    }
    const uint64_t* lsel = S->lsel ? S->lsel->bits : nullptr;
    LAUNCH("join_dm_count", k_dm_count, (int)n_tiles, JT, 0, D, S->lkeys[0], lsel, S->nl, outer,
           tcnt.as<int64_t>(), full ? matched.as<uint8_t>() : (uint8_t*)nullptr);
    rc = tdd::scan_i64(tcnt.as<int64_t>(), n_tiles, toff.as<int64_t>());
    if (rc) return rc;
    int64_t npairs = read_i64(toff.as<int64_t>() + n_tiles);
    if (npairs < 0) return TD_ERR_IO;
    int64_t n_unm = 0;
    if (full) {
        rc = append_unmatched(matched.as<uint8_t>(), S->nr, &n_unm, &flag, &pos);
        if (rc) return rc;
    }
    POLL_CANCEL();
    rc = alloc_pairs(npairs + n_unm, out);
    if (rc) return rc;
    if (npairs > 0)
        LAUNCH("join_dm_fill", k_dm_fill, (int)n_tiles, JT, 0, D, S->lkeys[0], lsel, S->nl,
               outer, (const int64_t*)toff.as<int64_t>(), out->l_idx, out->r_idx);
    if (n_unm > 0)
        LAUNCH("join_unmatched", k_unmatched_emit, (int)ceil_div(S->nr, 256), 256, 0,
               (const uint8_t*)matched.as<uint8_t>(), S->nr, (const int64_t*)pos.as<int64_t>(),
               npairs, out->l_idx, out->r_idx);
    out->path = 1;
    out->radix_bits = 0;
    return TD_OK;
}

/* Partition one side into (key word, row) runs.  Returns TD_OK. */
static int partition_side(const PartIn& I, uint64_t** key_out, uint32_t** row_out,
                          int64_t** off_out, int64_t* total) {
    const int P = 1 << I.bits;
    size_t nh = (size_t)P * I.nwg;
    SCRATCH(H, nh * 4);
    int64_t* off = (int64_t*)td__arena_alloc((nh + 1) * 8);
    if (!off) return TD_ERR_OOM;
    LAUNCH("join_part_hist", k_part_hist, I.nwg, JT, (size_t)P * 4, I, H.as<uint32_t>());
    int rc = tdd::scan_u32_i64(H.as<uint32_t>(), (int64_t)nh, off);
    if (rc) { td__arena_free(off); return rc; }
    int64_t n = read_i64(off + nh);
    uint64_t* key = (uint64_t*)td__arena_alloc((size_t)(n > 0 ? n : 1) * 8);
    uint32_t* row = (uint32_t*)td__arena_alloc((size_t)(n > 0 ? n : 1) * 4);
    if (!key || !row || n < 0) {
        td__arena_free(off); td__arena_free(key); td__arena_free(row);
        return n < 0 ? TD_ERR_IO : TD_ERR_OOM;
    }
    size_t lds = (size_t)P * 8 + (size_t)(2 * P + 2) * 4 + (size_t)JTILE * (8 + 4 + 2);
    LAUNCH("join_part_scatter", k_part_scatter, I.nwg, JT, lds, I,
           (const int64_t*)off, key, row);
    *key_out = key;
    *row_out = row;
    *off_out = off;
    *total = n;
    return TD_OK;
}

static int join_radix(const tdd_join_spec_t* S, tdd_join_result_t* out) {
    if (S->nl >= ((int64_t)1 << 32) || S->nr >= ((int64_t)1 << 32)) return TD_ERR_NYI;
    int bits = S->radix_bits;
    if (bits <= 0) {
        bits = 3;
        while (bits < 11 && ((int64_t)1 << bits) * 1536 < S->nr) bits++;
    }
    if (bits > 11) bits = 11;
    const int P = 1 << bits;
    JKeys LK, RK;
    memset(&LK, 0, sizeof(LK));
    memset(&RK, 0, sizeof(RK));
    LK.n = RK.n = S->n_keys;
    bool exact = S->n_keys == 1 && is_int(S->lkeys[0].type) && is_int(S->rkeys[0].type);
    LK.exact = RK.exact = exact;
    for (int k = 0; k < S->n_keys; k++) { LK.cols[k] = S->lkeys[k]; RK.cols[k] = S->rkeys[k]; }
    PartIn IL{LK, S->lsel ? S->lsel->bits : nullptr, S->nl, bits,
              (int32_t)ceil_div(S->nl > 0 ? S->nl : 1, JTILE)};
    PartIn IR{RK, nullptr, S->nr, bits, (int32_t)ceil_div(S->nr > 0 ? S->nr : 1, JTILE)};
    IL.plain = part_plain(LK);
    IR.plain = part_plain(RK);
    for (PartIn* I : {&IL, &IR}) {             /* a span of tiles per workgroup */
        const int64_t want = (int64_t)tdd::g_num_cu * 2;
        I->span = (int32_t)ceil_div(I->n_tiles, want);
        I->nwg = (int32_t)ceil_div(I->n_tiles, I->span);
    }
    uint64_t *lkey = nullptr, *rkey = nullptr;
    uint32_t *lrow = nullptr, *rrow = nullptr;
    int64_t *loff = nullptr, *roff = nullptr, nlp = 0, nrp = 0;
    int rc = partition_side(IL, &lkey, &lrow, &loff, &nlp);
    if (!rc) rc = partition_side(IR, &rkey, &rrow, &roff, &nrp);
    struct Frees { void* p[6]; ~Frees() { for (void* q : p) if (q) td__arena_free(q); } }
        fr{{lkey, lrow, loff, rkey, rrow, roff}};
    if (rc) return rc;
    POLL_CANCEL();
    BPArgs A;
    memset(&A, 0, sizeof(A));
    A.L = Parts{lkey, lrow, loff, IL.nwg};
    A.R = Parts{rkey, rrow, roff, IR.nwg};
    A.LK = LK;
    A.RK = RK;
    A.exact = exact;
    A.outer = S->join_type >= 1;
    A.full = S->join_type == 2;
    SCRATCH(pcount, (size_t)P * 8);
    SCRATCH(poff, (size_t)(P + 1) * 8);
    SCRATCH(lm, (size_t)(nlp > 0 ? nlp : 1));
    SCRATCH(rm, (size_t)(S->nr > 0 ? S->nr : 1));
    if (A.outer) { rc = tdd_fill(lm.p, 0, (size_t)(nlp > 0 ? nlp : 1)); if (rc) return rc; }
    if (A.full) { rc = tdd_fill(rm.p, 0, (size_t)(S->nr > 0 ? S->nr : 1)); if (rc) return rc; }
    A.lmatched = lm.as<uint8_t>();
    A.rmatched = rm.as<uint8_t>();
    A.pcount = pcount.as<int64_t>();
    A.fill = 0;
    LAUNCH("join_build_probe", k_bp, P, JT, 0, A);
    rc = tdd::scan_i64(pcount.as<int64_t>(), P, poff.as<int64_t>());
    if (rc) return rc;
    int64_t npairs = read_i64(poff.as<int64_t>() + P);
    if (npairs < 0) return TD_ERR_IO;
    int64_t n_unm = 0;
    SCRATCH(flag, (size_t)(S->nr > 0 ? S->nr : 1) * 4);
    SCRATCH(pos, (size_t)(S->nr + 1) * 8);
    if (A.full) {
        rc = append_unmatched(rm.as<uint8_t>(), S->nr, &n_unm, &flag, &pos);
        if (rc) return rc;
    }
    rc = alloc_pairs(npairs + n_unm, out);
    if (rc) return rc;
    if (A.outer) { rc = tdd_fill(lm.p, 0, (size_t)(nlp > 0 ? nlp : 1)); if (rc) return rc; }
    A.fill = 1;
    A.full = 0;                      /* matched build rows are already known */
    A.poff = poff.as<int64_t>();
    A.l_idx = out->l_idx;
    A.r_idx = out->r_idx;
    if (npairs > 0) LAUNCH("join_build_probe", k_bp, P, JT, 0, A);
    if (n_unm > 0)
        LAUNCH("join_unmatched", k_unmatched_emit, (int)ceil_div(S->nr, 256), 256, 0,
               (const uint8_t*)rm.as<uint8_t>(), S->nr, (const int64_t*)pos.as<int64_t>(), npairs,
               out->l_idx, out->r_idx);
    out->path = 2;
    out->radix_bits = bits;
    return TD_OK;
}

extern "C" int tdd_join(const tdd_join_spec_t* S, tdd_join_result_t* out) {
    memset(out, 0, sizeof(*out));
    if (S->n_keys < 1 || S->n_keys > TDD_MAX_KEYS || S->join_type < 0 || S->join_type > 2)
        return TD_ERR_NYI;
    for (int k = 0; k < S->n_keys; k++) {
        if (!S->lkeys[k].data || !S->rkeys[k].data) return TD_ERR_SCHEMA;
        if ((S->lkeys[k].type == TD_F64) != (S->rkeys[k].type == TD_F64)) return TD_ERR_TYPE;
    }
    int rc = TD_ERR_NYI;
    bool single_int = S->n_keys == 1 && is_int(S->lkeys[0].type) && is_int(S->rkeys[0].type);
    if (single_int && S->force_path != 2) rc = join_direct(S, out, S->force_path == 1);
    if (rc == TD_ERR_NYI) rc = join_radix(S, out);
    if (rc) return rc;
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}

/* ===================================================== fused join-group == */
/* Probe rows are filtered, looked up in the direct map and folded into a
 * direct-array group table.  slotmap mode: a unique build key whose group
 * keys all come from the build side maps straight to the group slot, so a
 * probe row costs one indexed load; chain mode walks the build rows. */

struct JGArgs {
    tdd_pred_t      pred;
    int32_t         has_pred;
    const uint64_t* lsel;
    int64_t         nl;
    tdd_col_t       lk;
    DMap            D;
    const int32_t*  slotmap;      /* slotmap mode, else NULL */
    int64_t         n_slots;
};

template <bool SLOTMAP>
__global__ __launch_bounds__(512) void k_jg_lds(JGArgs J, GroupPlan P, uint64_t* partials) {
    extern __shared__ uint64_t tab[];
    int64_t ncell_total = J.n_slots * P.ncells;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x)
        tab[i] = cell_identity(P.kind[i % P.ncells]);
    __syncthreads();
    int64_t n_words = (J.nl + 63) >> 6;
    int lane = lane_id();
    int64_t wave = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t w = wave; w < n_words; w += n_waves) {
        int64_t l = w * 64 + lane;
        if (l >= J.nl || !row_selected(J.lsel, l)) continue;
        if (J.has_pred && !pred_eval(J.pred, l)) continue;
        uint64_t i = (uint64_t)(ld_i64(J.lk, l) - J.D.kmin);
        if (i >= (uint64_t)J.D.range) continue;
        if (SLOTMAP) {
            int32_t s = J.slotmap[i];
            if (s >= 0) cells_accum2(&tab[(int64_t)s * P.ncells], P, l, 0);
        } else {
            for (int32_t r = J.D.head[i]; r >= 0; r = J.D.next[r])
                cells_accum2(&tab[da_slot2(P, l, r) * P.ncells], P, l, r);
        }
    }
    __syncthreads();
    uint64_t* dst = partials + (int64_t)blockIdx.x * ncell_total;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x) dst[i] = tab[i];
}

template <bool SLOTMAP>
__global__ __launch_bounds__(512) void k_jg_global(JGArgs J, GroupPlan P, uint64_t* tab) {
    int64_t n_words = (J.nl + 63) >> 6;
    int lane = lane_id();
    int64_t wave = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t w = wave; w < n_words; w += n_waves) {
        int64_t l = w * 64 + lane;
        if (l >= J.nl || !row_selected(J.lsel, l)) continue;
        if (J.has_pred && !pred_eval(J.pred, l)) continue;
        uint64_t i = (uint64_t)(ld_i64(J.lk, l) - J.D.kmin);
        if (i >= (uint64_t)J.D.range) continue;
        if (SLOTMAP) {
            int32_t s = J.slotmap[i];
            if (s >= 0) cells_accum2(&tab[(int64_t)s * P.ncells], P, l, 0);
        } else {
            for (int32_t r = J.D.head[i]; r >= 0; r = J.D.next[r])
                cells_accum2(&tab[da_slot2(P, l, r) * P.ncells], P, l, r);
        }
    }
}

/* ------------------------------------------------- slot map, tiled probe -- */

/* Unique-key slot map: one scatter pass stores every build row's group slot
 * at its key; the keys are unique exactly when nr entries end up non-empty
 * (rows sharing a key share an entry), which one streaming count over the
 * map decides.  Map entries are u16 when the slots fit (EMPTY = 0xffff),
 * else i32 (EMPTY = -1): the probe's random gathers touch half the lines. */
template <typename MT>
__global__ __launch_bounds__(256) void k_sm_scatter(TileCols T, int64_t nr, int64_t kmin,
                                                    GroupPlan P, MT* map) {
    const int R = 8;
    const int lane = lane_id();
    const int64_t n_tiles = (nr + 64 * R - 1) / (64 * R);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t n_waves = (int64_t)gridDim.x * blockDim.x / WAVE;
    for (int64_t t = wave; t < n_tiles; t += n_waves) {
        const int64_t base = t * 64 * R + (int64_t)lane * R;
        int64_t cv[TL_NC][R];
        load_tile<R>(T, base, nr, cv);
        uint32_t m = tile_live<R>(nullptr, base, nr);
#pragma unroll
        for (int j = 0; j < R; j++) {
            if (!((m >> j) & 1)) continue;
            int64_t slot = 0;
            for (int k = 0; k < P.n_keys; k++)          /* key k sits in tile slot k + 1 */
                slot += (cget<R>(cv, k + 1, j) - P.kmin[k]) * P.kstride[k];
            map[cget<R>(cv, 0, j) - kmin] = (MT)slot;
        }
    }
}

template <typename MT>
__global__ __launch_bounds__(256) void k_sm_count(const MT* map, int64_t range, MT empty,
                                                  int64_t* part) {
    __shared__ int64_t sc[256 / WAVE];
    int64_t c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < range;
         i += (int64_t)gridDim.x * blockDim.x)
        c += map[i] != empty;
    for (int d = WAVE / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, WAVE);
    if (lane_id() == 0) sc[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / WAVE; w++) c += sc[w];
        part[blockIdx.x] = c;
    }
}

__global__ void k_sum_final(const int64_t* part, int n, int64_t* out) {
    __shared__ int64_t s[256];
    int64_t c = 0;
    for (int i = threadIdx.x; i < n; i += 256) c += part[i];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

/* Slot map of unique build keys: map[key - kmin] = the row's group slot.
 * *unique tells whether every build key was distinct (occupied entries ==
 * build rows); with `deferred` the occupied count is left at *deferred on
 * the device instead (no host sync) and *unique is set for the caller to
 * check once its own work is queued. */
template <class MT>
static int sm_build(const tdd_col_t& rk, int64_t nr, int64_t kmin, int64_t range,
                    const GroupPlan& P, MT* map, int* unique, int64_t* deferred = nullptr) {
    TileCols T;
    memset(&T, 0, sizeof(T));
    tile_slot(T, rk);
    for (int k = 0; k < P.n_keys; k++) {     /* distinct slots: key k at k + 1 */
        if (T.n >= TL_NC) return TD_ERR_NYI;
        T.col[T.n] = P.keys[k];
        if (tile_vec_ok(P.keys[k])) T.vec |= 1u << T.n;
        T.n++;
    }
    int rc = tdd_fill(map, sizeof(MT) == 2 ? 0xffffffffffffffffull : ~0ull, (size_t)range * sizeof(MT));
    if (rc) return rc;
    if (nr > 0)
        LAUNCH("jg_slotmap", k_sm_scatter<MT>, stream_grid(ceil_div(nr, 256 * 8), 8), 256, 0, T,
               nr, kmin, P, map);
    int grid = stream_grid(ceil_div(range, 256 * 16), 4);
    SCRATCH(part, (size_t)(grid + 1) * 8);
    LAUNCH("jg_slotmap", k_sm_count<MT>, grid, 256, 0, (const MT*)map, range, (MT)~(MT)0,
           part.as<int64_t>() + 1);
    LAUNCH("jg_slotmap", k_sum_final, 1, 256, 0, (const int64_t*)part.as<int64_t>() + 1, grid,
           deferred ? deferred : part.as<int64_t>());
    if (deferred) {                      /* checked by the caller after its own launches */
        *unique = 1;
        return TD_OK;
    }
    int64_t cnt = 0;
    HC(hipMemcpyAsync(&cnt, part.p, 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    *unique = cnt == nr;
    return TD_OK;
}

/* Tiled fused probe (slot-map mode): probe-side columns -- the join key,
 * the predicate's columns and the aggregate inputs -- sit in the tile cache;
 * per tile a lane issues all its column loads, evaluates the predicate on
 * R rows at once, issues R independent slot-map gathers and folds the rows
 * into the workgroup's LDS table. */
struct JGTile {
    TileCols        T;
    tdd_pred_t      pred;          /* operands are tile slots */
    int32_t         has_pred;
    int32_t         key_c;         /* tile slot of the join key */
    const uint64_t* lsel;
    int64_t         nl;
    int64_t         kmin, range;
    const void*     slotmap;       /* MT entries: u16 (EMPTY 0xffff) or i32 (EMPTY -1) */
    int64_t         n_slots;
    uint8_t         agg_c[TDD_MAX_AGGS];   /* tile slot of each aggregate input */
    int32_t         map_nt;        /* slot-map gathers as non-temporal loads */
};

template <int R, typename MT>
__global__ __launch_bounds__(512) void k_jg_tile(JGTile A, GroupPlan P, uint64_t* partials) {
    const MT* slotmap = (const MT*)A.slotmap;
    extern __shared__ uint64_t tab[];
    const int64_t ncell_total = A.n_slots * P.ncells;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x)
        tab[i] = cell_identity(P.kind[i % P.ncells]);
    __syncthreads();
    const int lane = lane_id();
    const int64_t n_tiles = (A.nl + 64 * R - 1) / (64 * R);
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t t = wave; t < n_tiles; t += n_waves) {
        const int64_t base = t * 64 * R + (int64_t)lane * R;
        int64_t cv[TL_NC][R];
        load_tile<R>(A.T, base, A.nl, cv);
        uint32_t m = tile_live<R>(A.lsel, base, A.nl);
        if (A.has_pred) m &= pred_tile<R>(A.pred, A.T, cv);
        int32_t s[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            s[j] = -1;
            uint64_t i = (uint64_t)(cget<R>(cv, A.key_c, j) - A.kmin);
            if (((m >> j) & 1) && i < (uint64_t)A.range) {
                MT v = A.map_nt ? __builtin_nontemporal_load(slotmap + i) : slotmap[i];
                s[j] = v == (MT)~(MT)0 ? -1 : (int32_t)v;
            }
        }
        /* cell-major: the kind switch is uniform, the row loop unrolls */
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0) atomicAdd((unsigned long long*)&tab[(int64_t)s[j] * P.ncells], 1ull);
        for (int i = 1; i < P.ncells; i++) {
            const int cs = A.agg_c[P.col[i]];
            const bool f = A.T.col[cs].type == TD_F64;
            switch (P.kind[i]) {
            case C_ADD_F64:
#pragma unroll
                for (int j = 0; j < R; j++) {
                    if (s[j] < 0) continue;
                    int64_t raw = cget<R>(cv, cs, j);
                    atomicAdd((double*)&tab[(int64_t)s[j] * P.ncells + i],
                              f ? __longlong_as_double(raw) : (double)raw);
                }
                break;
            case C_SQ_F64:
#pragma unroll
                for (int j = 0; j < R; j++) {
                    if (s[j] < 0) continue;
                    int64_t raw = cget<R>(cv, cs, j);
                    double v = f ? __longlong_as_double(raw) : (double)raw;
                    atomicAdd((double*)&tab[(int64_t)s[j] * P.ncells + i], v * v);
                }
                break;
            case C_ADD_I64:
#pragma unroll
                for (int j = 0; j < R; j++)
                    if (s[j] >= 0)
                        atomicAdd((unsigned long long*)&tab[(int64_t)s[j] * P.ncells + i],
                                  (unsigned long long)cget<R>(cv, cs, j));
                break;
            case C_MIN_I64: case C_MAX_I64: case C_MIN_F64: case C_MAX_F64: {
                const bool is_min = P.kind[i] == C_MIN_I64 || P.kind[i] == C_MIN_F64;
                const bool ord = P.kind[i] == C_MIN_F64 || P.kind[i] == C_MAX_F64;
#pragma unroll
                for (int j = 0; j < R; j++) {
                    if (s[j] < 0) continue;
                    int64_t raw = cget<R>(cv, cs, j);
                    long long v = ord ? (long long)f64_ord(f ? __longlong_as_double(raw) : (double)raw)
                                      : (long long)raw;
                    long long* c = (long long*)&tab[(int64_t)s[j] * P.ncells + i];
                    if (is_min) atomicMin(c, v);
                    else atomicMax(c, v);
                }
                break;
            }
            default: break;              /* row cells (FIRST / LAST) are not fused */
            }
        }
    }
    __syncthreads();
    uint64_t* dst = partials + (int64_t)blockIdx.x * ncell_total;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x) dst[i] = tab[i];
}

/* Star-shaped slot-map probe: key in slot 0, a range-check predicate on
 * slots 1..NPRED, aggregate inputs reached by a uniform switch per cell. */
#define STAR_R 4          /* rows per lane per tile, star probe */

struct JGStar {
    TileCols        T;
    StarPred        SP;
    const uint64_t* lsel;
    int64_t         nl;
    int64_t         kmin, range;
    const void*     slotmap;       /* MT entries: u16 (EMPTY 0xffff) or i32 (EMPTY -1) */
    int64_t         n_slots;
    uint8_t         cell_slot[GP_MAX_CELLS];
};

template <int R, typename MT, int NPRED>
__global__ __launch_bounds__(512) void k_jg_star(JGStar A, GroupPlan P, uint64_t* partials) {
    extern __shared__ uint64_t tab[];
    const MT* slotmap = (const MT*)A.slotmap;
    const int64_t ncell_total = A.n_slots * P.ncells;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x)
        tab[i] = cell_identity(P.kind[i % P.ncells]);
    __syncthreads();
    const int lane = lane_id();
    const int64_t n_tiles = (A.nl + 64 * R - 1) / (64 * R);
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t t = wave; t < n_tiles; t += n_waves) {
        const int64_t base = t * 64 * R + (int64_t)lane * R;
        int64_t cv[TL_NC][R];
        load_tile<R>(A.T, base, A.nl, cv);
        const uint32_t m = tile_live<R>(A.lsel, base, A.nl) & star_pred<R, NPRED>(A.SP, A.T, cv);
        int32_t s[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            s[j] = -1;
            uint64_t i = (uint64_t)(cv[0][j] - A.kmin);
            if (((m >> j) & 1) && i < (uint64_t)A.range) {
                MT v = slotmap[i];
                s[j] = v == (MT)~(MT)0 ? -1 : (int32_t)v;
            }
        }
        acc_tile<R>(tab, P, A.cell_slot, A.T, cv, s);
    }
    __syncthreads();
    uint64_t* dst = partials + (int64_t)blockIdx.x * ncell_total;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x) dst[i] = tab[i];
}

/* Host: star plan of a slot-map probe (probe-side aggregates only). */
static bool jg_star_plan(const tdd_join_group_spec_t* S, const GroupPlan& P, JGStar& A) {
    memset(&A, 0, sizeof(A));
    if (!star_plan(S->join.lkeys[0], S->lpred, A.T, A.SP)) return false;
    int agg_slot[TDD_MAX_AGGS];
    for (int a = 0; a < P.n_aggs; a++) {
        agg_slot[a] = 0;
        if (!P.aggs[a].data) continue;
        if (P.agg_side[a]) return false;
        int c = tile_slot(A.T, P.aggs[a]);
        if (c < 0) return false;
        agg_slot[a] = c;
    }
    for (int i = 1; i < P.ncells; i++) A.cell_slot[i] = (uint8_t)agg_slot[P.col[i]];
    return true;
}

/* Tile description of a slot-map probe, when every probe-side column fits
 * the tile cache and the predicate's stack fits R-bit entries. */
template <int R>
static bool jg_tile_plan(const tdd_join_group_spec_t* S, const GroupPlan& P, JGTile& A) {
    memset(&A, 0, sizeof(A));
    int kc = tile_slot(A.T, S->join.lkeys[0]);
    if (kc < 0) return false;
    A.key_c = kc;
    if (S->lpred) {
        if (pred_depth(*S->lpred) * R > 64) return false;
        A.pred = *S->lpred;
        for (int i = 0; i < A.pred.n_ins; i++) {
            tdd_pred_ins_t& in = A.pred.ins[i];
            if (in.kind == TDD_P_CMP_CONST || in.kind == TDD_P_CMP_COL || in.kind == TDD_P_BOOLCOL) {
                int a = tile_slot(A.T, S->lpred->cols[in.a]);
                if (a < 0) return false;
                in.a = (uint8_t)a;
            }
            if (in.kind == TDD_P_CMP_COL) {
                int b = tile_slot(A.T, S->lpred->cols[in.b]);
                if (b < 0) return false;
                in.b = (uint8_t)b;
            }
        }
        A.has_pred = 1;
    }
    for (int a = 0; a < P.n_aggs; a++) {
        if (!P.aggs[a].data) continue;             /* COUNT without a column */
        if (P.agg_side[a]) return false;
        int c = tile_slot(A.T, P.aggs[a]);
        if (c < 0) return false;
        A.agg_c[a] = (uint8_t)c;
    }
    return true;
}

static int jg_compose(const tdd_join_group_spec_t* S, tdd_group_result_t* out);
int jg_partitioned(const tdd_join_group_spec_t* S, const GroupPlan& GP, int64_t kmin,
                   int64_t range, int64_t n_slots, bool forced, tdd_group_result_t* out);
int jg_xpart(const tdd_join_group_spec_t* S, const GroupPlan& GP, const TileCols& T0,
             const StarPred& SP, const uint8_t* cell_slot, const uint16_t* map, int64_t kmin,
             int64_t range, int64_t n_slots, tdd_group_result_t* out);

/* The fused join -> group-by.  `speculate`: build the slot map and probe
 * it without waiting for the build keys' uniqueness check, which is read
 * after the result is out; repeated keys then discard that result and run
 * the query again without speculation. */
static int jg_fused(const tdd_join_group_spec_t* S, tdd_group_result_t* out, bool speculate) {
    memset(out, 0, sizeof(*out));
    POLL_CANCEL();
    const tdd_join_spec_t* J = &S->join;
    bool fusable = J->join_type == 0 && J->n_keys == 1 && is_int(J->lkeys[0].type) &&
                   is_int(J->rkeys[0].type) && J->force_path != 2 && S->n_keys >= 1 &&
                   S->n_keys <= TDD_MAX_KEYS && S->n_aggs <= TDD_MAX_AGGS &&
                   J->nr < ((int64_t)1 << 31);
    for (int a = 0; a < S->n_aggs && fusable; a++)
        fusable = S->ops[a] != OP_FIRST && S->ops[a] != OP_LAST;
    for (int k = 0; k < S->n_keys && fusable; k++)
        fusable = is_int(S->keys[k].type);
    if (!fusable) return jg_compose(S, out);

    GroupPlan P;
    memset(&P, 0, sizeof(P));
    P.n_keys = S->n_keys;
    P.n_aggs = S->n_aggs;
    for (int k = 0; k < P.n_keys; k++) { P.keys[k] = S->keys[k]; P.key_side[k] = S->key_side[k]; }
    for (int a = 0; a < P.n_aggs; a++) {
        P.aggs[a] = S->aggs[a];
        P.ops[a] = S->ops[a];
        P.agg_side[a] = S->agg_side[a];
    }
    int rc = plan_cells(P);
    if (rc) return rc;

    /* Build-side key range: dense enough for a direct map? */
    int64_t mn, mx;
    rc = tdd_minmax(&J->rkeys[0], J->nr, nullptr, &mn, &mx);
    if (rc) return rc;
    if (mn > mx) { mn = 0; mx = 0; }
    int64_t range = (int64_t)((uint64_t)mx - (uint64_t)mn + 1);
    if (range <= 0 || range > ((int64_t)1 << 31) ||
        (range > 4 * J->nr + 4096 && J->force_path != 1))
        return jg_compose(S, out);

    /* Group slot space from the source columns' ranges (a superset of the
     * joined rows' ranges; empty slots are dropped at emit). */
    double slots = 1.0;
    for (int k = 0; k < P.n_keys; k++) {
        int64_t kmn, kmx;
        int64_t n = P.key_side[k] ? J->nr : J->nl;
        rc = tdd_minmax(&P.keys[k], n, P.key_side[k] ? nullptr : J->lsel, &kmn, &kmx);
        if (rc) return rc;
        if (kmn > kmx) { kmn = 0; kmx = 0; }
        P.kmin[k] = kmn;
        P.krange[k] = (int64_t)((uint64_t)kmx - (uint64_t)kmn + 1);
        if (P.krange[k] <= 0) return jg_compose(S, out);
        slots *= (double)P.krange[k];
    }
    if (slots > (double)g_da_max_slots || slots * P.ncells * 8 > 8.0e9) return jg_compose(S, out);
    int64_t st = 1;
    for (int k = P.n_keys - 1; k >= 0; k--) { P.kstride[k] = st; st *= P.krange[k]; }
    int64_t n_slots = (int64_t)slots;

    bool right_keys_only = true, left_aggs_only = true;
    for (int k = 0; k < P.n_keys; k++) right_keys_only = right_keys_only && P.key_side[k] == 1;
    for (int a = 0; a < P.n_aggs; a++) left_aggs_only = left_aggs_only && P.agg_side[a] == 0;
    bool slotmap = right_keys_only && left_aggs_only && n_slots < ((int64_t)1 << 31);

    size_t table_bytes = (size_t)n_slots * P.ncells * 8;
    const bool lds_table = table_bytes <= 64 * 1024;

    /* Large build key ranges: radix-partitioned probe with per-CU LDS
     * slices of the slot map (force_path 3 takes it at any size). */
    if (slotmap && (J->force_path == 0 || J->force_path == 3)) {
        rc = jg_partitioned(S, P, mn, range, n_slots, J->force_path == 3, out);
        if (rc != TD_ERR_NYI) return rc;
        memset(out, 0, sizeof(*out));
    }
    JGTile TA;
    JGStar SA;
    const int TR = TUNE_INT("TEIDE_TILE_ROWS", 8) == 4 ? 4 : 8;      /* rows per lane */
    const int TB = TUNE_INT("TEIDE_TILE_BLOCK", 512) == 256 ? 256 : 512;
    bool star = slotmap && lds_table && jg_star_plan(S, P, SA);
    bool tiled = slotmap && lds_table && !star &&
                 (TR == 8 ? jg_tile_plan<8>(S, P, TA) : jg_tile_plan<4>(S, P, TA));
    const bool small = (tiled || star) && n_slots < 0xffff;      /* u16 slot map */

    /* Slot map for unique build keys; duplicates fall back to chains. */
    Scratch smap(slotmap ? (size_t)range * (small ? 2 : 4) : 8);
    SCRATCH(smcnt, 8);
    if (!smap.ok()) return TD_ERR_OOM;
    const bool deferred = speculate && slotmap && (star || tiled) && lds_table &&
                          J->force_path != 4;
    if (slotmap) {
        int uniq = 0;
        int64_t* dc = deferred ? smcnt.as<int64_t>() : nullptr;
        rc = small ? sm_build<uint16_t>(J->rkeys[0], J->nr, mn, range, P, smap.as<uint16_t>(), &uniq, dc)
                   : sm_build<int32_t>(J->rkeys[0], J->nr, mn, range, P, smap.as<int32_t>(), &uniq, dc);
        if (rc == TD_ERR_NYI) rc = TD_OK;
        if (rc) return rc;
        slotmap = uniq != 0;
        tiled = tiled && slotmap;
        star = star && slotmap;
    }
    Scratch head(slotmap ? 8 : (size_t)range * 4);
    Scratch next(slotmap ? 8 : (size_t)(J->nr > 0 ? J->nr : 1) * 4);
    if (!head.ok() || !next.ok()) return TD_ERR_OOM;
    if (!slotmap) {
        SCRATCH(dup, 8);
        rc = tdd_fill(head.p, ~0ull, (size_t)range * 4);
        if (!rc) rc = tdd_fill(dup.p, 0, 8);
        if (rc) return rc;
        if (J->nr > 0)
            LAUNCH("jg_build", k_dm_build, stream_grid(ceil_div(J->nr, 256 * 4)), 256, 0,
                   J->rkeys[0], J->nr, mn, head.as<int32_t>(), next.as<int32_t>(), dup.as<int>());
    }

    JGArgs A;
    memset(&A, 0, sizeof(A));
    if (S->lpred) { A.pred = *S->lpred; A.has_pred = 1; }
    A.lsel = J->lsel ? J->lsel->bits : nullptr;
    A.nl = J->nl;
    A.lk = J->lkeys[0];
    A.D = DMap{head.as<int32_t>(), next.as<int32_t>(), mn, range};
    A.n_slots = n_slots;
    if (slotmap) A.slotmap = smap.as<int32_t>();

    /* XCD-partitioned probe: keeps each XCD's gathers inside its L2 at the
     * cost of writing and re-reading the passing rows; it measures level
     * with the single-pass star probe on 200 M x 16 M, so only a forced
     * path takes it. */
    if (star && small && lds_table && J->force_path == 4) {
        rc = jg_xpart(S, P, SA.T, SA.SP, SA.cell_slot, smap.as<uint16_t>(), mn, range, n_slots, out);
        if (rc != TD_ERR_NYI) return rc;
        memset(out, 0, sizeof(*out));
    }

    int64_t n_words = (J->nl + 63) / 64;
    if (lds_table) {
        int per_cu = (int)(160 * 1024 / (table_bytes + 1024));
        if (per_cu > 4) per_cu = 4;
        if (per_cu < 1) per_cu = 1;
        int grid = stream_grid(ceil_div(n_words > 0 ? n_words : 1, 8), per_cu);
        const int np = SA.SP.npred;
        if (star) {
            int64_t tiles = ceil_div(J->nl > 0 ? J->nl : 1, 64 * STAR_R);
            int64_t units = ceil_div(tiles, 512 / WAVE);
            grid = resident_grid(k_jg_star<STAR_R, int32_t, 2>, 512, table_bytes, units);
        } else if (tiled) {
            int64_t tiles = ceil_div(J->nl > 0 ? J->nl : 1, 64 * TR);
            int64_t units = ceil_div(tiles, TB / WAVE);
            grid = TR == 8 ? (small ? resident_grid(k_jg_tile<8, uint16_t>, TB, table_bytes, units)
                                    : resident_grid(k_jg_tile<8, int32_t>, TB, table_bytes, units))
                           : (small ? resident_grid(k_jg_tile<4, uint16_t>, TB, table_bytes, units)
                                    : resident_grid(k_jg_tile<4, int32_t>, TB, table_bytes, units));
        }
        SCRATCH(parts, (size_t)grid * table_bytes);
        if (star) {
            SA.lsel = A.lsel;
            SA.nl = J->nl;
            SA.kmin = mn;
            SA.range = range;
            SA.slotmap = smap.p;
            SA.n_slots = n_slots;
            uint64_t* pp = parts.as<uint64_t>();
#define STAR_LAUNCH(MT, NP) \
            LAUNCH("jg_probe_star", (k_jg_star<STAR_R, MT, NP>), grid, 512, table_bytes, SA, P, pp)
            if (small) {
                if (np == 0) STAR_LAUNCH(uint16_t, 0);
                else if (np == 1) STAR_LAUNCH(uint16_t, 1);
                else STAR_LAUNCH(uint16_t, 2);
            } else {
                if (np == 0) STAR_LAUNCH(int32_t, 0);
                else if (np == 1) STAR_LAUNCH(int32_t, 1);
                else STAR_LAUNCH(int32_t, 2);
            }
#undef STAR_LAUNCH
        } else if (tiled) {
            TA.lsel = A.lsel;
            TA.nl = J->nl;
            TA.kmin = mn;
            TA.range = range;
            TA.slotmap = smap.p;
            TA.n_slots = n_slots;
            TA.map_nt = TUNE_INT("TEIDE_MAP_NT", 0);
            uint64_t* pp = parts.as<uint64_t>();
            if (TR == 8 && small)
                LAUNCH("jg_probe_tile", (k_jg_tile<8, uint16_t>), grid, TB, table_bytes, TA, P, pp);
            else if (TR == 8)
                LAUNCH("jg_probe_tile", (k_jg_tile<8, int32_t>), grid, TB, table_bytes, TA, P, pp);
            else if (small)
                LAUNCH("jg_probe_tile", (k_jg_tile<4, uint16_t>), grid, TB, table_bytes, TA, P, pp);
            else
                LAUNCH("jg_probe_tile", (k_jg_tile<4, int32_t>), grid, TB, table_bytes, TA, P, pp);
        } else if (slotmap) {
            LAUNCH("jg_probe_lds", k_jg_lds<true>, grid, 512, table_bytes, A, P, parts.as<uint64_t>());
        } else {
            LAUNCH("jg_probe_lds", k_jg_lds<false>, grid, 512, table_bytes, A, P, parts.as<uint64_t>());
        }
        rc = da_finish(P, parts.as<uint64_t>(), grid, n_slots, out);
        out->path = 1;
        if (!rc && deferred) {
            int64_t cnt = 0;
            HC(hipMemcpyAsync(&cnt, smcnt.p, 8, hipMemcpyDeviceToHost, tdd::g_stream));
            HC(hipStreamSynchronize(tdd::g_stream));
            if (cnt != J->nr) {                  /* repeated build keys: the map was wrong */
                for (int k = 0; k < P.n_keys; k++) td__arena_free(out->keys[k]);
                for (int a = 0; a < P.n_aggs; a++) td__arena_free(out->vals[a]);
                return jg_fused(S, out, false);
            }
        }
    } else {
        SCRATCH(tab, table_bytes);
        rc = da_table_init(P, tab.as<uint64_t>(), n_slots);
        if (rc) return rc;
        int grid = stream_grid(ceil_div(n_words > 0 ? n_words : 1, 8), 4);
        if (slotmap)
            LAUNCH("jg_probe_global", k_jg_global<true>, grid, 512, 0, A, P, tab.as<uint64_t>());
        else
            LAUNCH("jg_probe_global", k_jg_global<false>, grid, 512, 0, A, P, tab.as<uint64_t>());
        rc = da_emit(P, tab.as<uint64_t>(), n_slots, out);
        out->path = 2;
    }
    if (rc) return rc;
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}

extern "C" int tdd_join_group(const tdd_join_group_spec_t* S, tdd_group_result_t* out) {
    return jg_fused(S, out, true);
}

/* General form on the device: filter to a selection, join to index pairs,
 * gather the key and aggregate columns of the joined rows, group them. */
/* Join -> group-by through the materialising radix join: the join writes
 * only the columns the group-by reads (each distinct column once), which
 * are then grouped; no pair arrays, no gathers.  TD_ERR_NYI when that join
 * does not take the query (the caller joins through pairs). */
static int jg_materialised(const tdd_join_group_spec_t* S, const tdd_join_spec_t* J,
                           tdd_group_result_t* out) {
    if (J->join_type != 0) return TD_ERR_NYI;
    tdd_join_mat_spec_t M;
    memset(&M, 0, sizeof(M));
    M.join = *J;
    int which[TDD_MAX_KEYS + TDD_MAX_AGGS];            /* side * 16 + column, -1: none */
    for (int i = 0; i < S->n_keys + S->n_aggs; i++) {
        const bool is_key = i < S->n_keys;
        const tdd_col_t& c = is_key ? S->keys[i] : S->aggs[i - S->n_keys];
        const int side = is_key ? S->key_side[i] : S->agg_side[i - S->n_keys];
        which[i] = -1;
        if (!c.data) continue;
        tdd_col_t* cols = side ? M.rcols : M.lcols;
        int32_t* n = side ? &M.n_rcols : &M.n_lcols;
        int at = -1;
        for (int j = 0; j < *n; j++)
            if (cols[j].data == c.data && cols[j].type == c.type && cols[j].attrs == c.attrs) at = j;
        if (at < 0) {
            if (*n >= 16) return TD_ERR_NYI;
            cols[*n] = c;
            at = (*n)++;
        }
        which[i] = side * 16 + at;
    }
    tdd_join_mat_result_t R;
    int rc = tdd_join_mat(&M, &R);
    if (rc == TD_ERR_OOM && J->force_path == 0) return TD_ERR_NYI;   /* pairs need less */
    if (rc) return rc;
    struct Free {
        tdd_join_mat_result_t* r; int nl, nr;
        ~Free() {
            for (int c = 0; c < nl; c++) td__arena_free(r->lcols[c]);
            for (int c = 0; c < nr; c++) td__arena_free(r->rcols[c]);
        }
    } fr{&R, M.n_lcols, M.n_rcols};
    tdd_group_spec_t G;
    memset(&G, 0, sizeof(G));
    G.nrows = R.n_rows;
    G.n_keys = S->n_keys;
    G.n_aggs = S->n_aggs;
    for (int i = 0; i < S->n_keys + S->n_aggs; i++) {
        const bool is_key = i < S->n_keys;
        tdd_col_t c = is_key ? S->keys[i] : S->aggs[i - S->n_keys];
        if (which[i] >= 0) {
            /* the source's bounds still hold for its joined copy */
            c.data = which[i] >= 16 ? R.rcols[which[i] - 16] : R.lcols[which[i]];
        }
        if (is_key) G.keys[i] = c;
        else { G.aggs[i - S->n_keys] = c; G.ops[i - S->n_keys] = S->ops[i - S->n_keys]; }
    }
    rc = tdd_group(&G, out);
    if (!rc) out->path += 10;                    /* reported: materialised join, then group */
    return rc;
}

static int jg_compose(const tdd_join_group_spec_t* S, tdd_group_result_t* out) {
    tdd_join_spec_t J = S->join;
    int64_t n_segs = (J.nl + 1023) / 1024;
    size_t sel_bytes = 16 + (size_t)((n_segs + 7) & ~7) + (size_t)((n_segs * 2 + 7) & ~7) +
                       (size_t)((J.nl + 63) / 64) * 8;
    Scratch selbuf(S->lpred ? sel_bytes : 8);
    if (!selbuf.ok()) return TD_ERR_OOM;
    tdd_sel_t sel;
    if (S->lpred) {
        uint8_t* b = selbuf.as<uint8_t>();
        sel.base = b;
        sel.flags = b + 16;
        sel.popcnt = (uint16_t*)(b + 16 + ((n_segs + 7) & ~7));
        sel.bits = (uint64_t*)(b + 16 + ((n_segs + 7) & ~7) + ((n_segs * 2 + 7) & ~7));
        sel.nrows = J.nl;
        sel.n_segs = (uint32_t)n_segs;
        int rc = tdd_sel_eval(S->lpred, J.lsel, &sel);
        if (rc) return rc;
        J.lsel = &sel;
    }
    int rc = jg_materialised(S, &J, out);
    if (rc != TD_ERR_NYI) return rc;
    memset(out, 0, sizeof(*out));
    tdd_join_result_t jr;
    rc = tdd_join(&J, &jr);
    if (rc) return rc;
    struct PairFree { tdd_join_result_t* r; ~PairFree() { td__arena_free(r->l_idx); td__arena_free(r->r_idx); } } pf{&jr};
    tdd_group_spec_t G;
    memset(&G, 0, sizeof(G));
    G.nrows = jr.n_pairs;
    G.n_keys = S->n_keys;
    G.n_aggs = S->n_aggs;
    void* bufs[TDD_MAX_KEYS + TDD_MAX_AGGS] = {};
    int nb = 0;
    struct BufFree { void** b; int* n; ~BufFree() { for (int i = 0; i < *n; i++) td__arena_free(b[i]); } } bf{bufs, &nb};
    for (int i = 0; i < S->n_keys + S->n_aggs; i++) {
        bool is_key = i < S->n_keys;
        const tdd_col_t& src = is_key ? S->keys[i] : S->aggs[i - S->n_keys];
        int side = is_key ? S->key_side[i] : S->agg_side[i - S->n_keys];
        tdd_col_t c = src;
        if (src.data) {
            void* d = td__arena_alloc((size_t)(jr.n_pairs > 0 ? jr.n_pairs : 1) * src.esz);
            if (!d) return TD_ERR_OOM;
            bufs[nb++] = d;
            rc = tdd_gather(&src, side ? jr.r_idx : jr.l_idx, jr.n_pairs, 0, d);
            if (rc) return rc;
            c.data = d;
        }
        if (is_key) G.keys[i] = c;
        else { G.aggs[i - S->n_keys] = c; G.ops[i - S->n_keys] = S->ops[i - S->n_keys]; }
    }
    return tdd_group(&G, out);
}
