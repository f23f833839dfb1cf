/*
 * k_select.hip -- predicate evaluation into TD_SEL bitmaps, selection
 * algebra, stream compaction and gathers.
 *
 * A TD_SEL holds one bit per row, LSB-first in 64-bit words, which is exactly
 * what a wave64 ballot returns: lane i of a wave evaluates row 64*w + i and
 * the ballot is word w.  A workgroup of 256 lanes (4 waves) covers one
 * 1024-row segment per step, four words per wave, and derives the segment's
 * popcount and NONE / ALL / MIX flag from the ballots it already holds.
 */
#include "pred_eval.h"
#include "sel_map.h"

#define SEG_ROWS 1024
#define SEL_T 256

/* One 1024-row segment per step: 4 waves x 4 ballots. */
__global__ __launch_bounds__(SEL_T) void k_sel_eval(tdd_pred_t P, int64_t nrows,
                                                    uint32_t n_segs, const uint64_t* and_bits,
                                                    uint8_t* flags, uint16_t* pop,
                                                    uint64_t* bits,
                                                    unsigned long long* total) {
    __shared__ uint32_t wcnt[SEL_T / WAVE];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t n_words = (nrows + 63) >> 6;
    unsigned long long my_total = 0;
    for (int64_t seg = blockIdx.x; seg < n_segs; seg += gridDim.x) {
        uint32_t cnt = 0;
        const int64_t r0 = (seg * 16 + w * 4) * 64 + lane;
        const uint32_t pm = pred_eval4(P, r0, nrows);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int64_t word = seg * 16 + w * 4 + k;
            int64_t row = word * 64 + lane;
            bool p = row < nrows && ((pm >> k) & 1);
            uint64_t m = __ballot(p);
            if (word < n_words) {
                if (and_bits) m &= and_bits[word];
                if (lane == 0) bits[word] = m;
            }
            cnt += __popcll(m);
        }
        if (lane == 0) wcnt[w] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t c = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
            int64_t seg_rows = nrows - seg * SEG_ROWS;
            if (seg_rows > SEG_ROWS) seg_rows = SEG_ROWS;
            pop[seg] = (uint16_t)c;
            flags[seg] = c == 0 ? TD_SEL_NONE : (c == (uint32_t)seg_rows ? TD_SEL_ALL : TD_SEL_MIX);
            my_total += c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && my_total) atomicAdd(total, my_total);
}

/* Word-parallel AND + recompute: 16 lanes per segment, one word each; the
 * popcount masks bits past nrows as td_sel_recompute does (sel.c:95-102). */
__global__ __launch_bounds__(SEL_T) void k_sel_words(const uint64_t* a, const uint64_t* b,
                                                     int64_t nrows, uint32_t n_segs,
                                                     uint8_t* flags, uint16_t* pop,
                                                     uint64_t* out,
                                                     unsigned long long* total) {
    const int64_t n_words = (nrows + 63) >> 6;
    int64_t seg = (int64_t)blockIdx.x * (SEL_T / 16) + threadIdx.x / 16;
    int sub = threadIdx.x & 15;
    int64_t word = seg * 16 + sub;
    uint32_t pc = 0;
    if (seg < n_segs && word < n_words) {
        uint64_t m = a[word];
        if (b) m &= b[word];
        out[word] = m;
        if (word == n_words - 1 && (nrows & 63)) m &= (1ull << (nrows & 63)) - 1;
        pc = __popcll(m);
    }
    for (int d = 8; d >= 1; d >>= 1) pc += __shfl_xor(pc, d, 16);
    if (sub == 0 && seg < n_segs) {
        int64_t seg_rows = nrows - seg * SEG_ROWS;
        if (seg_rows > SEG_ROWS) seg_rows = SEG_ROWS;
        pop[seg] = (uint16_t)pc;
        flags[seg] = pc == 0 ? TD_SEL_NONE : (pc == (uint32_t)seg_rows ? TD_SEL_ALL : TD_SEL_MIX);
        if (pc) atomicAdd(total, (unsigned long long)pc);
    }
}

/* meta: total_pass (i64) then n_segs (u32) and padding; the total is
 * accumulated by atomics so it starts from zero.  The flag and popcount
 * arrays' padding is zeroed too, as td_sel_new zeroes the whole payload
 * (sel.c:52-69); the arrays themselves are overwritten by the caller. */
__global__ void k_sel_meta(uint64_t* meta, uint32_t n_segs, int64_t head_words) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= head_words) return;
    meta[i] = i == 1 ? (uint64_t)n_segs : 0;
}

static int write_total_start(const tdd_sel_t* s) {
    int64_t head_words = ((char*)s->bits - (char*)s->base) / 8;
    LAUNCH("sel_meta", k_sel_meta, (int)ceil_div(head_words, 256), 256, 0, (uint64_t*)s->base,
           s->n_segs, head_words);
    return TD_OK;
}

extern "C" int tdd_sel_eval(const tdd_pred_t* pred, const tdd_sel_t* and_with, tdd_sel_t* out) {
    if (pred->n_ins < 1 || pred->n_ins > TDD_PRED_MAX_INS) return TD_ERR_RANGE;
    if (and_with && and_with->nrows != out->nrows) return TD_ERR_LENGTH;
    int rc = write_total_start(out);
    if (rc) return rc;
    if (out->n_segs == 0) return TD_OK;
    int grid = stream_grid(out->n_segs, 8);
    LAUNCH("sel_eval", k_sel_eval, grid, SEL_T, 0, *pred, out->nrows, out->n_segs,
           (const uint64_t*)(and_with ? and_with->bits : nullptr), out->flags, out->popcnt,
           out->bits, (unsigned long long*)out->base);
    return TD_OK;
}

extern "C" int tdd_sel_from_bool(const uint8_t* bools, tdd_sel_t* out) {
    tdd_pred_t p;
    memset(&p, 0, sizeof(p));
    p.n_ins = 1;
    p.n_cols = 1;
    p.cols[0].data = bools;
    p.cols[0].type = TD_U8;
    p.cols[0].esz = 1;
    p.ins[0].kind = TDD_P_BOOLCOL;
    return tdd_sel_eval(&p, nullptr, out);
}

static int sel_words(const uint64_t* a, const uint64_t* b, tdd_sel_t* out) {
    int rc = write_total_start(out);
    if (rc) return rc;
    if (out->n_segs == 0) return TD_OK;
    int grid = (int)ceil_div(out->n_segs, SEL_T / 16);
    LAUNCH("sel_words", k_sel_words, grid, SEL_T, 0, a, b, out->nrows, out->n_segs,
           out->flags, out->popcnt, out->bits, (unsigned long long*)out->base);
    return TD_OK;
}

extern "C" int tdd_sel_and(const tdd_sel_t* a, const tdd_sel_t* b, tdd_sel_t* out) {
    if (a->nrows != b->nrows || a->nrows != out->nrows) return TD_ERR_LENGTH;
    return sel_words(a->bits, b->bits, out);
}

extern "C" int tdd_sel_recompute(tdd_sel_t* s) {
    return sel_words(s->bits, nullptr, s);
}

extern "C" int tdd_sel_total(const tdd_sel_t* s, int64_t* total) {
    HC(hipMemcpyAsync(total, s->base, 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}

/* ------------------------------------------------------------ compaction -- */
/* seg_off[s] = selected rows before segment s (scan of the popcounts).  For
 * each segment the four waves first publish their words' popcounts, then
 * every selected lane writes to seg_off + (bits in earlier words) + (bits of
 * its own word in lower lanes). */

#define MAXC 16
struct GatherCols {
    const void* src[MAXC];
    void*       dst[MAXC];
    uint8_t     esz[MAXC];
    int32_t     n;
};

__device__ __forceinline__ void copy_elem(const GatherCols& G, int c, int64_t from, int64_t to) {
    switch (G.esz[c]) {
    case 1: ((uint8_t*)G.dst[c])[to] = ((const uint8_t*)G.src[c])[from]; break;
    case 2: ((uint16_t*)G.dst[c])[to] = ((const uint16_t*)G.src[c])[from]; break;
    case 4: ((uint32_t*)G.dst[c])[to] = ((const uint32_t*)G.src[c])[from]; break;
    case 8: ((uint64_t*)G.dst[c])[to] = ((const uint64_t*)G.src[c])[from]; break;
    default: {
        const uint8_t* s = (const uint8_t*)G.src[c] + from * G.esz[c];
        uint8_t* d = (uint8_t*)G.dst[c] + to * G.esz[c];
        for (int b = 0; b < G.esz[c]; b++) d[b] = s[b];
    }
    }
}

template <class T>
__device__ __forceinline__ void compact_col(const void* src, void* dst, const uint64_t* m,
                                            const int64_t* to, int64_t r0, int lane) {
    T v[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if ((m[k] >> lane) & 1) v[k] = ((const T*)src)[r0 + 64 * k];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if ((m[k] >> lane) & 1) ((T*)dst)[to[k]] = v[k];
}

/* Dense right column through the key map: the value at key - kmin where
 * the key is present, else 0. */
template <class T>
__device__ __forceinline__ void map_col(const void* src, void* dst, const uint64_t* m,
                                        const int64_t* to, const int64_t* at, const bool* hit,
                                        int lane) {
    T v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = hit[k] ? ((const T*)src)[at[k]] : (T)0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if ((m[k] >> lane) & 1) ((T*)dst)[to[k]] = v[k];
}

/* mode 0: gather columns (and the map's right columns); mode 1: write row
 * ids into idx.  Columns are copied one at a time, four rows per lane with
 * their loads in flight together. */
__global__ __launch_bounds__(SEL_T) void k_compact(const uint64_t* bits, const uint8_t* flags,
                                                   const int64_t* seg_off, int64_t nrows,
                                                   uint32_t n_segs, GatherCols G,
                                                   TddMapGather M, int64_t* idx, int mode) {
    __shared__ uint32_t wpc[SEL_T / WAVE];
    const int w = threadIdx.x / WAVE, lane = lane_id();
    const int64_t n_words = (nrows + 63) >> 6;
    for (int64_t seg = blockIdx.x; seg < n_segs; seg += gridDim.x) {
        if (flags[seg] == TD_SEL_NONE) continue;       /* uniform across the block */
        uint64_t m[4];
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int64_t word = seg * 16 + w * 4 + k;
            m[k] = word < n_words ? bits[word] : 0;
            if (word == n_words - 1 && (nrows & 63)) m[k] &= (1ull << (nrows & 63)) - 1;
            c += __popcll(m[k]);
        }
        if (lane == 0) wpc[w] = c;
        __syncthreads();
        int64_t pos = seg_off[seg];
        for (int i = 0; i < w; i++) pos += wpc[i];
        int64_t to[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            to[k] = pos + lanes_below(m[k]);
            pos += __popcll(m[k]);
        }
        const int64_t r0 = (seg * 16 + w * 4) * 64 + lane;
        if (mode == 1) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if ((m[k] >> lane) & 1) idx[to[k]] = r0 + 64 * k;
        } else {
            for (int cc = 0; cc < G.n; cc++) {
                switch (G.esz[cc]) {
                case 8: compact_col<uint64_t>(G.src[cc], G.dst[cc], m, to, r0, lane); break;
                case 4: compact_col<uint32_t>(G.src[cc], G.dst[cc], m, to, r0, lane); break;
                case 2: compact_col<uint16_t>(G.src[cc], G.dst[cc], m, to, r0, lane); break;
                case 1: compact_col<uint8_t>(G.src[cc], G.dst[cc], m, to, r0, lane); break;
                default:
                    for (int k = 0; k < 4; k++)
                        if ((m[k] >> lane) & 1) copy_elem(G, cc, r0 + 64 * k, to[k]);
                }
            }
            if (M.n > 0) {
                int64_t key[4], at[4];
                bool hit[4];
                ld4_i64(M.key, r0, nrows, key);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t i = (uint64_t)key[k] - (uint64_t)M.kmin;
                    hit[k] = ((m[k] >> lane) & 1) && i < (uint64_t)M.range &&
                             (!M.present || ((M.present[i >> 5] >> (i & 31)) & 1));
                    at[k] = hit[k] ? (int64_t)i : 0;
                }
                for (int cc = 0; cc < M.n; cc++) {
                    switch (M.esz[cc]) {
                    case 8: map_col<uint64_t>(M.src[cc], M.dst[cc], m, to, at, hit, lane); break;
                    case 4: map_col<uint32_t>(M.src[cc], M.dst[cc], m, to, at, hit, lane); break;
                    case 2: map_col<uint16_t>(M.src[cc], M.dst[cc], m, to, at, hit, lane); break;
                    default: map_col<uint8_t>(M.src[cc], M.dst[cc], m, to, at, hit, lane); break;
                    }
                }
            }
        }
        __syncthreads();
    }
}

static int compact_impl(const tdd_sel_t* s, const GatherCols& G, const TddMapGather* M,
                        int64_t* idx, int mode) {
    if (s->n_segs == 0) return TD_OK;
    SCRATCH(off, (size_t)(s->n_segs + 1) * 8);
    int rc = tdd::scan_u16_i64(s->popcnt, s->n_segs, off.as<int64_t>());
    if (rc) return rc;
    TddMapGather none;
    memset(&none, 0, sizeof(none));
    int grid = stream_grid(s->n_segs, 8);
    LAUNCH("compact", k_compact, grid, SEL_T, 0, s->bits, s->flags,
           (const int64_t*)off.as<int64_t>(), s->nrows, s->n_segs, G, M ? *M : none, idx, mode);
    return TD_OK;
}

extern "C" int tdd_sel_index(const tdd_sel_t* s, int64_t* idx) {
    GatherCols G;
    memset(&G, 0, sizeof(G));
    return compact_impl(s, G, nullptr, idx, 1);
}

static int fill_cols(GatherCols& G, int ncols, const tdd_col_t* src, void* const* dst) {
    if (ncols < 0 || ncols > MAXC) return TD_ERR_RANGE;
    memset(&G, 0, sizeof(G));
    for (int c = 0; c < ncols; c++) {
        G.src[c] = src[c].data;
        G.dst[c] = dst[c];
        G.esz[c] = src[c].esz;
    }
    G.n = ncols;
    return TD_OK;
}

int tdd__sel_compact_map(const tdd_sel_t* s, int ncols, const tdd_col_t* src, void* const* dst,
                         const TddMapGather* M) {
    GatherCols G;
    int rc = fill_cols(G, ncols, src, dst);
    if (rc) return rc;
    if (M->n < 0 || M->n > TDD_MAP_MAXC) return TD_ERR_RANGE;
    for (int c = 0; c < M->n; c++)
        if (M->esz[c] != 1 && M->esz[c] != 2 && M->esz[c] != 4 && M->esz[c] != 8)
            return TD_ERR_TYPE;
    return compact_impl(s, G, M, nullptr, 0);
}

extern "C" int tdd_sel_compact(const tdd_sel_t* s, int ncols, const tdd_col_t* src,
                               void* const* dst) {
    GatherCols G;
    int rc = fill_cols(G, ncols, src, dst);
    if (rc) return rc;
    return compact_impl(s, G, nullptr, nullptr, 0);
}

/* ---------------------------------------------------------------- gather -- */

template <int E>
__global__ __launch_bounds__(256) void k_gather(const uint8_t* src, const int64_t* idx,
                                                int64_t n, int nullable, int esz,
                                                uint8_t* dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        int64_t r = idx[i];
        if (E == 8) {
            ((uint64_t*)dst)[i] = (r < 0 && nullable) ? 0 : ((const uint64_t*)src)[r];
        } else if (E == 4) {
            ((uint32_t*)dst)[i] = (r < 0 && nullable) ? 0 : ((const uint32_t*)src)[r];
        } else {
            for (int b = 0; b < esz; b++)
                dst[i * esz + b] = (r < 0 && nullable) ? 0 : src[r * esz + b];
        }
    }
}

extern "C" int tdd_gather(const tdd_col_t* src, const int64_t* idx, int64_t n, int nullable,
                          void* dst) {
    if (n <= 0) return TD_OK;
    int grid = stream_grid(ceil_div(n, 256 * 4));
    if (src->esz == 8)
        LAUNCH("gather", k_gather<8>, grid, 256, 0, (const uint8_t*)src->data, idx, n, nullable,
               8, (uint8_t*)dst);
    else if (src->esz == 4)
        LAUNCH("gather", k_gather<4>, grid, 256, 0, (const uint8_t*)src->data, idx, n, nullable,
               4, (uint8_t*)dst);
    else
        LAUNCH("gather", k_gather<0>, grid, 256, 0, (const uint8_t*)src->data, idx, n, nullable,
               (int)src->esz, (uint8_t*)dst);
    return TD_OK;
}
