/*
 * tile.h -- row tiles for the streaming kernels.
 *
 * A wave takes 64*R consecutive rows; lane l owns rows base .. base+R-1 with
 * base = tile*64*R + l*R, so a column's loads for the wave cover one
 * contiguous 64*R*esz-byte span (whole cache lines) and each lane issues its
 * R loads back to back.  Every load of a tile is issued before any value is
 * used, which keeps 64*R*esz bytes per column in flight per wave: that, not
 * the instruction count, is what an HBM-bound scan on MI355X needs.
 *
 * Values are held widened to int64 (F64 as raw bits) in a small register
 * cache `cv[TL_NC][R]`; cget() reads it with a uniform column index through
 * a select chain so the cache never moves to scratch memory.
 */
#pragma once

#include "dev_common.h"
#include "pred_eval.h"

#define TL_NC 4   /* cached columns per tile */

typedef long long tl_v2i64 __attribute__((ext_vector_type(2)));
typedef int tl_v4i32 __attribute__((ext_vector_type(4)));

struct TileCols {
    tdd_col_t col[TL_NC];
    int32_t   n;          /* columns in use */
    uint32_t  vec;        /* bit c: 4- or 8-byte elements at a 16-byte aligned base */
};

/* Widen element r of c (any integer width, F64 raw bits) to int64. */
__device__ __forceinline__ int64_t ld_raw(const tdd_col_t& c, int64_t r) {
    return ld_i64(c, r);          /* F64 columns: ld_i64 returns the raw bits */
}

/* Rows base .. base+R-1 of c (rows >= n read as 0).  `vec`: the column has
 * 4- or 8-byte elements at a 16-byte aligned base (R a multiple of 4, base
 * a multiple of R), so whole tiles load as 16-byte vectors.  The loads are
 * non-temporal: streamed columns are read once, and the caches are left to
 * the lookup tables. */
template <int R>
__device__ __forceinline__ void load_rows(const tdd_col_t& c, int64_t base, int64_t n, bool vec,
                                          int64_t (&v)[R]) {
    if (vec && base + R <= n) {
        if (c.esz == 8) {
            const tl_v2i64* p = (const tl_v2i64*)((const int64_t*)c.data + base);
#pragma unroll
            for (int j = 0; j < R / 2; j++) {
                tl_v2i64 x = __builtin_nontemporal_load(p + j);
                v[2 * j] = x.x;
                v[2 * j + 1] = x.y;
            }
        } else {
            const tl_v4i32* p = (const tl_v4i32*)((const int32_t*)c.data + base);
            const bool uns = c.type == TD_SYM;       /* W32 indices are unsigned */
#pragma unroll
            for (int j = 0; j < R / 4; j++) {
                tl_v4i32 x = __builtin_nontemporal_load(p + j);
                v[4 * j + 0] = uns ? (int64_t)(uint32_t)x.x : (int64_t)x.x;
                v[4 * j + 1] = uns ? (int64_t)(uint32_t)x.y : (int64_t)x.y;
                v[4 * j + 2] = uns ? (int64_t)(uint32_t)x.z : (int64_t)x.z;
                v[4 * j + 3] = uns ? (int64_t)(uint32_t)x.w : (int64_t)x.w;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; j++) v[j] = base + j < n ? ld_raw(c, base + j) : 0;
    }
}

/* Host: whether a column takes the vector path of load_rows. */
static inline bool tile_vec_ok(const tdd_col_t& c) {
    return (c.esz == 8 || c.esz == 4) && ((uintptr_t)c.data & 15) == 0;
}

template <int R>
__device__ __forceinline__ void load_tile(const TileCols& T, int64_t base, int64_t n,
                                          int64_t (&cv)[TL_NC][R]) {
#pragma unroll
    for (int c = 0; c < TL_NC; c++) {
        if (c < T.n) {
            load_rows<R>(T.col[c], base, n, (T.vec >> c) & 1, cv[c]);
        } else {
#pragma unroll
            for (int j = 0; j < R; j++) cv[c][j] = 0;
        }
    }
}

/* cv[c][j] for a wave-uniform c and a compile-time j: the value is picked
 * with integer masks, never with an index into the cache, so the cache stays
 * in registers. */
template <int R>
__device__ __forceinline__ int64_t cget(const int64_t (&cv)[TL_NC][R], int c, int j) {
    uint64_t v = 0;
#pragma unroll
    for (int q = 0; q < TL_NC; q++) v |= (uint64_t)cv[q][j] & (0ull - (uint64_t)(c == q));
    return (int64_t)v;
}

/* A cached value as a double: F64 columns hold raw bits, others integers. */
__device__ __forceinline__ double cval_f64(const TileCols& T, int c, int64_t raw) {
    return T.col[c].type == TD_F64 ? __longlong_as_double(raw) : (double)raw;
}

/* Predicate over the tile's R rows, as an R-bit mask.  `P` is a compiled
 * predicate whose column operands a / b are tile-cache slots; the boolean
 * stack keeps one R-bit mask per entry in a 64-bit register (the host checks
 * that depth * R fits). */
template <int R>
__device__ __forceinline__ uint32_t pred_tile(const tdd_pred_t& P, const TileCols& T,
                                              const int64_t (&cv)[TL_NC][R]) {
    const uint64_t M = (1ull << R) - 1;
    uint64_t st = 0;
    for (int i = 0; i < P.n_ins; i++) {
        const tdd_pred_ins_t& in = P.ins[i];
        uint64_t m = 0;
        switch (in.kind) {
        case TDD_P_CMP_CONST:
#pragma unroll
            for (int j = 0; j < R; j++) {
                int64_t a = cget<R>(cv, in.a, j);
                bool b = in.as_f64 ? cmp_f(in.cmp, cval_f64(T, in.a, a), in.cf)
                                   : cmp_i(in.cmp, a, in.ci);
                m |= (uint64_t)b << j;
            }
            st = (st << R) | m;
            break;
        case TDD_P_CMP_COL:
#pragma unroll
            for (int j = 0; j < R; j++) {
                int64_t a = cget<R>(cv, in.a, j), b2 = cget<R>(cv, in.b, j);
                bool b = in.as_f64 ? cmp_f(in.cmp, cval_f64(T, in.a, a), cval_f64(T, in.b, b2))
                                   : cmp_i(in.cmp, a, b2);
                m |= (uint64_t)b << j;
            }
            st = (st << R) | m;
            break;
        case TDD_P_AND: m = (st & M) & ((st >> R) & M); st = ((st >> (2 * R)) << R) | m; break;
        case TDD_P_OR:  m = (st & M) | ((st >> R) & M); st = ((st >> (2 * R)) << R) | m; break;
        case TDD_P_NOT: st ^= M; break;
        case TDD_P_CONST: st = (st << R) | (in.ci ? M : 0); break;
        default: /* TDD_P_BOOLCOL */
#pragma unroll
            for (int j = 0; j < R; j++) m |= (uint64_t)(cget<R>(cv, in.a, j) != 0) << j;
            st = (st << R) | m;
            break;
        }
    }
    return (uint32_t)(st & M);
}

/* Rows of the tile that exist (base + j < n) and are set in `bits`. */
template <int R>
__device__ __forceinline__ uint32_t tile_live(const uint64_t* bits, int64_t base, int64_t n) {
    uint32_t m = (1u << R) - 1;
    if (base + R > n) m = base >= n ? 0u : (1u << (n - base)) - 1;
    if (bits && m) m &= (uint32_t)(bits[base >> 6] >> (base & 63));
    return m;
}

/* Host: maximum stack depth of a postfix predicate. */
static inline int pred_depth(const tdd_pred_t& P) {
    int d = 0, mx = 0;
    for (int i = 0; i < P.n_ins; i++) {
        switch (P.ins[i].kind) {
        case TDD_P_AND: case TDD_P_OR: d -= 1; break;
        case TDD_P_NOT: break;
        default: d += 1; break;
        }
        if (d > mx) mx = d;
    }
    return mx;
}

/* Host: slot of column c in T, adding it when there is room (-1: full). */
static inline int tile_slot(TileCols& T, const tdd_col_t& c) {
    for (int i = 0; i < T.n; i++)
        if (T.col[i].data == c.data && T.col[i].type == c.type && T.col[i].attrs == c.attrs)
            return i;
    if (T.n >= TL_NC) return -1;
    T.col[T.n] = c;
    if (tile_vec_ok(c)) T.vec |= 1u << T.n;
    return T.n++;
}

/* ------------------------------------------------------------ star tiles -- */
/* The common shape of a fused probe: the join key in tile slot 0 and a
 * predicate that is a conjunction of at most two comparisons of a probe
 * column with a constant, evaluated as range checks on slots 1..npred.
 * Everything the row loop touches then sits at a compile-time slot;
 * aggregate inputs are reached through a uniform switch per cell (see
 * acc_cell), so no per-row selection between cached columns remains. */

struct StarPred {
    int32_t npred;
    int32_t f64[2];            /* compare as doubles */
    int32_t neg[2];            /* negate (NE) */
    int32_t lo_open[2], hi_open[2];
    int64_t lo[2], hi[2];      /* integer bounds, inclusive */
    double  flo[2], fhi[2];    /* f64 bounds */
};

template <int R>
__device__ __forceinline__ uint32_t range_mask(const StarPred& S, int i, bool col_f64,
                                               const int64_t (&v)[R]) {
    uint32_t m = 0;
    if (S.f64[i]) {
#pragma unroll
        for (int j = 0; j < R; j++) {
            double x = col_f64 ? __longlong_as_double(v[j]) : (double)v[j];
            bool in = (S.lo_open[i] ? x > S.flo[i] : x >= S.flo[i]) &&
                      (S.hi_open[i] ? x < S.fhi[i] : x <= S.fhi[i]);
            m |= (uint32_t)(in != (S.neg[i] != 0)) << j;
        }
    } else {
        const uint64_t span = (uint64_t)(S.hi[i] - S.lo[i]);
#pragma unroll
        for (int j = 0; j < R; j++) {
            bool in = (uint64_t)(v[j] - S.lo[i]) <= span;
            m |= (uint32_t)(in != (S.neg[i] != 0)) << j;
        }
    }
    return m;
}

template <int R, int NPRED>
__device__ __forceinline__ uint32_t star_pred(const StarPred& S, const TileCols& T,
                                              const int64_t (&cv)[TL_NC][R]) {
    uint32_t m = (1u << R) - 1;
    if (NPRED >= 1) m &= range_mask<R>(S, 0, T.col[1].type == TD_F64, cv[1]);
    if (NPRED >= 2) m &= range_mask<R>(S, 1, T.col[2].type == TD_F64, cv[2]);
    return m;
}

/* Host: tile with the key at slot 0 and a star predicate at slots 1..npred;
 * false when the predicate is not a conjunction of <= 2 column-constant
 * comparisons. */
static inline bool star_plan(const tdd_col_t& key, const tdd_pred_t* pred, TileCols& T,
                             StarPred& S) {
    memset(&T, 0, sizeof(T));
    memset(&S, 0, sizeof(S));
    T.col[0] = key;
    if (tile_vec_ok(key)) T.vec |= 1u;
    T.n = 1;
    if (!pred) return true;
    const tdd_pred_t& P = *pred;
    int n = P.n_ins;
    bool shape1 = n == 1 && P.ins[0].kind == TDD_P_CMP_CONST;
    bool shape2 = n == 3 && P.ins[0].kind == TDD_P_CMP_CONST &&
                  P.ins[1].kind == TDD_P_CMP_CONST && P.ins[2].kind == TDD_P_AND;
    if (!shape1 && !shape2) return false;
    S.npred = shape1 ? 1 : 2;
    for (int i = 0; i < S.npred; i++) {
        const tdd_pred_ins_t& in = P.ins[i];
        const tdd_col_t& c = P.cols[in.a];
        T.col[T.n] = c;
        if (tile_vec_ok(c)) T.vec |= 1u << T.n;
        T.n++;
        if (in.as_f64) {
            const double inf = __builtin_inf(), x = in.cf;
            S.f64[i] = 1;
            S.flo[i] = -inf;
            S.fhi[i] = inf;
            switch (in.cmp) {
            case OP_EQ: S.flo[i] = S.fhi[i] = x; break;
            case OP_NE: S.flo[i] = S.fhi[i] = x; S.neg[i] = 1; break;
            case OP_LT: S.fhi[i] = x; S.hi_open[i] = 1; break;
            case OP_LE: S.fhi[i] = x; break;
            case OP_GT: S.flo[i] = x; S.lo_open[i] = 1; break;
            default:    S.flo[i] = x; break;                          /* GE */
            }
        } else {
            const int64_t mn = INT64_MIN, mx = INT64_MAX, x = in.ci;
            S.lo[i] = mn;
            S.hi[i] = mx;
            switch (in.cmp) {
            case OP_EQ: S.lo[i] = S.hi[i] = x; break;
            case OP_NE: S.lo[i] = S.hi[i] = x; S.neg[i] = 1; break;
            case OP_LT: if (x == mn) S.neg[i] = 1; else S.hi[i] = x - 1; break;
            case OP_LE: S.hi[i] = x; break;
            case OP_GT: if (x == mx) S.neg[i] = 1; else S.lo[i] = x + 1; break;
            default:    S.lo[i] = x; break;                           /* GE */
            }
        }
    }
    return true;
}
