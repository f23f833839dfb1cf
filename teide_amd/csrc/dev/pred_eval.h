/*
 * pred_eval.h -- per-row evaluation of a compiled filter predicate
 * (tdd_pred_t), shared by the selection kernels and the fused join kernels.
 */
#pragma once

#include "dev_common.h"

__device__ __forceinline__ bool cmp_i(int op, int64_t a, int64_t b) {
    switch (op) {
    case OP_EQ: return a == b;
    case OP_NE: return a != b;
    case OP_LT: return a < b;
    case OP_LE: return a <= b;
    case OP_GT: return a > b;
    default:    return a >= b;
    }
}

__device__ __forceinline__ bool cmp_f(int op, double a, double b) {
    switch (op) {
    case OP_EQ: return a == b;
    case OP_NE: return a != b;
    case OP_LT: return a < b;
    case OP_LE: return a <= b;
    case OP_GT: return a > b;
    default:    return a >= b;
    }
}

/* Postfix evaluation with the boolean stack kept in the bits of one
 * register (no private-memory array, so nothing spills). */
__device__ __forceinline__ bool pred_eval(const tdd_pred_t& P, int64_t r) {
    uint32_t st = 0;
    for (int i = 0; i < P.n_ins; i++) {
        const tdd_pred_ins_t& in = P.ins[i];
        uint32_t v;
        switch (in.kind) {
        case TDD_P_CMP_CONST:
            v = in.as_f64 ? cmp_f(in.cmp, ld_f64(P.cols[in.a], r), in.cf)
                          : cmp_i(in.cmp, ld_i64(P.cols[in.a], r), in.ci);
            st = (st << 1) | v;
            break;
        case TDD_P_CMP_COL:
            v = in.as_f64 ? cmp_f(in.cmp, ld_f64(P.cols[in.a], r), ld_f64(P.cols[in.b], r))
                          : cmp_i(in.cmp, ld_i64(P.cols[in.a], r), ld_i64(P.cols[in.b], r));
            st = (st << 1) | v;
            break;
        case TDD_P_AND: v = (st & 1) & ((st >> 1) & 1); st = (st >> 2 << 1) | v; break;
        case TDD_P_OR:  v = (st & 1) | ((st >> 1) & 1); st = (st >> 2 << 1) | v; break;
        case TDD_P_NOT: st ^= 1; break;
        case TDD_P_CONST: st = (st << 1) | (in.ci != 0); break;
        default: /* TDD_P_BOOLCOL */
            st = (st << 1) | (ld_i64(P.cols[in.a], r) != 0);
            break;
        }
    }
    return st & 1;
}

/* Four rows r0 + 64k (k = 0..3) of one column: the type is dispatched once
 * and the four loads are issued together.  Rows at or past nrows read
 * nothing and come back 0. */
template <class T>
__device__ __forceinline__ void ld4_as(const void* d, int64_t r0, int64_t nrows, int64_t* v) {
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = r0 + 64 * k < nrows ? (int64_t)((const T*)d)[r0 + 64 * k] : 0;
}

__device__ __forceinline__ void ld4_i64(const tdd_col_t& c, int64_t r0, int64_t nrows, int64_t* v) {
    switch (c.type) {
    case TD_I64: case TD_TIMESTAMP: case TD_F64: ld4_as<int64_t>(c.data, r0, nrows, v); break;
    case TD_SYM:
        switch (c.attrs & TD_SYM_W_MASK) {
        case TD_SYM_W8:  ld4_as<uint8_t>(c.data, r0, nrows, v); break;
        case TD_SYM_W16: ld4_as<uint16_t>(c.data, r0, nrows, v); break;
        case TD_SYM_W32: ld4_as<uint32_t>(c.data, r0, nrows, v); break;
        default:         ld4_as<int64_t>(c.data, r0, nrows, v); break;
        }
        break;
    case TD_I32: case TD_DATE: case TD_TIME: ld4_as<int32_t>(c.data, r0, nrows, v); break;
    case TD_I16: ld4_as<int16_t>(c.data, r0, nrows, v); break;
    default: ld4_as<uint8_t>(c.data, r0, nrows, v); break;
    }
}

__device__ __forceinline__ void ld4_f64(const tdd_col_t& c, int64_t r0, int64_t nrows, double* v) {
    int64_t b[4];
    ld4_i64(c, r0, nrows, b);
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = c.type == TD_F64 ? __longlong_as_double(b[k]) : (double)b[k];
}

/* pred_eval over the four rows r0 + 64k at once, one instruction at a time
 * for all four, so each instruction's loads are in flight together; bit k
 * of the result is row r0 + 64k's value (rows past nrows: unspecified). */
template <class Pred>
__device__ __forceinline__ uint32_t pred_eval4(Pred& P, int64_t r0, int64_t nrows) {
    uint32_t st[4] = {0, 0, 0, 0};
    for (int i = 0; i < P.n_ins; i++) {
        const tdd_pred_ins_t in = P.ins[i];
        uint32_t v[4];
        switch (in.kind) {
        case TDD_P_CMP_CONST:
        case TDD_P_CMP_COL:
            if (in.as_f64) {
                double a[4], b[4];
                const tdd_col_t ca = P.cols[in.a];
                ld4_f64(ca, r0, nrows, a);
                if (in.kind == TDD_P_CMP_COL) {
                    const tdd_col_t cb = P.cols[in.b];
                    ld4_f64(cb, r0, nrows, b);
                }
                else for (int k = 0; k < 4; k++) b[k] = in.cf;
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = cmp_f(in.cmp, a[k], b[k]);
            } else {
                int64_t a[4], b[4];
                const tdd_col_t ca = P.cols[in.a];
                ld4_i64(ca, r0, nrows, a);
                if (in.kind == TDD_P_CMP_COL) {
                    const tdd_col_t cb = P.cols[in.b];
                    ld4_i64(cb, r0, nrows, b);
                }
                else for (int k = 0; k < 4; k++) b[k] = in.ci;
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = cmp_i(in.cmp, a[k], b[k]);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) st[k] = (st[k] << 1) | v[k];
            break;
        case TDD_P_AND:
#pragma unroll
            for (int k = 0; k < 4; k++) st[k] = (st[k] >> 2 << 1) | (st[k] & (st[k] >> 1) & 1);
            break;
        case TDD_P_OR:
#pragma unroll
            for (int k = 0; k < 4; k++) st[k] = (st[k] >> 2 << 1) | ((st[k] | (st[k] >> 1)) & 1);
            break;
        case TDD_P_NOT:
#pragma unroll
            for (int k = 0; k < 4; k++) st[k] ^= 1;
            break;
        case TDD_P_CONST:
#pragma unroll
            for (int k = 0; k < 4; k++) st[k] = (st[k] << 1) | (in.ci != 0);
            break;
        default: { /* TDD_P_BOOLCOL */
            int64_t a[4];
            const tdd_col_t ca = P.cols[in.a];
            ld4_i64(ca, r0, nrows, a);
#pragma unroll
            for (int k = 0; k < 4; k++) st[k] = (st[k] << 1) | (a[k] != 0);
        }
        }
    }
    return (st[0] & 1) | (st[1] & 1) << 1 | (st[2] & 1) << 2 | (st[3] & 1) << 3;
}
