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
