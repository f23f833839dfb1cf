/*
 * k_expr.hip -- element-wise expression evaluation.
 *
 * One row per thread; the registers of every thread live in LDS (a column
 * per thread), so the interpreter indexes them freely without private
 * memory.  All lanes of a wave run the same instruction stream, so the
 * opcode switch is uniform.  Arithmetic follows the reference's morsel
 * evaluator (exec.c:960-1100) operation by operation; HIP code here is
 * built without FP contraction, so a * b + c rounds twice as it does there.
 */
#include "dev_common.h"

#define EX_T 256

__device__ __forceinline__ double as_f(uint64_t u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ uint64_t f_as(double d) { return (uint64_t)__double_as_longlong(d); }

__global__ __launch_bounds__(EX_T) void k_expr(tdd_expr_t E, int64_t n, void* out) {
    __shared__ uint64_t rf[TDD_EXPR_MAX_REGS][EX_T];
    const int t = threadIdx.x;
    const int8_t ot = E.regs[E.out_reg].type;
    for (int64_t row = (int64_t)blockIdx.x * EX_T + t; row < n; row += (int64_t)gridDim.x * EX_T) {
        for (int r = 0; r < E.n_regs; r++) {
            const tdd_expr_reg_t& R = E.regs[r];
            if (R.kind == TDD_R_COL) {
                if (R.type == TD_F64)
                    rf[r][t] = f_as(R.col.type == TD_F64 ? ((const double*)R.col.data)[row]
                                                          : (double)ld_i64(R.col, row));
                else
                    rf[r][t] = (uint64_t)ld_i64(R.col, row);
            } else if (R.kind == TDD_R_CONST) {
                rf[r][t] = R.type == TD_F64 ? f_as(R.cf) : (uint64_t)R.ci;
            }
        }
        for (int i = 0; i < E.n_ins; i++) {
            const tdd_expr_ins_t I = E.ins[i];
            const int8_t dt = E.regs[I.dst].type;
            const int8_t t1 = E.regs[I.src1].type;
            const uint64_t a = rf[I.src1][t];
            uint64_t res = 0;
            if (I.src2 != 0xFF) {
                const uint64_t b = rf[I.src2][t];
                if (dt == TD_F64) {
                    const double x = as_f(a), y = as_f(b);
                    double z = 0.0;
                    switch (I.op) {
                    case OP_ADD: z = x + y; break;
                    case OP_SUB: z = x - y; break;
                    case OP_MUL: z = x * y; break;
                    case OP_DIV: z = y != 0 ? x / y : 0; break;
                    case OP_MOD: z = y != 0 ? fmod(x, y) : 0; break;
                    case OP_MIN2: z = x < y ? x : y; break;
                    case OP_MAX2: z = x > y ? x : y; break;
                    default: break;
                    }
                    res = f_as(z);
                } else if (dt == TD_I64) {
                    const int64_t x = (int64_t)a, y = (int64_t)b;
                    const bool ok = y != 0 && !(x == INT64_MIN && y == -1);
                    switch (I.op) {
                    case OP_ADD: res = a + b; break;
                    case OP_SUB: res = a - b; break;
                    case OP_MUL: res = a * b; break;
                    case OP_DIV: res = ok ? (uint64_t)(x / y) : 0; break;
                    case OP_MOD: res = ok ? (uint64_t)(x % y) : 0; break;
                    case OP_MIN2: res = (uint64_t)(x < y ? x : y); break;
                    case OP_MAX2: res = (uint64_t)(x > y ? x : y); break;
                    default: break;
                    }
                } else {                                        /* TD_BOOL */
                    bool z = false;
                    if (I.op == OP_AND || I.op == OP_OR) {
                        z = I.op == OP_AND ? (a && b) : (a || b);
                    } else if (t1 == TD_F64) {
                        const double x = as_f(a), y = as_f(b);
                        switch (I.op) {
                        case OP_EQ: z = x == y; break;
                        case OP_NE: z = x != y; break;
                        case OP_LT: z = x < y; break;
                        case OP_LE: z = x <= y; break;
                        case OP_GT: z = x > y; break;
                        case OP_GE: z = x >= y; break;
                        default: break;
                        }
                    } else if (t1 == TD_I64) {
                        const int64_t x = (int64_t)a, y = (int64_t)b;
                        switch (I.op) {
                        case OP_EQ: z = x == y; break;
                        case OP_NE: z = x != y; break;
                        case OP_LT: z = x < y; break;
                        case OP_LE: z = x <= y; break;
                        case OP_GT: z = x > y; break;
                        case OP_GE: z = x >= y; break;
                        default: break;
                        }
                    } else {                                    /* bool = bool */
                        switch (I.op) {
                        case OP_EQ: z = a == b; break;
                        case OP_NE: z = a != b; break;
                        default: break;
                        }
                    }
                    res = z;
                }
            } else {
                if (dt == TD_F64) {
                    if (t1 == TD_F64) {
                        const double x = as_f(a);
                        double z = 0.0;
                        switch (I.op) {
                        case OP_NEG: z = -x; break;
                        case OP_ABS: z = fabs(x); break;
                        case OP_SQRT: z = sqrt(x); break;
                        case OP_LOG: z = log(x); break;
                        case OP_EXP: z = exp(x); break;
                        case OP_CEIL: z = ceil(x); break;
                        case OP_FLOOR: z = floor(x); break;
                        default: z = x; break;
                        }
                        res = f_as(z);
                    } else {
                        res = f_as((double)(int64_t)a);           /* CAST i64 -> f64 */
                    }
                } else if (dt == TD_I64) {
                    if (t1 != TD_F64) {                         /* i64, or bool as 0 / 1 */
                        const int64_t x = (int64_t)a;
                        switch (I.op) {
                        case OP_NEG: res = 0ull - a; break;
                        case OP_ABS: res = x < 0 ? 0ull - a : a; break;
                        default: res = a; break;
                        }
                    } else {                                    /* CAST f64 -> i64, clamped */
                        /* NaN takes INT64_MIN, the value x86's truncating
                         * conversion gives the reference there */
                        const double x = as_f(a);
                        res = x >= 9223372036854775807.0 ? (uint64_t)INT64_MAX
                            : x > -9223372036854775808.0 ? (uint64_t)(int64_t)x
                            : (uint64_t)INT64_MIN;
                    }
                } else {
                    res = I.op == OP_NOT ? !a : a;
                }
            }
            rf[I.dst][t] = res;
        }
        const uint64_t o = rf[E.out_reg][t];
        if (ot == TD_BOOL) ((uint8_t*)out)[row] = (uint8_t)(o != 0);
        else ((uint64_t*)out)[row] = o;
    }
}

extern "C" int tdd_expr_eval(const tdd_expr_t* e, int64_t n, void* out) {
    if (n <= 0) return TD_OK;
    if (e->n_regs < 1 || e->n_regs > TDD_EXPR_MAX_REGS || e->n_ins < 0 ||
        e->n_ins > TDD_EXPR_MAX_INS || e->out_reg < 0 || e->out_reg >= e->n_regs)
        return TD_ERR_NYI;
    for (int i = 0; i < e->n_ins; i++) {
        const tdd_expr_ins_t& I = e->ins[i];
        if (I.dst >= e->n_regs || I.src1 >= e->n_regs || (I.src2 != 0xFF && I.src2 >= e->n_regs))
            return TD_ERR_NYI;
    }
    LAUNCH("expr", k_expr, stream_grid(ceil_div(n, EX_T), 4), EX_T, 0, *e, n, out);
    return TD_OK;
}
