/*
 * k_expr.hip -- element-wise expression evaluation.
 *
 * An interpreter whose instruction stream is the same for every lane, so
 * its dispatch is uniform (scalar branches).  Each thread evaluates R rows
 * at a time and every instruction is dispatched once for all R: the
 * per-instruction scalar work (descriptor load, dispatch) is shared by R
 * rows.  Register values live in LDS, R row slots per register slot, each
 * thread touching its own slots only (no barriers); a register whose last
 * reader has run gives its slot to a later result; constants are scalars
 * in the descriptor.  Arithmetic follows the reference's morsel evaluator
 * (exec.c:960-1100) operation by operation; HIP code here is built without
 * FP contraction, so a * b + c rounds twice as it does there.
 */
#include <type_traits>

#include "dev_common.h"

#define EX_T    256
#define EX_MAXC 4          /* most columns of the pipelined kernel */
#define EX_R    4          /* rows per thread per step, pipelined kernel */

__device__ __forceinline__ double as_f(uint64_t u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ uint64_t f_as(double d) { return (uint64_t)__double_as_longlong(d); }

/* What an instruction computes, resolved on the host from its opcode and
 * its result and first operand types (the branches of exec.c:980-1100). */
enum ExOp : uint32_t {
    X_ZERO = 0,
    X_FADD, X_FSUB, X_FMUL, X_FDIV, X_FMOD, X_FMIN, X_FMAX,
    X_IADD, X_ISUB, X_IMUL, X_IDIV, X_IMOD, X_IMIN, X_IMAX,
    X_FEQ, X_FNE, X_FLT, X_FLE, X_FGT, X_FGE,
    X_IEQ, X_INE, X_ILT, X_ILE, X_IGT, X_IGE,
    X_BEQ, X_BNE, X_AND, X_OR,
    X_FNEG, X_FABS, X_FSQRT, X_FLOG, X_FEXP, X_FCEIL, X_FFLOOR, X_FCOPY,
    X_I2F, X_INEG, X_IABS, X_ICOPY, X_F2I, X_NOT, X_BCOPY,
};

/* An instruction in eight dwords, read with one scalar load (gfx950 scalar
 * loads have no byte form; byte fields would be fetched with vector loads,
 * whose waits would also wait for the column loads in flight):
 *   w[0] ExOp | src1 constant << 8 | src2 constant << 9 | binary << 10
 *   w[1] LDS element offset of src1's slot | of src2's slot << 16
 *   w[2] LDS element offset of the result's slot
 *   w[4..5] src1's constant bits, w[6..7] src2's
 * A slot holds `rows` row values EX_T elements apart; the thread index is
 * added in the kernel. */
struct __align__(32) ExIns {
    uint32_t w[8];
};

/* The expression's columns and instructions, and the facts the row loops
 * need, as dwords. */
struct ExPlan {
    int32_t   n_cols, n_ins, out_off, out_bool;
    int32_t   col_off[TDD_EXPR_MAX_REGS];    /* LDS element offset of each column's slot */
    int32_t   conv[TDD_EXPR_MAX_REGS];       /* integer column into an f64 register */
    tdd_col_t cols[TDD_EXPR_MAX_REGS];
    ExIns     ins[TDD_EXPR_MAX_INS];
};

/* The register value of a raw column load (integer widths widened, f64
 * bits as is): an integer column read into an f64 register converts. */
__device__ __forceinline__ uint64_t ex_value(int conv, uint64_t raw) {
    return conv ? f_as((double)(int64_t)raw) : raw;
}

/* CAST f64 -> i64, clamped (exec.c:1072-1077); NaN takes INT64_MIN, the
 * value x86's truncating conversion gives the reference there. */
__device__ __forceinline__ uint64_t ex_f2i(double x) {
    return x >= 9223372036854775807.0 ? (uint64_t)INT64_MAX
         : x > -9223372036854775808.0 ? (uint64_t)(int64_t)x
         : (uint64_t)INT64_MIN;
}

#define EX_LOOP(expr)                                   \
    _Pragma("unroll") for (int r = 0; r < R; r++) {     \
        const uint64_t a = A[r], b = B[r];              \
        (void)a; (void)b;                               \
        Z[r] = (expr);                                  \
    }                                                   \
    break

/* One instruction over R rows of this thread: operands from the slots (or
 * the constants), one dispatch, results into the destination slot. */
template <int R>
__device__ __forceinline__ void ex_exec(const ExIns& J, uint64_t* L) {
    const uint4 h = *(const uint4*)J.w;
    const uint4 k = *(const uint4*)(J.w + 4);
    const uint32_t op = h.x & 0xFF;
    uint64_t A[R], B[R], Z[R];
    if (h.x & 0x100) {
        const uint64_t c = (uint64_t)k.y << 32 | k.x;
#pragma unroll
        for (int r = 0; r < R; r++) A[r] = c;
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) A[r] = L[(h.y & 0xFFFF) + r * EX_T];
    }
    if (h.x & 0x200) {
        const uint64_t c = (uint64_t)k.w << 32 | k.z;
#pragma unroll
        for (int r = 0; r < R; r++) B[r] = c;
    } else if (h.x & 0x400) {
#pragma unroll
        for (int r = 0; r < R; r++) B[r] = L[(h.y >> 16) + r * EX_T];
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) B[r] = 0;
    }
    switch (op) {
    case X_FADD: EX_LOOP(f_as(as_f(a) + as_f(b)));
    case X_FSUB: EX_LOOP(f_as(as_f(a) - as_f(b)));
    case X_FMUL: EX_LOOP(f_as(as_f(a) * as_f(b)));
    case X_FDIV: EX_LOOP(f_as(as_f(b) != 0 ? as_f(a) / as_f(b) : 0.0));
    case X_FMOD: EX_LOOP(f_as(as_f(b) != 0 ? fmod(as_f(a), as_f(b)) : 0.0));
    case X_FMIN: EX_LOOP(f_as(as_f(a) < as_f(b) ? as_f(a) : as_f(b)));
    case X_FMAX: EX_LOOP(f_as(as_f(a) > as_f(b) ? as_f(a) : as_f(b)));
    case X_IADD: EX_LOOP(a + b);
    case X_ISUB: EX_LOOP(a - b);
    case X_IMUL: EX_LOOP(a * b);
    case X_IDIV:
        EX_LOOP((int64_t)b != 0 && !((int64_t)a == INT64_MIN && (int64_t)b == -1)
                    ? (uint64_t)((int64_t)a / (int64_t)b) : 0);
    case X_IMOD:
        EX_LOOP((int64_t)b != 0 && !((int64_t)a == INT64_MIN && (int64_t)b == -1)
                    ? (uint64_t)((int64_t)a % (int64_t)b) : 0);
    case X_IMIN: EX_LOOP((int64_t)a < (int64_t)b ? a : b);
    case X_IMAX: EX_LOOP((int64_t)a > (int64_t)b ? a : b);
    case X_FEQ: EX_LOOP((uint64_t)(as_f(a) == as_f(b)));
    case X_FNE: EX_LOOP((uint64_t)(as_f(a) != as_f(b)));
    case X_FLT: EX_LOOP((uint64_t)(as_f(a) < as_f(b)));
    case X_FLE: EX_LOOP((uint64_t)(as_f(a) <= as_f(b)));
    case X_FGT: EX_LOOP((uint64_t)(as_f(a) > as_f(b)));
    case X_FGE: EX_LOOP((uint64_t)(as_f(a) >= as_f(b)));
    case X_IEQ: EX_LOOP((uint64_t)((int64_t)a == (int64_t)b));
    case X_INE: EX_LOOP((uint64_t)((int64_t)a != (int64_t)b));
    case X_ILT: EX_LOOP((uint64_t)((int64_t)a < (int64_t)b));
    case X_ILE: EX_LOOP((uint64_t)((int64_t)a <= (int64_t)b));
    case X_IGT: EX_LOOP((uint64_t)((int64_t)a > (int64_t)b));
    case X_IGE: EX_LOOP((uint64_t)((int64_t)a >= (int64_t)b));
    case X_BEQ: EX_LOOP((uint64_t)(a == b));
    case X_BNE: EX_LOOP((uint64_t)(a != b));
    case X_AND: EX_LOOP((uint64_t)(a && b));
    case X_OR: EX_LOOP((uint64_t)(a || b));
    case X_FNEG: EX_LOOP(f_as(-as_f(a)));
    case X_FABS: EX_LOOP(f_as(fabs(as_f(a))));
    case X_FSQRT: EX_LOOP(f_as(sqrt(as_f(a))));
    case X_FLOG: EX_LOOP(f_as(log(as_f(a))));
    case X_FEXP: EX_LOOP(f_as(exp(as_f(a))));
    case X_FCEIL: EX_LOOP(f_as(ceil(as_f(a))));
    case X_FFLOOR: EX_LOOP(f_as(floor(as_f(a))));
    case X_I2F: EX_LOOP(f_as((double)(int64_t)a));
    case X_INEG: EX_LOOP(0ull - a);
    case X_IABS: EX_LOOP((int64_t)a < 0 ? 0ull - a : a);
    case X_F2I: EX_LOOP(ex_f2i(as_f(a)));
    case X_NOT: EX_LOOP((uint64_t)(a == 0));
    case X_FCOPY: case X_ICOPY: case X_BCOPY: EX_LOOP(a);
    default: EX_LOOP((uint64_t)0);
    }
    uint64_t* D = L + J.w[2];
#pragma unroll
    for (int r = 0; r < R; r++) D[r * EX_T] = Z[r];
}

/* Any expression, one row per thread at a time: each row's columns (any
 * width) are read into their slots, then the instructions run. */
__global__ __launch_bounds__(EX_T) void k_expr_any(ExPlan X, int64_t n, void* out) {
    extern __shared__ uint64_t lds[];
    uint64_t* L = lds + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * EX_T;
    for (int64_t row = (int64_t)blockIdx.x * EX_T + threadIdx.x; row < n; row += stride) {
        for (int c = 0; c < X.n_cols; c++)
            L[X.col_off[c]] = ex_value(X.conv[c], (uint64_t)ld_i64(X.cols[c], row));
        for (int i = 0; i < X.n_ins; i++) ex_exec<1>(X.ins[i], L);
        const uint64_t o = L[X.out_off];
        if (X.out_bool) ((uint8_t*)out)[row] = (uint8_t)(o != 0);
        else ((uint64_t*)out)[row] = o;
    }
}

/* The R results of the step starting at `row`, stored (rows past the end
 * clamped to the last, whose value they hold). */
template <int R>
__device__ __forceinline__ void ex_store(const ExPlan& X, const uint64_t* v, int64_t row,
                                         int64_t stride, int64_t n, void* out) {
    if (X.out_bool) {
#pragma unroll
        for (int r = 0; r < R; r++)
            ((uint8_t*)out)[min(row + r * stride, n - 1)] = (uint8_t)(v[r] != 0);
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) ((uint64_t*)out)[min(row + r * stride, n - 1)] = v[r];
    }
}

/* NC columns, all of 8-byte elements; a thread's step covers rows
 * row, row + S, ..., row + (R-1) S (S = grid threads).  A step (1) takes
 * the previous step's results out of LDS, (2) moves the loaded column
 * values into their slots, (3) stores the previous results, (4) issues the
 * loads of the next step and (5) runs the instructions.  The wait in (2) is
 * for memory operations issued before the last (5), so the loads of one
 * step overlap the evaluation of the one before.  Past the last row the
 * loads and stores are clamped to it (the same inputs give the same value,
 * so the duplicate store writes what is already there). */
template <int NC, int R>
__global__ __launch_bounds__(EX_T) void k_expr_tile(ExPlan X, int64_t n, void* out) {
    extern __shared__ uint64_t lds[];
    uint64_t* L = lds + threadIdx.x;
    const uint64_t* src[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) src[c] = (const uint64_t*)X.cols[c].data;
    const int64_t stride = (int64_t)gridDim.x * EX_T;
    const int64_t row0 = (int64_t)blockIdx.x * EX_T + threadIdx.x;
    uint64_t ld[R][NC], res[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t rr = min(row0 + r * stride, n - 1);
#pragma unroll
        for (int c = 0; c < NC; c++) ld[r][c] = src[c][rr];
    }
    int64_t prev = -1;
    for (int64_t row = row0; row < n; row += R * stride) {
        if (prev >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) res[r] = L[X.out_off + r * EX_T];
        }
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int r = 0; r < R; r++) L[X.col_off[c] + r * EX_T] = ex_value(X.conv[c], ld[r][c]);
        if (prev >= 0) ex_store<R>(X, res, prev, stride, n, out);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t rr = min(row + (R + r) * stride, n - 1);
#pragma unroll
            for (int c = 0; c < NC; c++) ld[r][c] = src[c][rr];
        }
        for (int i = 0; i < X.n_ins; i++) ex_exec<R>(X.ins[i], L);
        prev = row;
    }
    if (prev >= 0) {
#pragma unroll
        for (int r = 0; r < R; r++) res[r] = L[X.out_off + r * EX_T];
        ex_store<R>(X, res, prev, stride, n, out);
    }
}

template <int NC>
static int launch_tile(const ExPlan& X, int64_t n, void* out, size_t shmem) {
    const int64_t units = ceil_div(n, (int64_t)EX_T * EX_R);
    LAUNCH("expr", (k_expr_tile<NC, EX_R>),
           resident_grid(k_expr_tile<NC, EX_R>, EX_T, shmem, units), EX_T, shmem, X, n, out);
    return TD_OK;
}

/* ExOp of an instruction: its opcode under its result type and first
 * operand type, as the reference's evaluator branches (exec.c:980-1100). */
static uint32_t ex_code(uint8_t op, int8_t dt, int8_t t1, bool binary) {
    if (binary) {
        if (dt == TD_F64) {
            switch (op) {
            case OP_ADD: return X_FADD; case OP_SUB: return X_FSUB; case OP_MUL: return X_FMUL;
            case OP_DIV: return X_FDIV; case OP_MOD: return X_FMOD; case OP_MIN2: return X_FMIN;
            case OP_MAX2: return X_FMAX; default: return X_ZERO;
            }
        }
        if (dt == TD_I64) {
            switch (op) {
            case OP_ADD: return X_IADD; case OP_SUB: return X_ISUB; case OP_MUL: return X_IMUL;
            case OP_DIV: return X_IDIV; case OP_MOD: return X_IMOD; case OP_MIN2: return X_IMIN;
            case OP_MAX2: return X_IMAX; default: return X_ZERO;
            }
        }
        if (op == OP_AND) return X_AND;
        if (op == OP_OR) return X_OR;
        if (op < OP_EQ || op > OP_GE) return X_ZERO;
        if (t1 == TD_F64) return X_FEQ + (op - OP_EQ);
        if (t1 == TD_I64) return X_IEQ + (op - OP_EQ);
        return op == OP_EQ ? X_BEQ : op == OP_NE ? X_BNE : X_ZERO;
    }
    if (dt == TD_F64) {
        if (t1 != TD_F64) return X_I2F;
        switch (op) {
        case OP_NEG: return X_FNEG; case OP_ABS: return X_FABS; case OP_SQRT: return X_FSQRT;
        case OP_LOG: return X_FLOG; case OP_EXP: return X_FEXP; case OP_CEIL: return X_FCEIL;
        case OP_FLOOR: return X_FFLOOR; default: return X_FCOPY;
        }
    }
    if (dt == TD_I64) {
        if (t1 == TD_F64) return X_F2I;
        return op == OP_NEG ? X_INEG : op == OP_ABS ? X_IABS : X_ICOPY;
    }
    return op == OP_NOT ? X_NOT : X_BCOPY;
}

static uint64_t const_bits(const tdd_expr_reg_t& R) {
    if (R.type != TD_F64) return (uint64_t)R.ci;
    uint64_t u;
    memcpy(&u, &R.cf, 8);
    return u;
}

extern "C" int tdd_expr_eval(const tdd_expr_t* e, int64_t n, void* out) {
    if (n <= 0) return TD_OK;
    if (e->n_regs < 1 || e->n_regs > TDD_EXPR_MAX_REGS || e->n_ins < 1 ||
        e->n_ins > TDD_EXPR_MAX_INS || e->out_reg < 0 || e->out_reg >= e->n_regs)
        return TD_ERR_NYI;
    for (int i = 0; i < e->n_ins; i++) {
        const tdd_expr_ins_t& I = e->ins[i];
        if (I.dst >= e->n_regs || I.src1 >= e->n_regs || (I.src2 != 0xFF && I.src2 >= e->n_regs) ||
            e->regs[I.dst].kind != TDD_R_TMP)
            return TD_ERR_NYI;
    }
    ExPlan X;
    memset(&X, 0, sizeof(X));
    X.n_ins = e->n_ins;
    X.out_bool = e->regs[e->out_reg].type == TD_BOOL;
    bool tiled = true;
    for (int r = 0; r < e->n_regs; r++)
        if (e->regs[r].kind == TDD_R_COL) {
            const tdd_col_t& col = e->regs[r].col;
            tiled = tiled && (col.type == TD_I64 || col.type == TD_F64 || col.type == TD_TIMESTAMP ||
                              (col.type == TD_SYM && (col.attrs & TD_SYM_W_MASK) == TD_SYM_W64));
            X.n_cols++;
        }
    tiled = tiled && X.n_cols >= 1 && X.n_cols <= EX_MAXC;
    const int rows = tiled ? EX_R : 1;

    /* Slots: a column takes one at the start, a result when it is written;
     * a register's slot is freed after its last reader (the result of the
     * last instruction is read by the store), so a result may take a slot
     * its own operands free. */
    int last[TDD_EXPR_MAX_REGS];
    for (int r = 0; r < e->n_regs; r++) last[r] = -1;
    for (int i = 0; i < e->n_ins; i++) {
        last[e->ins[i].src1] = i;
        if (e->ins[i].src2 != 0xFF) last[e->ins[i].src2] = i;
    }
    last[e->out_reg] = e->n_ins;
    int slot[TDD_EXPR_MAX_REGS], free_slots[TDD_EXPR_MAX_REGS], n_slots = 0, n_free = 0;
    for (int r = 0; r < e->n_regs; r++) slot[r] = -1;
    auto take = [&]() { return n_free ? free_slots[--n_free] : n_slots++; };
    X.n_cols = 0;
    for (int r = 0; r < e->n_regs; r++)
        if (e->regs[r].kind == TDD_R_COL) {
            slot[r] = take();
            X.cols[X.n_cols] = e->regs[r].col;
            X.conv[X.n_cols] = e->regs[r].type == TD_F64 && e->regs[r].col.type != TD_F64;
            X.col_off[X.n_cols] = slot[r] * rows * EX_T;
            X.n_cols++;
        }
    for (int i = 0; i < e->n_ins; i++) {
        const tdd_expr_ins_t& I = e->ins[i];
        const bool binary = I.src2 != 0xFF;
        const tdd_expr_reg_t& R1 = e->regs[I.src1];
        const bool c1 = R1.kind == TDD_R_CONST;
        const bool c2 = binary && e->regs[I.src2].kind == TDD_R_CONST;
        ExIns& J = X.ins[i];
        J.w[0] = ex_code(I.op, e->regs[I.dst].type, R1.type, binary) | (uint32_t)c1 << 8 |
                 (uint32_t)c2 << 9 | (uint32_t)binary << 10;
        uint32_t o1 = 0, o2 = 0;
        if (c1) {
            const uint64_t u = const_bits(R1);
            J.w[4] = (uint32_t)u;
            J.w[5] = (uint32_t)(u >> 32);
        } else {
            if (slot[I.src1] < 0) return TD_ERR_NYI;          /* read before it is written */
            o1 = (uint32_t)(slot[I.src1] * rows * EX_T);
        }
        if (c2) {
            const uint64_t u = const_bits(e->regs[I.src2]);
            J.w[6] = (uint32_t)u;
            J.w[7] = (uint32_t)(u >> 32);
        } else if (binary) {
            if (slot[I.src2] < 0) return TD_ERR_NYI;
            o2 = (uint32_t)(slot[I.src2] * rows * EX_T);
        }
        /* operands read for the last time give their slots back first */
        if (!c1 && last[I.src1] == i) free_slots[n_free++] = slot[I.src1];
        if (binary && !c2 && last[I.src2] == i && I.src2 != I.src1)
            free_slots[n_free++] = slot[I.src2];
        if (slot[I.dst] >= 0) return TD_ERR_NYI;              /* written twice */
        slot[I.dst] = take();
        J.w[1] = o1 | o2 << 16;
        J.w[2] = (uint32_t)(slot[I.dst] * rows * EX_T);
        if (last[I.dst] < 0) free_slots[n_free++] = slot[I.dst];   /* never read */
    }
    if (slot[e->out_reg] < 0) return TD_ERR_NYI;
    X.out_off = slot[e->out_reg] * rows * EX_T;
    const size_t shmem = (size_t)n_slots * rows * EX_T * sizeof(uint64_t);
    if (tiled) {
        switch (X.n_cols) {
        case 1: return launch_tile<1>(X, n, out, shmem);
        case 2: return launch_tile<2>(X, n, out, shmem);
        case 3: return launch_tile<3>(X, n, out, shmem);
        default: return launch_tile<4>(X, n, out, shmem);
        }
    }
    LAUNCH("expr", k_expr_any, resident_grid(k_expr_any, EX_T, shmem, ceil_div(n, EX_T)), EX_T,
           shmem, X, n, out);
    return TD_OK;
}

/* ------------------------------------------------- affine aggregate bias -- */

__global__ void k_affine_fix(void* vals, int is_f64, int avg, const int64_t* counts, int64_t n,
                             int64_t bias_i, double bias_f) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    if (is_f64) {
        double* v = (double*)vals;
        v[g] = avg ? v[g] + bias_f : v[g] + bias_f * (double)counts[g];
    } else {
        uint64_t* v = (uint64_t*)vals;
        v[g] += (uint64_t)bias_i * (uint64_t)counts[g];
    }
}

extern "C" int tdd_affine_fix(void* vals, int8_t type, int avg, const int64_t* counts,
                              int64_t n, int64_t bias_i, double bias_f) {
    if (n <= 0) return TD_OK;
    if (!avg && !counts) return TD_ERR_NYI;
    LAUNCH("affine_fix", k_affine_fix, ceil_div(n, 256), 256, 0, vals, type == TD_F64, avg,
           counts, n, bias_i, bias_f);
    return TD_OK;
}

/* ----------------------------------------------------- partition column -- */

/* out[r] = kv[p] for the partition p whose rows [offs[p], offs[p+1])
 * hold r: the per-row key of a MAPCOMMON column (materialize_mapcommon,
 * exec.c:201-241), esz 1, 2, 4 or 8. */
#define MC_LDS 1024                 /* partitions whose offsets and keys go to LDS */

/* Key of element r, advancing the partition index p (offs ascending,
 * offs[0] = 0). */
__device__ __forceinline__ int mc_part(const int64_t* offs, int n_parts, int p, int64_t r) {
    while (p + 1 < n_parts && offs[p + 1] <= r) p++;
    return p;
}

/* Each thread fills 16-byte chunks: one binary search for the chunk's first
 * element, then a linear advance; 4- and 8-byte keys leave as one vector
 * store per chunk.  Offsets and keys are read from LDS when they fit. */
template <int ESZ, bool LDS>
__global__ __launch_bounds__(256) void k_mapcommon(void* out, const void* kv, const int64_t* goffs,
                                                   int n_parts, int64_t n, bool vec) {
    typedef typename std::conditional<ESZ == 8, uint64_t,
            typename std::conditional<ESZ == 4, uint32_t,
            typename std::conditional<ESZ == 2, uint16_t, uint8_t>::type>::type>::type E;
    constexpr int PER = 16 / ESZ;
    __shared__ int64_t so[LDS ? MC_LDS : 1];
    __shared__ E sk[LDS ? MC_LDS : 1];
    const int64_t* offs = goffs;
    const E* keys = (const E*)kv;
    if (LDS) {
        for (int i = threadIdx.x; i < n_parts; i += blockDim.x) {
            so[i] = goffs[i];
            sk[i] = ((const E*)kv)[i];
        }
        __syncthreads();
        offs = so;
        keys = sk;
    }
    E* o = (E*)out;
    const int64_t n_chunks = (n + PER - 1) / PER;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks; c += stride) {
        const int64_t r0 = c * PER;
        int lo = 0, hi = n_parts - 1;                 /* last p with offs[p] <= r0 */
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= r0) lo = mid;
            else hi = mid - 1;
        }
        if (vec && r0 + PER <= n && (ESZ == 4 || ESZ == 8)) {
            E v[PER];
#pragma unroll
            for (int e = 0; e < PER; e++) {
                lo = mc_part(offs, n_parts, lo, r0 + e);
                v[e] = keys[lo];
            }
            if (ESZ == 8) {
                ((ulonglong2*)o)[c] = make_ulonglong2((unsigned long long)v[0],
                                                      (unsigned long long)v[PER > 1 ? 1 : 0]);
            } else {
                ((uint4*)o)[c] = make_uint4((uint32_t)v[0], (uint32_t)v[PER > 1 ? 1 : 0],
                                            (uint32_t)v[PER > 2 ? 2 : 0], (uint32_t)v[PER > 3 ? 3 : 0]);
            }
        } else {
            for (int e = 0; e < PER && r0 + e < n; e++) {
                lo = mc_part(offs, n_parts, lo, r0 + e);
                o[r0 + e] = keys[lo];
            }
        }
    }
}

extern "C" int tdd_mapcommon_fill(void* out, int esz, const void* kv, const int64_t* offs,
                                  int n_parts, int64_t n) {
    if (n <= 0) return TD_OK;
    if (n_parts < 1 || (esz != 1 && esz != 2 && esz != 4 && esz != 8)) return TD_ERR_NYI;
    const bool vec = ((uintptr_t)out & 15) == 0;
    const int grid = stream_grid(ceil_div(ceil_div(n, 16 / esz), 256), 8);
    const bool lds = n_parts <= MC_LDS;
#define MC(ESZ_)                                                                            \
    do {                                                                                    \
        if (lds) LAUNCH("mapcommon", (k_mapcommon<ESZ_, true>), grid, 256, 0, out, kv, offs, \
                        n_parts, n, vec);                                                   \
        else LAUNCH("mapcommon", (k_mapcommon<ESZ_, false>), grid, 256, 0, out, kv, offs,   \
                    n_parts, n, vec);                                                       \
    } while (0)
    switch (esz) {
    case 1: MC(1); break;
    case 2: MC(2); break;
    case 4: MC(4); break;
    default: MC(8); break;
    }
#undef MC
    return TD_OK;
}
