/*
 * group_plan.h -- accumulator layout shared by the group-by kernels and the
 * fused join -> group-by kernels.
 *
 * Every group owns `ncells` 8-byte cells: cell 0 is the row count, the
 * others are one or two cells per aggregate:
 *   SUM / AVG          add (f64 or wrapping i64, by the input's type)
 *   MIN / MAX          atomic min / max (f64 through an order-preserving
 *                      int64 image)
 *   FIRST / LAST       min / max row index; the value is gathered at emit
 *   VAR / STDDEV (pop) add + add of squares (f64)
 *   COUNT              uses cell 0
 * These are the reference's accumulators (da_accum_t, exec.c:4608-4620;
 * ght_layout_t, exec.c:3898-3952) arranged for atomics.
 */
#pragma once

#include "dev_common.h"

enum {
    C_COUNT = 0, C_ADD_F64, C_ADD_I64, C_MIN_I64, C_MAX_I64, C_MIN_F64, C_MAX_F64,
    C_MIN_ROW, C_MAX_ROW, C_SQ_F64
};

#define GP_MAX_CELLS (1 + 2 * TDD_MAX_AGGS)

struct GroupPlan {
    int32_t   n_keys, n_aggs, ncells, pad;
    tdd_col_t keys[TDD_MAX_KEYS];
    int64_t   kmin[TDD_MAX_KEYS];      /* direct array: per-key minimum */
    int64_t   kstride[TDD_MAX_KEYS];   /* direct array: slot stride */
    int64_t   krange[TDD_MAX_KEYS];
    tdd_col_t aggs[TDD_MAX_AGGS];
    uint16_t  ops[TDD_MAX_AGGS];
    int8_t    out_types[TDD_MAX_AGGS];
    uint8_t   cell_of[TDD_MAX_AGGS];   /* first cell of each aggregate */
    /* per-cell facts read with a run-time cell index: dwords, so the reads
     * are scalar loads (gfx950 scalar loads have no byte form; a byte array
     * is read with vector loads, whose waits also wait for the row loads
     * in flight) */
    int32_t   kind[GP_MAX_CELLS];
    int32_t   col[GP_MAX_CELLS];       /* aggregate column feeding the cell */
    uint8_t   key_side[TDD_MAX_KEYS];  /* fused join: 0 probe row, 1 build row */
    uint8_t   agg_side[TDD_MAX_AGGS];
};

/* Direct-array slot of a row (pair of rows for a fused join). */
__device__ __forceinline__ int64_t da_slot2(const GroupPlan& P, int64_t lrow, int64_t rrow) {
    int64_t s = 0;
    for (int k = 0; k < P.n_keys; k++)
        s += (ld_i64(P.keys[k], P.key_side[k] ? rrow : lrow) - P.kmin[k]) * P.kstride[k];
    return s;
}

__host__ __device__ __forceinline__ uint64_t cell_identity(int kind) {
    switch (kind) {
    case C_MIN_I64: case C_MIN_F64: case C_MIN_ROW: return 0x7fffffffffffffffull;
    case C_MAX_I64: case C_MAX_F64: case C_MAX_ROW: return 0x8000000000000000ull;
    default: return 0;
    }
}

/* MIN / MAX into a cell, skipping the atomic when a plain read already
 * shows the value cannot improve it.  A cell only ever moves toward its
 * final value, so a stale read is never past the current one and a skip is
 * never wrong; most rows of a group stop issuing atomics after the first
 * few.  (The read may see another workgroup's update late, which only
 * costs an atomic that changes nothing.) */
__device__ __forceinline__ void cell_min(uint64_t* c, int64_t v) {
    if (v < *(volatile int64_t*)c) atomicMin((long long*)c, (long long)v);
}
__device__ __forceinline__ void cell_max(uint64_t* c, int64_t v) {
    if (v > *(volatile int64_t*)c) atomicMax((long long*)c, (long long)v);
}
template <bool SKIP>
__device__ __forceinline__ void min_cell(uint64_t* c, int64_t v) {
    if (SKIP) cell_min(c, v);
    else atomicMin((long long*)c, (long long)v);
}
template <bool SKIP>
__device__ __forceinline__ void max_cell(uint64_t* c, int64_t v) {
    if (SKIP) cell_max(c, v);
    else atomicMax((long long*)c, (long long)v);
}

/* Fold one row into the cells at `c` (LDS or global; the address space is
 * resolved after inlining).  A fused join passes the probe and build rows;
 * each aggregate reads the row of its side.  SKIP: MIN / MAX through
 * cell_min / cell_max -- a win where the cell read is cheap (LDS, a table
 * that stays in cache); off where the read would be one more dependent
 * trip to memory per row (a large hash table). */
template <bool SKIP = true>
__device__ __forceinline__ void cells_accum2(uint64_t* c, const GroupPlan& P, int64_t lrow,
                                             int64_t rrow) {
    atomicAdd((unsigned long long*)&c[0], 1ull);
    for (int i = 1; i < P.ncells; i++) {
        const tdd_col_t& a = P.aggs[P.col[i]];
        int64_t row = P.agg_side[P.col[i]] ? rrow : lrow;
        switch (P.kind[i]) {
        case C_ADD_F64: atomicAdd((double*)&c[i], ld_f64(a, row)); break;
        case C_ADD_I64: atomicAdd((unsigned long long*)&c[i], (unsigned long long)ld_i64(a, row)); break;
        case C_MIN_I64: min_cell<SKIP>(&c[i], ld_i64(a, row)); break;
        case C_MAX_I64: max_cell<SKIP>(&c[i], ld_i64(a, row)); break;
        case C_MIN_F64: min_cell<SKIP>(&c[i], f64_ord(ld_f64(a, row))); break;
        case C_MAX_F64: max_cell<SKIP>(&c[i], f64_ord(ld_f64(a, row))); break;
        case C_MIN_ROW: min_cell<SKIP>(&c[i], row); break;
        case C_MAX_ROW: max_cell<SKIP>(&c[i], row); break;
        default: { double v = ld_f64(a, row); atomicAdd((double*)&c[i], v * v); } break;
        }
    }
}

template <bool SKIP = true>
__device__ __forceinline__ void cells_accum(uint64_t* c, const GroupPlan& P, int64_t row) {
    cells_accum2<SKIP>(c, P, row, row);
}

/* Wave-wide reductions (every lane gets the result; fixed butterfly order). */
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = (int64_t)((uint64_t)v + (uint64_t)__shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { int64_t u = __shfl_xor(v, o); v = u < v ? u : v; }
    return v;
}
__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { int64_t u = __shfl_xor(v, o); v = u > v ? u : v; }
    return v;
}

/* A wave whose selected rows (`act`, mask m) all fall in one group: the
 * rows folded across the wave first, then one update per cell by the
 * `leader` lane.  Every lane of the wave calls this together.  Runs of one
 * key -- sorted or partitioned data, a partition key repeated over its
 * rows -- otherwise send all 64 lanes' atomics to the same cells. */
__device__ __forceinline__ void cells_accum_wave(uint64_t* c, const GroupPlan& P, int64_t row,
                                                 bool act, uint64_t m, bool leader) {
    if (leader) atomicAdd((unsigned long long*)&c[0], (unsigned long long)__popcll(m));
    for (int i = 1; i < P.ncells; i++) {
        const tdd_col_t& a = P.aggs[P.col[i]];
        switch (P.kind[i]) {
        case C_ADD_F64: {
            double v = wave_sum_f64(act ? ld_f64(a, row) : 0.0);
            if (leader) atomicAdd((double*)&c[i], v);
        } break;
        case C_SQ_F64: {
            double x = act ? ld_f64(a, row) : 0.0;
            double v = wave_sum_f64(x * x);
            if (leader) atomicAdd((double*)&c[i], v);
        } break;
        case C_ADD_I64: {
            int64_t v = wave_sum_i64(act ? ld_i64(a, row) : 0);
            if (leader) atomicAdd((unsigned long long*)&c[i], (unsigned long long)v);
        } break;
        case C_MIN_I64: case C_MIN_F64: case C_MIN_ROW: {
            int64_t x = !act ? INT64_MAX : P.kind[i] == C_MIN_I64 ? ld_i64(a, row)
                      : P.kind[i] == C_MIN_F64 ? f64_ord(ld_f64(a, row)) : row;
            int64_t v = wave_min_i64(x);
            if (leader) cell_min(&c[i], v);
        } break;
        default: {                                  /* C_MAX_I64 / C_MAX_F64 / C_MAX_ROW */
            int64_t x = !act ? INT64_MIN : P.kind[i] == C_MAX_I64 ? ld_i64(a, row)
                      : P.kind[i] == C_MAX_F64 ? f64_ord(ld_f64(a, row)) : row;
            int64_t v = wave_max_i64(x);
            if (leader) cell_max(&c[i], v);
        } break;
        }
    }
}

/* One wave of rows w * 64 + lane (selected by m) into a direct-array table:
 * one group for the whole wave takes the wave-folded path, else each row
 * its own atomics.  Every lane of the wave calls this together. */
template <bool SKIP = true>
__device__ __forceinline__ void da_wave(uint64_t* tab, const GroupPlan& P, int64_t w,
                                        uint64_t m, int lane, int64_t (*slot_of)(const GroupPlan&,
                                                                                 int64_t)) {
    const bool act = (m >> lane) & 1;
    const int64_t row = w * 64 + lane;
    const int64_t slot = act ? slot_of(P, row) : -1;
    const int first = __ffsll((unsigned long long)m) - 1;
    const int64_t s0 = __shfl(slot, first);
    if (__ballot(act && slot != s0) == 0) {
        cells_accum_wave(&tab[s0 * P.ncells], P, row, act, m, lane == first);
    } else if (act) {
        cells_accum<SKIP>(&tab[slot * P.ncells], P, row);
    }
}

/* ----------------------------------------------------------- hash keys -- */

#define EMPTY_KEY 0xffffffffffffffffull

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return z;
}

/* 64-bit group key of a row: the single key's bits, or the mixed-radix
 * composite of several integer keys (stored in kmin / kstride). */
__device__ __forceinline__ uint64_t ht_key(const GroupPlan& P, int64_t row) {
    if (P.n_keys == 1 && P.kstride[0] == 0) return (uint64_t)ld_i64(P.keys[0], row);
    return (uint64_t)da_slot2(P, row, row);
}

/* Slot of key in a global open-addressing table (linear probing, keys
 * claimed with CAS), inserting it if absent; the all-ones key has the
 * extra slot after the table; -1 when the table is full. */
__device__ __forceinline__ int64_t ht_find(unsigned long long* keys, uint64_t mask, uint64_t key) {
    if (key == EMPTY_KEY) return (int64_t)mask + 1;
    uint64_t i = mix64(key) & mask;
    for (uint64_t probe = 0; probe <= mask; probe++) {
        unsigned long long k = keys[i];
        if (k == key) return (int64_t)i;
        if (k == EMPTY_KEY) {
            unsigned long long prev = atomicCAS(&keys[i], EMPTY_KEY, (unsigned long long)key);
            if (prev == EMPTY_KEY || prev == key) return (int64_t)i;
        }
        i = (i + 1) & mask;
    }
    return -1;
}

/* Merge a partial cell block `src` into `dst` with atomics. */
__device__ __forceinline__ void cells_merge(uint64_t* dst, const uint64_t* src, const GroupPlan& P) {
    atomicAdd((unsigned long long*)&dst[0], (unsigned long long)src[0]);
    for (int i = 1; i < P.ncells; i++) {
        switch (P.kind[i]) {
        case C_ADD_F64: case C_SQ_F64:
            atomicAdd((double*)&dst[i], __longlong_as_double((long long)src[i])); break;
        case C_ADD_I64: atomicAdd((unsigned long long*)&dst[i], (unsigned long long)src[i]); break;
        case C_MIN_I64: case C_MIN_F64: case C_MIN_ROW:
            cell_min(&dst[i], (int64_t)src[i]); break;
        default:
            cell_max(&dst[i], (int64_t)src[i]); break;
        }
    }
}

/* Non-atomic fold of src into an accumulator held in registers / memory. */
__device__ __forceinline__ void cells_fold(uint64_t* acc, const uint64_t* src, const GroupPlan& P) {
    acc[0] += src[0];
    for (int i = 1; i < P.ncells; i++) {
        uint64_t s = src[i];
        switch (P.kind[i]) {
        case C_ADD_F64: case C_SQ_F64:
            acc[i] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)acc[i]) +
                                                    __longlong_as_double((long long)s));
            break;
        case C_ADD_I64: acc[i] += s; break;
        case C_MIN_I64: case C_MIN_F64: case C_MIN_ROW:
            if ((int64_t)s < (int64_t)acc[i]) acc[i] = s; break;
        default:
            if ((int64_t)s > (int64_t)acc[i]) acc[i] = s; break;
        }
    }
}

/* Aggregate a of the group whose cells are `c`, written to out[a][g]
 * (emit_agg_columns, exec.c:4632-4773). */
__device__ __forceinline__ void cells_emit(const uint64_t* c, const GroupPlan& P, int a,
                                           void* dst, int64_t g) {
    const tdd_col_t& col = P.aggs[a];
    int ci = P.cell_of[a];
    int64_t cnt = (int64_t)c[0];
    double dcnt = (double)cnt;
    bool src_f = col.type == TD_F64;
    switch (P.ops[a]) {
    case OP_COUNT: ((int64_t*)dst)[g] = cnt; return;
    case OP_SUM:
        if (src_f) ((uint64_t*)dst)[g] = c[ci];
        else ((int64_t*)dst)[g] = (int64_t)c[ci];
        return;
    case OP_AVG: {
        double s = src_f ? __longlong_as_double((long long)c[ci]) : (double)(int64_t)c[ci];
        ((double*)dst)[g] = s / dcnt;
        return;
    }
    case OP_MIN: case OP_MAX:
        if (src_f) ((double*)dst)[g] = ord_f64((int64_t)c[ci]);
        else st_int(dst, g, col.esz, (int64_t)c[ci]);
        return;
    case OP_FIRST: case OP_LAST: {
        int64_t row = (int64_t)c[ci];
        if (src_f) ((double*)dst)[g] = ((const double*)col.data)[row];
        else st_int(dst, g, col.esz, ld_i64(col, row));
        return;
    }
    default: {                                         /* VAR family */
        double s = src_f ? __longlong_as_double((long long)c[ci]) : (double)(int64_t)c[ci];
        double sq = __longlong_as_double((long long)c[ci + 1]);
        double mean = cnt > 0 ? s / dcnt : 0.0;
        double vp = cnt > 0 ? sq / dcnt - mean * mean : 0.0;
        if (vp < 0) vp = 0;
        double v;
        uint16_t op = P.ops[a];
        if (op == OP_VAR_POP) v = cnt > 0 ? vp : __builtin_nan("");
        else if (op == OP_VAR) v = cnt > 1 ? vp * dcnt / (dcnt - 1) : __builtin_nan("");
        else if (op == OP_STDDEV_POP) v = cnt > 0 ? sqrt(vp) : __builtin_nan("");
        else v = cnt > 1 ? sqrt(vp * dcnt / (dcnt - 1)) : __builtin_nan("");
        ((double*)dst)[g] = v;
    }
    }
}

/* Host: fill the cell layout and output types of a plan from the aggregate
 * ops and input columns.  Returns TD_OK or TD_ERR_NYI. */
int plan_cells(GroupPlan& P);
/* Host: allocate typed output arrays for n groups. */
int alloc_group_outputs(const GroupPlan& P, int64_t n, tdd_group_result_t* out);
/* Host: direct-array table (n_slots x ncells) -> compacted outputs in slot order. */
int da_emit(const GroupPlan& P, const uint64_t* tab, int64_t n_slots, tdd_group_result_t* out);
/* Host: fold nparts per-workgroup partial tables in order, then emit. */
int da_finish(const GroupPlan& P, const uint64_t* partials, int nparts, int64_t n_slots,
              tdd_group_result_t* out);
/* Host: fill a global slot table with the cells' identities. */
int da_table_init(const GroupPlan& P, uint64_t* tab, int64_t n_slots);
extern int64_t g_da_max_slots;
