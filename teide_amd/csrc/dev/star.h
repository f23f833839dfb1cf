/*
 * star.h -- folding tiles of probe rows into a group table (LDS or global)
 * for the star-shaped fused kernels: each aggregate cell reads its input
 * from a tile slot chosen by a uniform switch, so every access inside the
 * unrolled row loop has a compile-time register index.
 */
#pragma once

#include "group_plan.h"
#include "tile.h"

/* Fold the R rows of one tile into cell i of their slots s[j] (< 0: skip),
 * reading the input from tile slot Q. */
template <int R, int Q>
__device__ __forceinline__ void acc_cell(uint64_t* tab, int ncells, int i, int kind, bool col_f64,
                                         const int64_t (&cv)[TL_NC][R], const int32_t (&s)[R]) {
    switch (kind) {
    case C_ADD_F64:
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0)
                atomicAdd((double*)&tab[(int64_t)s[j] * ncells + i],
                          col_f64 ? __longlong_as_double(cv[Q][j]) : (double)cv[Q][j]);
        break;
    case C_SQ_F64:
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0) {
                double v = col_f64 ? __longlong_as_double(cv[Q][j]) : (double)cv[Q][j];
                atomicAdd((double*)&tab[(int64_t)s[j] * ncells + i], v * v);
            }
        break;
    case C_ADD_I64:
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0)
                atomicAdd((unsigned long long*)&tab[(int64_t)s[j] * ncells + i],
                          (unsigned long long)cv[Q][j]);
        break;
    case C_MIN_I64:
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0) atomicMin((long long*)&tab[(int64_t)s[j] * ncells + i], (long long)cv[Q][j]);
        break;
    case C_MAX_I64:
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0) atomicMax((long long*)&tab[(int64_t)s[j] * ncells + i], (long long)cv[Q][j]);
        break;
    case C_MIN_F64:
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0)
                atomicMin((long long*)&tab[(int64_t)s[j] * ncells + i],
                          (long long)f64_ord(col_f64 ? __longlong_as_double(cv[Q][j]) : (double)cv[Q][j]));
        break;
    case C_MAX_F64:
#pragma unroll
        for (int j = 0; j < R; j++)
            if (s[j] >= 0)
                atomicMax((long long*)&tab[(int64_t)s[j] * ncells + i],
                          (long long)f64_ord(col_f64 ? __longlong_as_double(cv[Q][j]) : (double)cv[Q][j]));
        break;
    default: break;                     /* row cells (FIRST / LAST) are not fused */
    }
}

/* All cells of a tile: COUNT, then each aggregate cell from its slot. */
template <int R>
__device__ __forceinline__ void acc_tile(uint64_t* tab, const GroupPlan& P, const uint8_t* cell_slot,
                                         const TileCols& T, const int64_t (&cv)[TL_NC][R],
                                         const int32_t (&s)[R]) {
#pragma unroll
    for (int j = 0; j < R; j++)
        if (s[j] >= 0) atomicAdd((unsigned long long*)&tab[(int64_t)s[j] * P.ncells], 1ull);
    for (int i = 1; i < P.ncells; i++) {
        const int q = cell_slot[i];
        const bool f = T.col[q].type == TD_F64;
        switch (q) {
        case 0:  acc_cell<R, 0>(tab, P.ncells, i, P.kind[i], f, cv, s); break;
        case 1:  acc_cell<R, 1>(tab, P.ncells, i, P.kind[i], f, cv, s); break;
        case 2:  acc_cell<R, 2>(tab, P.ncells, i, P.kind[i], f, cv, s); break;
        default: acc_cell<R, 3>(tab, P.ncells, i, P.kind[i], f, cv, s); break;
        }
    }
}
