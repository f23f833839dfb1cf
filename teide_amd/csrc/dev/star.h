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

/* Cell i of one slot for a whole wave's tile (all live rows of the wave in
 * that slot): rows folded in registers, then across the wave, one atomic
 * by the leader.  Every lane of the wave calls this together. */
template <int R, int Q>
__device__ __forceinline__ void acc_cell_wave(uint64_t* c, int kind, bool col_f64,
                                              const int64_t (&cv)[TL_NC][R], uint32_t live,
                                              bool leader) {
    switch (kind) {
    case C_ADD_F64: case C_SQ_F64: {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((live >> j) & 1) {
                const double v = col_f64 ? __longlong_as_double(cv[Q][j]) : (double)cv[Q][j];
                a += kind == C_SQ_F64 ? v * v : v;
            }
        a = wave_sum_f64(a);
        if (leader) atomicAdd((double*)c, a);
    } break;
    case C_ADD_I64: {
        int64_t a = 0;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((live >> j) & 1) a += cv[Q][j];
        a = wave_sum_i64(a);
        if (leader) atomicAdd((unsigned long long*)c, (unsigned long long)a);
    } break;
    case C_MIN_I64: case C_MIN_F64: {
        int64_t a = INT64_MAX;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((live >> j) & 1) {
                const int64_t x = kind == C_MIN_I64 ? cv[Q][j]
                    : f64_ord(col_f64 ? __longlong_as_double(cv[Q][j]) : (double)cv[Q][j]);
                a = x < a ? x : a;
            }
        a = wave_min_i64(a);
        if (leader) cell_min(c, a);
    } break;
    case C_MAX_I64: case C_MAX_F64: {
        int64_t a = INT64_MIN;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((live >> j) & 1) {
                const int64_t x = kind == C_MAX_I64 ? cv[Q][j]
                    : f64_ord(col_f64 ? __longlong_as_double(cv[Q][j]) : (double)cv[Q][j]);
                a = x > a ? x : a;
            }
        a = wave_max_i64(a);
        if (leader) cell_max(c, a);
    } break;
    default: break;                     /* row cells (FIRST / LAST) are not fused */
    }
}

/* All cells of one slot `c` for a wave's tile whose live rows (bits of
 * `live`, per lane) all fall in that slot. */
template <int R>
__device__ __forceinline__ void acc_tile_wave(uint64_t* c, const GroupPlan& P,
                                              const uint8_t* cell_slot, const TileCols& T,
                                              const int64_t (&cv)[TL_NC][R], uint32_t live,
                                              bool leader) {
    const int64_t n = wave_sum_i64((int64_t)__popc(live));
    if (leader) atomicAdd((unsigned long long*)&c[0], (unsigned long long)n);
    for (int i = 1; i < P.ncells; i++) {
        const int q = cell_slot[i];
        const bool f = T.col[q].type == TD_F64;
        switch (q) {
        case 0:  acc_cell_wave<R, 0>(&c[i], P.kind[i], f, cv, live, leader); break;
        case 1:  acc_cell_wave<R, 1>(&c[i], P.kind[i], f, cv, live, leader); break;
        case 2:  acc_cell_wave<R, 2>(&c[i], P.kind[i], f, cv, live, leader); break;
        default: acc_cell_wave<R, 3>(&c[i], P.kind[i], f, cv, live, leader); break;
        }
    }
}
