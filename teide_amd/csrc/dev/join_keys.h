/*
 * join_keys.h -- key words, hashing and key comparison shared by the join
 * kernels (k_join.hip, k_rjoin.hip).
 */
#pragma once

#include "dev_common.h"

__device__ __forceinline__ uint64_t jmix(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return z;
}

/* Keys of one side.  exact: a single integer key compared by value (the key
 * word is the value); otherwise the key word is a hash of all key columns
 * and candidate pairs are verified against the columns. */
struct JKeys {
    tdd_col_t cols[TDD_MAX_KEYS];
    int32_t   n, exact;
};

__device__ __forceinline__ uint64_t key_word(const JKeys& K, int64_t row) {
    if (K.exact) return (uint64_t)ld_i64(K.cols[0], row);
    uint64_t h = 0x2545f4914f6cdd1dull;
    for (int k = 0; k < K.n; k++) {
        uint64_t w;
        if (K.cols[k].type == TD_F64) {
            double d = ((const double*)K.cols[k].data)[row];
            if (d == 0.0) d = 0.0;               /* -0.0 == 0.0 must hash alike */
            w = (uint64_t)__double_as_longlong(d);
        } else {
            w = (uint64_t)ld_i64(K.cols[k], row);
        }
        h = jmix(h ^ w) + (uint64_t)k;
    }
    return h;
}

/* join_keys_eq (exec.c:7754-7768): F64 keys with ==, others as int64. */
__device__ __forceinline__ bool keys_equal(const JKeys& L, int64_t l, const JKeys& R, int64_t r) {
    for (int k = 0; k < L.n; k++) {
        if (L.cols[k].type == TD_F64) {
            if (((const double*)L.cols[k].data)[l] != ((const double*)R.cols[k].data)[r]) return false;
        } else if (ld_i64(L.cols[k], l) != ld_i64(R.cols[k], r)) {
            return false;
        }
    }
    return true;
}

__device__ __forceinline__ bool row_selected(const uint64_t* bits, int64_t row) {
    return !bits || ((bits[row >> 6] >> (row & 63)) & 1ull);
}

