/*
 * k_sort.hip -- stable multi-key sort permutation for SORT / HEAD(SORT)
 * over (group) result tables.
 *
 * Each key is mapped to an order-preserving 64-bit word exactly as the
 * reference's radix path encodes it (radix_encode_fn, exec.c:2739-2870):
 * signed integers with the sign bit flipped, f64 by its IEEE bits (negative
 * values inverted, NaN placed first or last by nulls_first), SYM by the
 * rank of its string, descending keys inverted.  The permutation is built
 * least-significant key first, each pass a stable LSD radix sort of
 * (word, row) pairs, so rows equal on every key keep their input order --
 * the order a stable multi-key sort gives.
 */
#include <rocprim/device/device_radix_sort.hpp>

#include "dev_common.h"

__device__ __forceinline__ uint64_t sort_word(const tdd_sort_key_t& K, int64_t r) {
    const tdd_col_t& c = K.col;
    uint64_t e;
    switch (c.type) {
    case TD_F64: {
        const uint64_t bits = ((const uint64_t*)c.data)[r];
        if ((bits & 0x7FF0000000000000ull) == 0x7FF0000000000000ull &&
            (bits & 0x000FFFFFFFFFFFFFull)) {                    /* NaN */
            const uint64_t nan_e = (K.nulls_first ^ K.desc) ? 0ull : ~0ull;
            return K.desc ? ~nan_e : nan_e;
        }
        const uint64_t mask = (0ull - (bits >> 63)) | (1ull << 63);
        e = bits ^ mask;
        break;
    }
    case TD_SYM:
        e = (uint64_t)K.rank[(uint32_t)ld_i64(c, r)];
        break;
    case TD_I32: case TD_DATE: case TD_TIME:
        e = (uint64_t)((uint32_t)((const int32_t*)c.data)[r] ^ 0x80000000u);
        break;
    case TD_I16:
        e = (uint64_t)((uint16_t)((const int16_t*)c.data)[r] ^ 0x8000u);
        break;
    case TD_BOOL: case TD_U8:
        e = (uint64_t)((const uint8_t*)c.data)[r];
        break;
    default:                                                      /* I64, TIMESTAMP */
        e = (uint64_t)((const int64_t*)c.data)[r] ^ (1ull << 63);
        break;
    }
    return K.desc ? ~e : e;
}

__global__ void k_sort_iota(int64_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = i;
}

/* words[i] = key word of row perm[i] */
__global__ void k_sort_encode(tdd_sort_key_t K, const int64_t* perm, int64_t n, uint64_t* words) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        words[i] = sort_word(K, perm[i]);
}

extern "C" int tdd_sort_perm(const tdd_sort_key_t* keys, int nk, int64_t n, int64_t* perm) {
    if (n <= 0) return TD_OK;
    if (nk < 1) return TD_ERR_NYI;
    const int grid = stream_grid(ceil_div(n, 256 * 4));
    LAUNCH("sort", k_sort_iota, grid, 256, 0, perm, n);
    SCRATCH(wa, (size_t)n * 8);
    SCRATCH(wb, (size_t)n * 8);
    SCRATCH(pb, (size_t)n * 8);
    size_t temp = 0;
    HC(rocprim::radix_sort_pairs(nullptr, temp, wa.as<uint64_t>(), wb.as<uint64_t>(), perm,
                                 pb.as<int64_t>(), (size_t)n, 0, 64, tdd::g_stream));
    SCRATCH(tmp, temp);
    for (int k = nk - 1; k >= 0; k--) {
        LAUNCH("sort", k_sort_encode, grid, 256, 0, keys[k], (const int64_t*)perm, n,
               wa.as<uint64_t>());
        int prof = tdd::prof_begin("sort");
        hipError_t e = rocprim::radix_sort_pairs(tmp.p, temp, wa.as<uint64_t>(), wb.as<uint64_t>(),
                                                 (const int64_t*)perm, pb.as<int64_t>(), (size_t)n,
                                                 0, 64, tdd::g_stream);
        tdd::prof_end(prof);
        HC(e);
        HC(hipMemcpyAsync(perm, pb.p, (size_t)n * 8, hipMemcpyDeviceToDevice, tdd::g_stream));
    }
    return TD_OK;
}
