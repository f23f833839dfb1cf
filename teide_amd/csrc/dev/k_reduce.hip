/*
 * k_reduce.hip -- scalar reductions that the group-by kernels do not cover:
 * the product of a column (OP_PROD) and the canonical f64 words that COUNT
 * DISTINCT counts.
 *
 * PROD runs as a fixed-shape two-level tree: every lane multiplies a
 * grid-strided run of rows, the wave folds its 64 lanes with cross-lane
 * shuffles, the workgroup folds its waves through LDS and writes one
 * partial; a single workgroup then folds the partials in order.  The grid
 * depends only on the row count, so a given input always gives the same
 * f64 bits.  i64 products wrap (two's complement), which makes them exact
 * whatever the order.
 */
#include "dev_common.h"

#define RED_T 256
#define RED_MAX_BLOCKS 1024

template <class T> struct MulOp;
template <> struct MulOp<double> {
    __device__ static double one() { return 1.0; }
    __device__ static double mul(double a, double b) { return a * b; }
};
template <> struct MulOp<uint64_t> {
    __device__ static uint64_t one() { return 1ull; }
    __device__ static uint64_t mul(uint64_t a, uint64_t b) { return a * b; }
};

template <class T> __device__ __forceinline__ T ld_val(const tdd_col_t& c, int64_t r);
template <> __device__ __forceinline__ double ld_val<double>(const tdd_col_t& c, int64_t r) {
    return ((const double*)c.data)[r];
}
template <> __device__ __forceinline__ uint64_t ld_val<uint64_t>(const tdd_col_t& c, int64_t r) {
    return (uint64_t)ld_i64(c, r);
}

/* Product of a wave's lanes, in a fixed butterfly order. */
template <class T> __device__ __forceinline__ T wave_prod(T v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T o;
        if constexpr (sizeof(T) == 8) {
            uint64_t bits;
            __builtin_memcpy(&bits, &v, 8);
            uint32_t lo = __shfl_xor((uint32_t)bits, d), hi = __shfl_xor((uint32_t)(bits >> 32), d);
            uint64_t ob = ((uint64_t)hi << 32) | lo;
            __builtin_memcpy(&o, &ob, 8);
        }
        v = MulOp<T>::mul(v, o);
    }
    return v;
}

/* Block-level product of `in` (n values of column c, or n partials when
 * c is null) into out[blockIdx.x]. */
template <class T>
__global__ __launch_bounds__(RED_T) void k_prod(tdd_col_t c, const T* partials, int64_t n,
                                                T* out) {
    __shared__ T wp[RED_T / WAVE];
    T acc = MulOp<T>::one();
    const int64_t stride = (int64_t)gridDim.x * RED_T;
    for (int64_t r = (int64_t)blockIdx.x * RED_T + threadIdx.x; r < n; r += stride)
        acc = MulOp<T>::mul(acc, partials ? partials[r] : ld_val<T>(c, r));
    acc = wave_prod(acc);
    const int w = threadIdx.x / WAVE;
    if (lane_id() == 0) wp[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T p = wp[0];
        for (int i = 1; i < RED_T / WAVE; i++) p = MulOp<T>::mul(p, wp[i]);
        out[blockIdx.x] = p;
    }
}

template <class T> static int prod_run(const tdd_col_t* col, int64_t n, void* out8) {
    int blocks = (int)(n / (RED_T * 8));
    if (blocks < 1) blocks = 1;
    if (blocks > RED_MAX_BLOCKS) blocks = RED_MAX_BLOCKS;
    SCRATCH(part, (size_t)(blocks + 1) * sizeof(T));
    T* p = part.as<T>();
    LAUNCH("reduce_prod", (k_prod<T>), blocks, RED_T, 0, *col, (const T*)nullptr, n, p);
    LAUNCH("reduce_prod", (k_prod<T>), 1, RED_T, 0, *col, (const T*)p, (int64_t)blocks,
           p + blocks);
    HC(hipMemcpyAsync(out8, p + blocks, 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}

extern "C" int tdd_reduce_prod(const tdd_col_t* col, int64_t n, void* out8) {
    if (!col || !out8 || n < 0) return TD_ERR_DOMAIN;
    if (n == 0 || !col->data) {
        if (col->type == TD_F64) { double one = 1.0; memcpy(out8, &one, 8); }
        else { int64_t one = 1; memcpy(out8, &one, 8); }
        return TD_OK;
    }
    return col->type == TD_F64 ? prod_run<double>(col, n, out8) : prod_run<uint64_t>(col, n, out8);
}

/* out[i] = bits of in[i] with every NaN made the one canonical NaN and -0.0
 * made +0.0, so that equal values (and all NaNs) share one word. */
__global__ void k_f64_canon(const double* in, int64_t n, int64_t* out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double d = in[i];
        int64_t w = __double_as_longlong(d);
        if (d != d) w = 0x7ff8000000000000LL;
        else if (d == 0.0) w = 0;
        out[i] = w;
    }
}

extern "C" int tdd_f64_canon(const double* in, int64_t n, int64_t* out) {
    if (n <= 0) return TD_OK;
    if (!in || !out) return TD_ERR_DOMAIN;
    LAUNCH("f64_canon", k_f64_canon, stream_grid(ceil_div(n, 256)), 256, 0, in, n, out);
    return TD_OK;
}
