/*
 * dev_common.h -- internals shared by the HIP translation units of the
 * device layer (gfx950 / MI355X, wave64).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>

#include <teide/td.h>
#include <teide/tdd.h>

/* ----------------------------------------------------------- runtime -- */

namespace tdd {

extern hipStream_t g_stream;
extern int g_num_cu;
extern int g_lds_block;                 /* LDS bytes one workgroup may allocate */

int map_err(hipError_t e, const char* what, const char* file, int line);

/* Kernel timing (tdd_prof_*): returns an event pair index or -1. */
int  prof_begin(const char* name);
void prof_end(int slot);

void* scratch(size_t bytes);          /* arena-backed device scratch */
void  release(void* p);

}  // namespace tdd

#define HC(x)                                                             \
    do {                                                                  \
        hipError_t e__ = (x);                                             \
        if (e__ != hipSuccess) return tdd::map_err(e__, #x, __FILE__, __LINE__); \
    } while (0)

/* Launch on the library stream, bracketed by timing events when enabled. */
#define LAUNCH(name, kern, grid, block, shmem, ...)                           \
    do {                                                                      \
        int ps__ = tdd::prof_begin(name);                                     \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shmem, tdd::g_stream, \
                           __VA_ARGS__);                                      \
        tdd::prof_end(ps__);                                                  \
        HC(hipGetLastError());                                                \
    } while (0)

/* A cancellation point between launches (td_cancel, host engine): the
 * caller's scratch is freed by its destructors on the way out. */
extern "C" int td__cancel_poll(void);
#define POLL_CANCEL()                                            \
    do {                                                         \
        if (td__cancel_poll()) return TD_ERR_CANCEL;             \
    } while (0)

/* Scoped arena scratch: freed on scope exit. */
struct Scratch {
    void* p = nullptr;
    explicit Scratch(size_t bytes) { p = tdd::scratch(bytes ? bytes : 8); }
    ~Scratch() { if (p) tdd::release(p); }
    template <class T> T* as() const { return static_cast<T*>(p); }
    bool ok() const { return p != nullptr; }
    /* replace the block by a new one of `bytes` */
    bool reset(size_t bytes) {
        if (p) tdd::release(p);
        p = tdd::scratch(bytes ? bytes : 8);
        return p != nullptr;
    }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
};

#define SCRATCH(var, bytes)                                      \
    Scratch var(bytes);                                          \
    if (!var.ok()) return TD_ERR_OOM

/* ------------------------------------------------------ device helpers -- */

#define WAVE 64

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

/* Count of set bits of m in lanes below this one. */
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* Widen any integer-representable column element to int64, the reference's
 * read_col_i64 rule (exec.c:90-109): SYM by index width (unsigned below
 * W64), I32/DATE/TIME and I16 sign-extended, BOOL/U8 zero-extended. */
__device__ __forceinline__ int64_t ld_i64(const tdd_col_t& c, int64_t r) {
    switch (c.type) {
    case TD_I64: case TD_TIMESTAMP: case TD_F64:
        return ((const int64_t*)c.data)[r];
    case TD_SYM:
        switch (c.attrs & TD_SYM_W_MASK) {
        case TD_SYM_W8:  return ((const uint8_t*)c.data)[r];
        case TD_SYM_W16: return ((const uint16_t*)c.data)[r];
        case TD_SYM_W32: return ((const uint32_t*)c.data)[r];
        default:         return ((const int64_t*)c.data)[r];
        }
    case TD_I32: case TD_DATE: case TD_TIME:
        return ((const int32_t*)c.data)[r];
    case TD_I16:
        return ((const int16_t*)c.data)[r];
    default:
        return ((const uint8_t*)c.data)[r];
    }
}

__device__ __forceinline__ double ld_f64(const tdd_col_t& c, int64_t r) {
    if (c.type == TD_F64) return ((const double*)c.data)[r];
    return (double)ld_i64(c, r);
}

/* Store an int64 into a column element of the given width. */
__device__ __forceinline__ void st_int(void* dst, int64_t i, int esz, int64_t v) {
    switch (esz) {
    case 1: ((uint8_t*)dst)[i] = (uint8_t)v; break;
    case 2: ((uint16_t*)dst)[i] = (uint16_t)v; break;
    case 4: ((uint32_t*)dst)[i] = (uint32_t)v; break;
    default: ((int64_t*)dst)[i] = v; break;
    }
}

/* Order-preserving map of a double to int64 (signed compare of the images
 * orders like the doubles, -0.0 below +0.0); used for atomic MIN/MAX. */
__device__ __forceinline__ int64_t f64_ord(double d) {
    int64_t i = __double_as_longlong(d);
    return i ^ ((i >> 63) & 0x7fffffffffffffffLL);
}
__device__ __forceinline__ double ord_f64(int64_t i) {
    return __longlong_as_double(i ^ ((i >> 63) & 0x7fffffffffffffffLL));
}

__device__ __forceinline__ bool sel_bit(const uint64_t* bits, int64_t r) {
    return (bits[r >> 6] >> (r & 63)) & 1ull;
}

static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

/* Grid size for a streaming kernel: enough workgroups to fill the 256 CUs
 * several times over, never more than the work. */
static inline int stream_grid(int64_t units, int per_cu = 8) {
    int64_t cap = (int64_t)tdd::g_num_cu * per_cu;
    if (units < 1) return 1;
    return (int)(units < cap ? units : cap);
}

/* Workgroups of `kern` that are resident on the whole chip at once, capped
 * by the work available: the grid of a grid-stride kernel. */
template <class K>
static inline int resident_grid(K kern, int block, size_t shmem, int64_t units) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, block, shmem) != hipSuccess ||
        per < 1)
        per = 1;
    int64_t cap = (int64_t)tdd::g_num_cu * per;
    if (units < 1) return 1;
    return (int)(units < cap ? units : cap);
}

/* Integer launch parameter from the environment (read once per call site),
 * else the default: used for sweeps of tile shapes. */
#define TUNE_INT(name, def)                                           \
    ([]() {                                                           \
        static int v__ = -1;                                          \
        if (v__ < 0) {                                                \
            const char* e__ = getenv(name);                           \
            v__ = e__ ? atoi(e__) : (def);                            \
        }                                                             \
        return v__;                                                   \
    }())

/* ----------------------------------------------------- host-side scans -- */

namespace tdd {
/* Exclusive prefix sums on the device: out[i] = sum(in[0..i)), out[n] = total
 * (out must hold n + 1 entries). */
int scan_u16_i64(const uint16_t* in, int64_t n, int64_t* out);
int scan_i64(const int64_t* in, int64_t n, int64_t* out);
int scan_u32_i64(const uint32_t* in, int64_t n, int64_t* out);
/* the same, each prefix also written to out2[0 .. n) */
int scan_u32_i64_dup(const uint32_t* in, int64_t n, int64_t* out, int64_t* out2);
}  // namespace tdd
