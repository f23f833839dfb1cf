/*
 * dev_runtime.hip -- device runtime of the C ABI in tdd.h: device selection,
 * the library stream, raw device memory for the arena, copies, a kernel
 * fill, event timing, and the device prefix scans shared by the kernels.
 */
#include "dev_common.h"

#include <mutex>
#include <thread>
#include <unistd.h>
#include <stdio.h>
#include <string.h>
#include <vector>

namespace tdd {

hipStream_t g_stream = nullptr;
int g_num_cu = 256;
int g_lds_block = 65536;
static int g_ready = 0;                /* 1 ok, -1 unavailable */
static std::mutex g_init_mu;
static char g_arch[32];
static char g_name[64];
static hipDeviceProp_t g_props;

int map_err(hipError_t e, const char* what, const char* file, int line) {
    fprintf(stderr, "teide_amd: HIP error %d (%s) at %s:%d in %s\n", (int)e,
            hipGetErrorString(e), file, line, what);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return TD_ERR_OOM;
    return TD_ERR_IO;
}

/* ----------------------------------------------------------- profiling -- */

struct ProfPair { hipEvent_t a, b; int name_idx; };
static bool g_prof = false;
static std::vector<ProfPair> g_prof_open;        /* recorded, not yet read */
static std::vector<hipEvent_t> g_event_pool;
static std::vector<std::string> g_prof_names;
static std::vector<int64_t> g_prof_launches;
static std::vector<double> g_prof_ms;

static hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static int name_index(const char* name) {
    for (size_t i = 0; i < g_prof_names.size(); i++)
        if (g_prof_names[i] == name) return (int)i;
    g_prof_names.emplace_back(name);
    g_prof_launches.push_back(0);
    g_prof_ms.push_back(0.0);
    return (int)g_prof_names.size() - 1;
}

int prof_begin(const char* name) {
    if (!g_prof) return -1;
    hipEvent_t a = take_event(), b = take_event();
    if (!a || !b) return -1;
    if (hipEventRecord(a, g_stream) != hipSuccess) return -1;
    g_prof_open.push_back(ProfPair{a, b, name_index(name)});
    return (int)g_prof_open.size() - 1;
}

void prof_end(int slot) {
    if (slot < 0 || slot >= (int)g_prof_open.size()) return;
    (void)hipEventRecord(g_prof_open[slot].b, g_stream);
}

static void prof_drain() {
    if (g_prof_open.empty()) return;
    (void)hipStreamSynchronize(g_stream);
    for (auto& p : g_prof_open) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g_prof_launches[p.name_idx] += 1;
            g_prof_ms[p.name_idx] += ms;
        }
        g_event_pool.push_back(p.a);
        g_event_pool.push_back(p.b);
    }
    g_prof_open.clear();
}

void* scratch(size_t bytes) { return td__arena_alloc(bytes); }
void release(void* p) { td__arena_free(p); }

}  // namespace tdd

using namespace tdd;

extern "C" int tdd_init(void) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_ready) return g_ready > 0 ? TD_OK : TD_ERR_IO;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        g_ready = -1;
        return TD_ERR_IO;
    }
    if (hipSetDevice(0) != hipSuccess || hipGetDeviceProperties(&g_props, 0) != hipSuccess) {
        g_ready = -1;
        return TD_ERR_IO;
    }
    g_num_cu = g_props.multiProcessorCount > 0 ? g_props.multiProcessorCount : 256;
    if (g_props.sharedMemPerBlock >= 32768) g_lds_block = (int)g_props.sharedMemPerBlock;
    snprintf(g_arch, sizeof(g_arch), "%s", g_props.gcnArchName);
    snprintf(g_name, sizeof(g_name), "%s", g_props.name);
    /* One stream at the default priority for everything the library runs. */
    if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
        g_ready = -1;
        return TD_ERR_IO;
    }
    g_ready = 1;
    return TD_OK;
}

extern "C" int tdd_available(void) { return tdd_init() == TD_OK; }

extern "C" int tdd_info(tdd_info_t* out) {
    if (tdd_init() != TD_OK) return TD_ERR_IO;
    memset(out, 0, sizeof(*out));
    snprintf(out->name, sizeof(out->name), "%s", g_name);
    snprintf(out->arch, sizeof(out->arch), "%s", g_arch);
    out->n_cu = g_num_cu;
    out->n_xcd_hint = 8;
    out->hbm_bytes = g_props.totalGlobalMem;
    out->lds_bytes_per_cu = g_props.maxSharedMemoryPerMultiProcessor;
    out->clock_mhz = g_props.clockRate / 1000;
    return TD_OK;
}

extern "C" int tdd_malloc(void** dptr, size_t bytes) {
    if (tdd_init() != TD_OK) return TD_ERR_IO;
    *dptr = nullptr;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *dptr = nullptr;
        return e == hipErrorOutOfMemory ? TD_ERR_OOM : TD_ERR_IO;
    }
    return TD_OK;
}

extern "C" int tdd_free(void* dptr) {
    if (!dptr) return TD_OK;
    HC(hipStreamSynchronize(g_stream));
    HC(hipFree(dptr));
    return TD_OK;
}

extern "C" int tdd_h2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return TD_OK;
    HC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
    HC(hipStreamSynchronize(g_stream));
    return TD_OK;
}

/* ------------------------------------------------------ file staging -- */

#define STAGE_BYTES   ((size_t)64 << 20)   /* per staging buffer */
#define STAGE_READERS 4                    /* reader threads per chunk */

static std::mutex g_stage_lock;
static void* g_stage[2];
static hipEvent_t g_stage_ev[2];

static bool pread_all(int fd, char* p, size_t n, int64_t off) {
    while (n) {
        ssize_t r = pread(fd, p, n, (off_t)off);
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += r;
    }
    return true;
}

extern "C" int tdd_file_h2d(int fd, int64_t off, void* dst, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_stage_lock);
    if (!g_stage[0]) {
        for (int b = 0; b < 2; b++) {
            HC(hipHostMalloc(&g_stage[b], STAGE_BYTES, hipHostMallocDefault));
            HC(hipEventCreateWithFlags(&g_stage_ev[b], hipEventDisableTiming));
            HC(hipEventRecord(g_stage_ev[b], g_stream));
        }
    }
    size_t pos = 0;
    int b = 0;
    while (pos < bytes) {
        const size_t n = bytes - pos < STAGE_BYTES ? bytes - pos : STAGE_BYTES;
        HC(hipEventSynchronize(g_stage_ev[b]));          /* its previous copy is done */
        char* buf = (char*)g_stage[b];
        const size_t part = ((n + STAGE_READERS - 1) / STAGE_READERS + 4095) & ~(size_t)4095;
        bool ok[STAGE_READERS];
        std::thread th[STAGE_READERS];
        for (int r = 0; r < STAGE_READERS; r++) {
            const size_t s0 = (size_t)r * part;
            ok[r] = true;
            if (s0 >= n) continue;
            const size_t len = n - s0 < part ? n - s0 : part;
            th[r] = std::thread([=, &ok] { ok[r] = pread_all(fd, buf + s0, len, off + (int64_t)(pos + s0)); });
        }
        bool good = true;
        for (int r = 0; r < STAGE_READERS; r++) {
            if (th[r].joinable()) th[r].join();
            good = good && ok[r];
        }
        if (!good) {
            HC(hipStreamSynchronize(g_stream));
            return TD_ERR_IO;
        }
        HC(hipMemcpyAsync((char*)dst + pos, buf, n, hipMemcpyHostToDevice, g_stream));
        HC(hipEventRecord(g_stage_ev[b], g_stream));
        pos += n;
        b ^= 1;
    }
    HC(hipStreamSynchronize(g_stream));
    return TD_OK;
}

extern "C" int tdd_d2h(void* dst, const void* src, size_t bytes) {
    if (!bytes) return TD_OK;
    HC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
    HC(hipStreamSynchronize(g_stream));
    return TD_OK;
}

/* Streaming copy: each thread moves 4 x 16 bytes per step, all loads issued
 * before the stores. */
typedef int rt_v4i32 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy(rt_v4i32* dst, const rt_v4i32* src, int64_t n16) {
    int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        rt_v4i32 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

extern "C" int tdd_d2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return TD_OK;
    if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15) == 0) {
        int64_t n16 = (int64_t)(bytes / 16);
        int grid = stream_grid(ceil_div(n16, 256 * 4), 8);
        LAUNCH("copy", k_copy, grid, 256, 0, (rt_v4i32*)dst, (const rt_v4i32*)src, n16);
        return TD_OK;
    }
    HC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
    return TD_OK;
}

/* Fill with an 8-byte pattern by kernel (not a memset node): bytes beyond
 * the last whole word are filled from the pattern's low bytes. */
__global__ void k_fill(uint64_t* dst, uint64_t pattern, int64_t nwords, uint8_t* tail,
                       int ntail) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride) dst[i] = pattern;
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail)
        tail[threadIdx.x] = (uint8_t)(pattern >> (8 * threadIdx.x));
}

extern "C" int tdd_fill(void* dst, uint64_t pattern, size_t bytes) {
    if (!bytes) return TD_OK;
    if (((uintptr_t)dst & 7) != 0) return TD_ERR_RANGE;
    int64_t nw = (int64_t)(bytes / 8);
    int ntail = (int)(bytes % 8);
    int grid = stream_grid(ceil_div(nw > 0 ? nw : 1, 256 * 4));
    LAUNCH("fill", k_fill, grid, 256, 0, (uint64_t*)dst, pattern, nw,
           (uint8_t*)dst + nw * 8, ntail);
    return TD_OK;
}

extern "C" int tdd_sync(void) {
    HC(hipStreamSynchronize(g_stream));
    return TD_OK;
}

extern "C" int tdd_mem_info(uint64_t* free_bytes, uint64_t* total_bytes) {
    if (tdd_init() != TD_OK) return TD_ERR_IO;
    size_t f = 0, t = 0;
    HC(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return TD_OK;
}

extern "C" int tdd_prof_enable(int on) {
    prof_drain();
    g_prof = on != 0;
    return TD_OK;
}

extern "C" int tdd_prof_reset(void) {
    prof_drain();
    for (auto& v : g_prof_launches) v = 0;
    for (auto& v : g_prof_ms) v = 0.0;
    return TD_OK;
}

extern "C" int tdd_prof_read(tdd_prof_row_t* rows, int cap, int* n_rows) {
    prof_drain();
    int n = 0;
    for (size_t i = 0; i < g_prof_names.size() && n < cap; i++) {
        if (!g_prof_launches[i]) continue;
        memset(&rows[n], 0, sizeof(rows[n]));
        snprintf(rows[n].name, sizeof(rows[n].name), "%s", g_prof_names[i].c_str());
        rows[n].launches = g_prof_launches[i];
        rows[n].total_ms = g_prof_ms[i];
        n++;
    }
    *n_rows = n;
    return TD_OK;
}

/* ---------------------------------------------------------------- scans -- */
/* Three-phase exclusive scan: per-block sums, one block scans those sums,
 * then each block rescans its chunk from its base.  Chunk = 256 threads x
 * 16 elements.  Inputs of at most SCAN_ONE chunks take one launch: a single
 * block walks the chunks in order, carrying the running total. */

#define SCAN_T 256
#define SCAN_V 16
#define SCAN_CHUNK (SCAN_T * SCAN_V)
#define SCAN_ONE 4

template <class T>
__device__ __forceinline__ int64_t scan_ld(const T* in, int64_t i, int64_t n) {
    return i < n ? (int64_t)in[i] : 0;
}

/* Block-wide exclusive scan of one value per thread; returns the prefix,
 * *total receives the block sum. */
__device__ int64_t block_excl_scan(int64_t v, int64_t* total) {
    __shared__ int64_t wsum[SCAN_T / WAVE];
    int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    int64_t x = v;
    for (int d = 1; d < WAVE; d <<= 1) {
        int64_t y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    if (lane == WAVE - 1) wsum[w] = x;
    __syncthreads();
    int64_t base = 0, tot = 0;
    for (int i = 0; i < SCAN_T / WAVE; i++) {
        if (i < w) base += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_scan_sums(const T* in, int64_t n, int64_t* sums) {
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_V;
    int64_t s = 0;
    for (int j = 0; j < SCAN_V; j++) s += scan_ld(in, base + j, n);
    int64_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_top(int64_t* sums, int64_t nb) {
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += SCAN_T) {
        int64_t i = b0 + threadIdx.x;
        int64_t v = i < nb ? sums[i] : 0;
        int64_t tot;
        int64_t p = block_excl_scan(v, &tot);
        if (i < nb) sums[i] = carry + p;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const T* in, int64_t n,
                                                       const int64_t* sums, int64_t* out,
                                                       int64_t* out2) {
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_V;
    int64_t v[SCAN_V], s = 0;
    for (int j = 0; j < SCAN_V; j++) { v[j] = scan_ld(in, base + j, n); s += v[j]; }
    int64_t tot;
    int64_t p = block_excl_scan(s, &tot) + sums[blockIdx.x];
    for (int j = 0; j < SCAN_V; j++) {
        if (base + j < n) {
            out[base + j] = p;
            if (out2) out2[base + j] = p;
        }
        p += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_T - 1) out[n] = sums[gridDim.x];
}

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_scan_one(const T* in, int64_t n, int64_t* out,
                                                     int64_t* out2) {
    int64_t carry = 0;
    for (int64_t c0 = 0; c0 < n; c0 += SCAN_CHUNK) {     /* uniform across the block */
        const int64_t base = c0 + threadIdx.x * SCAN_V;
        int64_t v[SCAN_V], s = 0;
        for (int j = 0; j < SCAN_V; j++) { v[j] = scan_ld(in, base + j, n); s += v[j]; }
        int64_t tot;
        int64_t p = block_excl_scan(s, &tot) + carry;
        for (int j = 0; j < SCAN_V; j++) {
            if (base + j < n) {
                out[base + j] = p;
                if (out2) out2[base + j] = p;
            }
            p += v[j];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

template <class T>
static int scan_impl(const T* in, int64_t n, int64_t* out, const char* name,
                     int64_t* out2 = nullptr) {
    if (n <= 0) {
        int64_t z = 0;
        HC(hipMemcpyAsync(out, &z, 8, hipMemcpyHostToDevice, g_stream));
        HC(hipStreamSynchronize(g_stream));
        return TD_OK;
    }
    int64_t nb = ceil_div(n, SCAN_CHUNK);
    if (nb <= SCAN_ONE) {
        LAUNCH(name, k_scan_one<T>, 1, SCAN_T, 0, in, n, out, out2);
        return TD_OK;
    }
    SCRATCH(sums, (size_t)(nb + 1) * 8);
    LAUNCH(name, k_scan_sums<T>, (int)nb, SCAN_T, 0, in, n, sums.as<int64_t>());
    LAUNCH(name, k_scan_top, 1, SCAN_T, 0, sums.as<int64_t>(), nb);
    LAUNCH(name, k_scan_apply<T>, (int)nb, SCAN_T, 0, in, n, (const int64_t*)sums.as<int64_t>(), out,
           out2);
    return TD_OK;
}

namespace tdd {
int scan_u16_i64(const uint16_t* in, int64_t n, int64_t* out) { return scan_impl(in, n, out, "scan"); }
int scan_i64(const int64_t* in, int64_t n, int64_t* out) { return scan_impl(in, n, out, "scan"); }
int scan_u32_i64(const uint32_t* in, int64_t n, int64_t* out) { return scan_impl(in, n, out, "scan"); }
int scan_u32_i64_dup(const uint32_t* in, int64_t n, int64_t* out, int64_t* out2) {
    return scan_impl(in, n, out, "scan", out2);
}
}  // namespace tdd
