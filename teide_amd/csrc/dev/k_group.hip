/*
 * k_group.hip -- group-by kernels.
 *
 * Direct-array path (integer keys with a small slot space, the device form
 * of the reference's DA path, exec.c:5948-6404): slot = sum_k (key_k -
 * min_k) * stride_k.  When the slot table fits in LDS every workgroup
 * aggregates into its own LDS copy with LDS atomics and writes it out as a
 * partial; a second kernel folds the partials in workgroup order.  Larger
 * slot spaces aggregate with global atomics into one table.
 *
 * Hash path (everything else, the device form of exec.c:3868-4526): a
 * global open-addressing table of 64-bit group keys sized from a
 * HyperLogLog estimate, with cells updated by global atomics.  A small
 * per-workgroup LDS table absorbs heavy hitters first (a Zipf key can own
 * a tenth of all rows) and is merged into the global table at the end.
 */
#include "group_plan.h"
#include "star.h"
#include "tile.h"

#include <math.h>
#include <stdio.h>

#define GT 512                      /* threads per group-by workgroup */
#define LDS_TABLE_BYTES (64 * 1024) /* LDS slot table budget per workgroup */

static bool is_int_type(int8_t t) {
    return t == TD_I64 || t == TD_SYM || t == TD_I32 || t == TD_TIMESTAMP || t == TD_DATE ||
           t == TD_TIME || t == TD_BOOL || t == TD_U8 || t == TD_I16;
}

int plan_cells(GroupPlan& P) {
    int n = 1;
    P.kind[0] = C_COUNT;
    P.col[0] = 0;
    for (int a = 0; a < P.n_aggs; a++) {
        const tdd_col_t& c = P.aggs[a];
        bool f = c.data && c.type == TD_F64;
        uint16_t op = P.ops[a];
        if (op != OP_COUNT && !c.data) return TD_ERR_NYI;
        if (op != OP_COUNT && c.type != TD_F64 && !is_int_type(c.type)) return TD_ERR_TYPE;
        P.cell_of[a] = (uint8_t)n;
        switch (op) {
        case OP_COUNT: P.out_types[a] = TD_I64; P.cell_of[a] = 0; break;
        case OP_SUM:
            P.out_types[a] = f ? TD_F64 : TD_I64;
            P.kind[n] = f ? C_ADD_F64 : C_ADD_I64; P.col[n++] = (uint8_t)a; break;
        case OP_AVG:
            P.out_types[a] = TD_F64;
            P.kind[n] = f ? C_ADD_F64 : C_ADD_I64; P.col[n++] = (uint8_t)a; break;
        case OP_MIN:
            P.out_types[a] = c.type;
            P.kind[n] = f ? C_MIN_F64 : C_MIN_I64; P.col[n++] = (uint8_t)a; break;
        case OP_MAX:
            P.out_types[a] = c.type;
            P.kind[n] = f ? C_MAX_F64 : C_MAX_I64; P.col[n++] = (uint8_t)a; break;
        case OP_FIRST:
            P.out_types[a] = c.type;
            P.kind[n] = C_MIN_ROW; P.col[n++] = (uint8_t)a; break;
        case OP_LAST:
            P.out_types[a] = c.type;
            P.kind[n] = C_MAX_ROW; P.col[n++] = (uint8_t)a; break;
        case OP_VAR: case OP_VAR_POP: case OP_STDDEV: case OP_STDDEV_POP:
            P.out_types[a] = TD_F64;
            P.kind[n] = f ? C_ADD_F64 : C_ADD_I64; P.col[n++] = (uint8_t)a;
            P.kind[n] = C_SQ_F64; P.col[n++] = (uint8_t)a; break;
        default:
            return TD_ERR_NYI;       /* PROD, COUNT_DISTINCT */
        }
    }
    P.ncells = n;
    return TD_OK;
}

/* Output columns of n groups; on failure none is left allocated. */
int alloc_group_outputs(const GroupPlan& P, int64_t n, tdd_group_result_t* out) {
    bool ok = true;
    for (int k = 0; k < P.n_keys; k++) {
        out->keys[k] = td__arena_alloc((size_t)(n > 0 ? n : 1) * P.keys[k].esz);
        ok = ok && out->keys[k];
    }
    for (int a = 0; a < P.n_aggs; a++) {
        int esz = P.out_types[a] == TD_F64 || P.out_types[a] == TD_I64 ? 8 : P.aggs[a].esz;
        out->vals[a] = td__arena_alloc((size_t)(n > 0 ? n : 1) * esz);
        ok = ok && out->vals[a];
        out->out_types[a] = P.out_types[a];
    }
    if (!ok) {
        for (int k = 0; k < P.n_keys; k++) { td__arena_free(out->keys[k]); out->keys[k] = nullptr; }
        for (int a = 0; a < P.n_aggs; a++) { td__arena_free(out->vals[a]); out->vals[a] = nullptr; }
        return TD_ERR_OOM;
    }
    out->n_groups = n;
    return TD_OK;
}

/* ---------------------------------------------------------- row walking -- */
/* Every kernel walks rows in 64-row words (one per wave step) so the
 * selection word can gate a whole wave at once. */

struct RowMask {
    const uint64_t* bits;   /* NULL: all rows */
    int64_t nrows;
};

__device__ __forceinline__ uint64_t word_mask(const RowMask& M, int64_t word) {
    int64_t base = word * 64;
    uint64_t m = M.bits ? M.bits[word] : ~0ull;
    int64_t left = M.nrows - base;
    if (left < 64) m &= (1ull << left) - 1;
    return m;
}

/* ---------------------------------------------------------------- minmax -- */

#define MM_R 8
#define MM_T 256

/* Per-workgroup min / max (no atomics: same-address atomics from thousands
 * of waves serialise at the memory side), folded by k_minmax_final. */
__global__ __launch_bounds__(MM_T) void k_minmax(tdd_col_t col, int vec, RowMask M,
                                                 long long* part) {
    __shared__ long long smn[MM_T / WAVE], smx[MM_T / WAVE];
    const int lane = lane_id();
    long long mn = 0x7fffffffffffffffLL, mx = (long long)0x8000000000000000ull;
    const int64_t n_tiles = (M.nrows + 64 * MM_R - 1) / (64 * MM_R);
    const int64_t wave_id = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t n_waves = (int64_t)gridDim.x * blockDim.x / WAVE;
    for (int64_t t = wave_id; t < n_tiles; t += n_waves) {
        const int64_t base = t * 64 * MM_R + (int64_t)lane * MM_R;
        int64_t v[MM_R];
        load_rows<MM_R>(col, base, M.nrows, vec != 0, v);
        uint32_t m = tile_live<MM_R>(M.bits, base, M.nrows);
#pragma unroll
        for (int j = 0; j < MM_R; j++) {
            if ((m >> j) & 1) {
                mn = v[j] < mn ? v[j] : mn;
                mx = v[j] > mx ? v[j] : mx;
            }
        }
    }
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
        long long a = __shfl_xor(mn, d, WAVE), b = __shfl_xor(mx, d, WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 0) { smn[threadIdx.x / WAVE] = mn; smx[threadIdx.x / WAVE] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < MM_T / WAVE; w++) {
            mn = smn[w] < mn ? smn[w] : mn;
            mx = smx[w] > mx ? smx[w] : mx;
        }
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}

__global__ __launch_bounds__(MM_T) void k_minmax_final(const long long* part, int n,
                                                       long long* out) {
    __shared__ long long smn[MM_T], smx[MM_T];
    long long mn = 0x7fffffffffffffffLL, mx = (long long)0x8000000000000000ull;
    for (int i = threadIdx.x; i < n; i += MM_T) {
        mn = part[2 * i] < mn ? part[2 * i] : mn;
        mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
    }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int d = MM_T / 2; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d) {
            smn[threadIdx.x] = smn[threadIdx.x + d] < smn[threadIdx.x] ? smn[threadIdx.x + d]
                                                                         : smn[threadIdx.x];
            smx[threadIdx.x] = smx[threadIdx.x + d] > smx[threadIdx.x] ? smx[threadIdx.x + d]
                                                                         : smx[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = smn[0]; out[1] = smx[0]; }
}

static RowMask row_mask(const tdd_sel_t* s, int64_t nrows) {
    RowMask M;
    M.bits = s ? s->bits : nullptr;
    M.nrows = nrows;
    return M;
}

/* The partial min / max pairs of one launch plus the folded pair: buf holds
 * 2 + 2 * grid int64 words (folded pair first). */
static int minmax_run(const tdd_col_t* col, int64_t nrows, const tdd_sel_t* mask, int grid,
                      long long* buf, int64_t* mn, int64_t* mx) {
    long long* out = buf;
    long long* part = buf + 2;
    LAUNCH("minmax", k_minmax, grid, MM_T, 0, *col, (int)tile_vec_ok(*col),
           row_mask(mask, nrows), part);
    LAUNCH("minmax", k_minmax_final, 1, MM_T, 0, (const long long*)part, grid, out);
    long long h[2];
    HC(hipMemcpyAsync(h, out, 16, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    *mn = h[0];
    *mx = h[1];
    return TD_OK;
}

static int minmax_grid(int64_t nrows) {
    return stream_grid(ceil_div(nrows > 0 ? nrows : 1, MM_T * MM_R * 2), 4);
}

extern "C" int tdd_minmax(const tdd_col_t* col, int64_t nrows, const tdd_sel_t* mask,
                          int64_t* mn, int64_t* mx) {
    if (col->ranged && !mask && nrows > 0) {          /* known column statistics */
        *mn = col->lo;
        *mx = col->hi;
        return TD_OK;
    }
    const int grid = minmax_grid(nrows);
    SCRATCH(buf, 16 + (size_t)grid * 16);
    return minmax_run(col, nrows, mask, grid, buf.as<long long>(), mn, mx);
}

/* ------------------------------------------------ allocation-free min/max -- */
/* tdd_minmax without any device allocation: its partial pairs live in a
 * file-scope device array, one per GPU, looked up once per device.  Every
 * call on a GPU reuses that GPU's array (calls that overlap on one GPU
 * race); each launch writes every partial it folds, so nothing needs
 * clearing between calls. */
#define MM_NOALLOC_PARTS 2048                 /* partial pairs held: 32 KiB + 16 B */
#define MM_MAX_DEVICES 64

static __device__ __align__(256) long long scratch[2 + 2 * MM_NOALLOC_PARTS];
static long long* g_scratch_dev[MM_MAX_DEVICES];   /* device address of scratch, per GPU */

static int noalloc_scratch(long long** out) {
    int dev = 0;
    HC(hipGetDevice(&dev));
    if (dev < 0 || dev >= MM_MAX_DEVICES) return TD_ERR_RANGE;
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return TD_ERR_IO;
        g_scratch_dev[dev] = (long long*)p;
    }
    *out = g_scratch_dev[dev];
    return TD_OK;
}

extern "C" int tdd_minmax_noalloc(const tdd_col_t* col, int64_t nrows, const tdd_sel_t* mask,
                                  int64_t* mn, int64_t* mx) {
    if (col->ranged && !mask && nrows > 0) {
        *mn = col->lo;
        *mx = col->hi;
        return TD_OK;
    }
    const int grid = minmax_grid(nrows);
    if (grid > MM_NOALLOC_PARTS) {
        fprintf(stderr, "tdd_minmax_noalloc: %d partial pairs needed, the static scratch "
                        "holds %d\n", grid, MM_NOALLOC_PARTS);
        return TD_ERR_RANGE;
    }
    long long* buf = nullptr;
    int rc = noalloc_scratch(&buf);
    if (rc) return rc;
    return minmax_run(col, nrows, mask, grid, buf, mn, mx);
}

/* --------------------------------------------------------- direct array -- */

__device__ __forceinline__ int64_t da_slot(const GroupPlan& P, int64_t row) {
    return da_slot2(P, row, row);
}

/* Identity of every cell: the cell index advances by the grid stride (one
 * division per thread), the identities come from LDS. */
__global__ __launch_bounds__(256) void k_cells_init(uint64_t* tab, int64_t n_slots, GroupPlan P) {
    __shared__ uint64_t id[GP_MAX_CELLS];
    const int nc = P.ncells;
    if (threadIdx.x < GP_MAX_CELLS) id[threadIdx.x] = cell_identity(threadIdx.x < nc ? P.kind[threadIdx.x] : 0);
    __syncthreads();
    const int64_t n = n_slots * nc;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int step = (int)(stride % nc);
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = (int)(i % nc);
    for (; i < n; i += stride) {
        tab[i] = id[c];
        c += step;
        if (c >= nc) c -= nc;
    }
}

/* Each workgroup: private LDS table, then a plain store of the table into
 * its partial slice (merged in workgroup order by k_da_fold). */
__global__ __launch_bounds__(GT) void k_da_lds(GroupPlan P, RowMask M, int64_t n_slots,
                                               uint64_t* partials) {
    extern __shared__ uint64_t tab[];
    int64_t ncell_total = n_slots * P.ncells;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x)
        tab[i] = cell_identity(P.kind[i % P.ncells]);
    __syncthreads();
    int64_t n_words = (M.nrows + 63) >> 6;
    int lane = lane_id();
    int64_t wave = (int64_t)blockIdx.x * (GT / WAVE) + threadIdx.x / WAVE;
    int64_t n_waves = (int64_t)gridDim.x * (GT / WAVE);
    for (int64_t w = wave; w < n_words; w += n_waves) {
        uint64_t m = word_mask(M, w);
        if (m) da_wave(tab, P, w, m, lane, da_slot);
    }
    __syncthreads();
    uint64_t* dst = partials + (int64_t)blockIdx.x * ncell_total;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x) dst[i] = tab[i];
}

/* The same fold for the common shape -- at most two keys and four distinct
 * columns in all, no FIRST / LAST cells -- with R consecutive rows per
 * lane: every column load of a tile is issued before any cell is touched
 * and the cells read their inputs from registers (star.h), where k_da_lds
 * waits on one 64-row word per wave and on each input inside its cell
 * loop.  Keys sit at tile slots 0 (and 1), aggregate inputs after them.
 * A wave whose live rows all fall in one slot (sorted or partitioned keys)
 * folds them in registers and across the wave first, one atomic per cell.
 * Measured on MI355X, 200 M rows: 1,000 random groups of 4 aggregates
 * 2.15 -> 1.27 ms (5.0 TB/s); 30 groups, 1.17 -> 0.83 ms; the 30-day date
 * key of a partitioned table (one slot per wave) 1.47 -> 0.80 ms. */
struct DATile {
    TileCols T;
    uint8_t  cell_slot[GP_MAX_CELLS];
};

template <int R>
__global__ __launch_bounds__(GT) void k_da_tile(GroupPlan P, DATile A, RowMask M, int64_t n_slots,
                                                uint64_t* partials) {
    extern __shared__ uint64_t tab[];
    const int64_t ncell_total = n_slots * P.ncells;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x)
        tab[i] = cell_identity(P.kind[i % P.ncells]);
    __syncthreads();
    const int lane = lane_id();
    const int64_t n_tiles = (M.nrows + 64 * R - 1) / (64 * R);
    const int64_t wave = (int64_t)blockIdx.x * (GT / WAVE) + threadIdx.x / WAVE;
    const int64_t n_waves = (int64_t)gridDim.x * (GT / WAVE);
    const bool two = P.n_keys > 1;
    for (int64_t t = wave; t < n_tiles; t += n_waves) {
        const int64_t base = t * 64 * R + (int64_t)lane * R;
        int64_t cv[TL_NC][R];
        load_tile<R>(A.T, base, M.nrows, cv);
        const uint32_t m = tile_live<R>(M.bits, base, M.nrows);
        int32_t s[R];
        int32_t ls = -1;                               /* the lane's slot, if one */
        bool mixed = false;
#pragma unroll
        for (int j = 0; j < R; j++) {
            int64_t slot = (cv[0][j] - P.kmin[0]) * P.kstride[0];
            if (two) slot += (cv[1][j] - P.kmin[1]) * P.kstride[1];
            s[j] = ((m >> j) & 1) ? (int32_t)slot : -1;
            if (s[j] >= 0) {
                mixed |= ls >= 0 && s[j] != ls;
                ls = s[j];
            }
        }
        /* sorted or partitioned keys: the wave's whole tile in one slot */
        const uint64_t has = __ballot(ls >= 0);
        if (!has) continue;
        const int first = __ffsll((unsigned long long)has) - 1;
        const int32_t s0 = __shfl(ls, first);
        if (__ballot(mixed || (ls >= 0 && ls != s0)) == 0)
            acc_tile_wave<R>(&tab[(int64_t)s0 * P.ncells], P, A.cell_slot, A.T, cv, m, lane == first);
        else
            acc_tile<R>(tab, P, A.cell_slot, A.T, cv, s);
    }
    __syncthreads();
    uint64_t* dst = partials + (int64_t)blockIdx.x * ncell_total;
    for (int64_t i = threadIdx.x; i < ncell_total; i += blockDim.x) dst[i] = tab[i];
}

#define DA_R 4

/* Host: the tile plan of P, or false when the shape is not the common one. */
static bool da_tile_plan(const GroupPlan& P, DATile& A) {
    memset(&A, 0, sizeof(A));
    if (P.n_keys < 1 || P.n_keys > 2) return false;
    for (int k = 0; k < P.n_keys; k++) {
        if (P.keys[k].type == TD_F64 || !P.keys[k].data) return false;
        A.T.col[A.T.n] = P.keys[k];
        if (tile_vec_ok(P.keys[k])) A.T.vec |= 1u << A.T.n;
        A.T.n++;
    }
    for (int i = 1; i < P.ncells; i++) {
        if (P.kind[i] == C_MIN_ROW || P.kind[i] == C_MAX_ROW) return false;
        const tdd_col_t& c = P.aggs[P.col[i]];
        if (!c.data) return false;
        int q = -1;
        for (int k = P.n_keys; k < A.T.n; k++)          /* aggregate slots only */
            if (A.T.col[k].data == c.data && A.T.col[k].type == c.type &&
                A.T.col[k].attrs == c.attrs) q = k;
        if (q < 0) {
            if (A.T.n >= TL_NC) return false;
            A.T.col[A.T.n] = c;
            if (tile_vec_ok(c)) A.T.vec |= 1u << A.T.n;
            q = A.T.n++;
        }
        A.cell_slot[i] = (uint8_t)q;
    }
    return true;
}

__device__ __forceinline__ uint64_t fold_cell(int kind, uint64_t acc, uint64_t s) {
    switch (kind) {
    case C_ADD_F64: case C_SQ_F64:
        return (uint64_t)__double_as_longlong(__longlong_as_double((long long)acc) +
                                              __longlong_as_double((long long)s));
    case C_COUNT: case C_ADD_I64: return acc + s;
    case C_MIN_I64: case C_MIN_F64: case C_MIN_ROW:
        return (int64_t)s < (int64_t)acc ? s : acc;
    default:
        return (int64_t)s > (int64_t)acc ? s : acc;
    }
}

/* Partials are folded in a fixed order, so the result does not depend on
 * which workgroup finished first: chunk c of FOLD_CHUNK consecutive
 * workgroups is folded in workgroup order (one thread per cell, the cells of
 * a partial read contiguously), then the chunks in chunk order. */
#define FOLD_CHUNK 32

__global__ void k_da_fold(GroupPlan P, const uint64_t* partials, int nparts, int64_t stride,
                          int chunk, uint64_t* out) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= stride) return;
    int p0 = (int)blockIdx.y * chunk;
    int p1 = p0 + chunk < nparts ? p0 + chunk : nparts;
    int kind = P.kind[idx % P.ncells];
    uint64_t acc = partials[(int64_t)p0 * stride + idx];
    for (int b = p0 + 1; b < p1; b++) acc = fold_cell(kind, acc, partials[(int64_t)b * stride + idx]);
    out[(int64_t)blockIdx.y * stride + idx] = acc;
}

/* tab = fold of nparts partial tables of n_slots * ncells cells. */
static int fold_partials(const GroupPlan& P, const uint64_t* partials, int nparts,
                         int64_t n_slots, uint64_t* tab) {
    int64_t stride = n_slots * P.ncells;
    int bx = (int)ceil_div(stride, 256);
    int nch = (int)ceil_div(nparts, FOLD_CHUNK);
    if (nch <= 1) {
        LAUNCH("group_da_fold", k_da_fold, dim3(bx, 1), 256, 0, P, partials, nparts, stride,
               nparts, tab);
        return TD_OK;
    }
    SCRATCH(tmp, (size_t)nch * stride * 8);
    LAUNCH("group_da_fold", k_da_fold, dim3(bx, nch), 256, 0, P, partials, nparts, stride,
           FOLD_CHUNK, tmp.as<uint64_t>());
    LAUNCH("group_da_fold", k_da_fold, dim3(bx, 1), 256, 0, P,
           (const uint64_t*)tmp.as<uint64_t>(), nch, stride, nch, tab);
    return TD_OK;
}

__global__ __launch_bounds__(GT) void k_da_global(GroupPlan P, RowMask M, uint64_t* tab) {
    int64_t n_words = (M.nrows + 63) >> 6;
    int lane = lane_id();
    int64_t wave = (int64_t)blockIdx.x * (GT / WAVE) + threadIdx.x / WAVE;
    int64_t n_waves = (int64_t)gridDim.x * (GT / WAVE);
    for (int64_t w = wave; w < n_words; w += n_waves) {
        uint64_t m = word_mask(M, w);
        if (m) da_wave(tab, P, w, m, lane, da_slot);
    }
}

/* Non-empty slot flags for the output compaction. */
__global__ void k_nonempty(const uint64_t* tab, int64_t n_slots, int ncells, uint32_t* flag) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_slots) flag[s] = tab[s * ncells] != 0;
}

/* Direct-array emit: keys decoded from the slot, in ascending slot order. */
/* A block's slots: their cells (row-major) through LDS with consecutive
 * lanes on consecutive words, then each lane emits its slot from there. */
__device__ __forceinline__ const uint64_t* emit_stage(uint64_t* lc, const uint64_t* tab,
                                                      int64_t n_slots, int nc) {
    const int64_t s0 = (int64_t)blockIdx.x * 256;
    const int64_t ns = n_slots - s0 < 256 ? n_slots - s0 : 256;
    for (int64_t w = threadIdx.x; w < ns * nc; w += 256) lc[w] = tab[s0 * nc + w];
    __syncthreads();
    return lc + (int64_t)threadIdx.x * nc;
}

__global__ __launch_bounds__(256) void k_da_emit(GroupPlan P, const uint64_t* tab, int64_t n_slots,
                                                 const int64_t* pos, tdd_group_result_t out) {
    extern __shared__ uint64_t lc[];              /* 256 x ncells */
    const uint64_t* c = emit_stage(lc, tab, n_slots, P.ncells);
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    if (c[0] == 0) return;
    int64_t g = pos[s];
    for (int k = 0; k < P.n_keys; k++) {
        int64_t v = P.kmin[k] + (s / P.kstride[k]) % P.krange[k];
        st_int(out.keys[k], g, P.keys[k].esz, v);
    }
    for (int a = 0; a < P.n_aggs; a++) cells_emit(c, P, a, out.vals[a], g);
}

/* Table -> outputs, shared by both direct-array variants. */
int da_emit(const GroupPlan& P, const uint64_t* tab, int64_t n_slots,
                   tdd_group_result_t* out) {
    SCRATCH(flag, (size_t)n_slots * 4);
    SCRATCH(pos, (size_t)(n_slots + 1) * 8);
    int blocks = (int)ceil_div(n_slots, 256);
    LAUNCH("group_emit", k_nonempty, blocks, 256, 0, tab, n_slots, P.ncells, flag.as<uint32_t>());
    int rc = tdd::scan_u32_i64(flag.as<uint32_t>(), n_slots, pos.as<int64_t>());
    if (rc) return rc;
    int64_t ng = 0;
    HC(hipMemcpyAsync(&ng, pos.as<int64_t>() + n_slots, 8, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    rc = alloc_group_outputs(P, ng, out);
    if (rc) return rc;
    if (ng > 0)
        LAUNCH("group_emit", k_da_emit, blocks, 256, (size_t)256 * P.ncells * 8, P, tab, n_slots,
               (const int64_t*)pos.as<int64_t>(), *out);
    return TD_OK;
}

int64_t g_da_max_slots = (int64_t)1 << 24;

/* Fold per-workgroup LDS partials in workgroup order, then emit. */
int da_finish(const GroupPlan& P, const uint64_t* partials, int nparts, int64_t n_slots,
              tdd_group_result_t* out) {
    SCRATCH(tab, (size_t)n_slots * P.ncells * 8);
    int rc = fold_partials(P, partials, nparts, n_slots, tab.as<uint64_t>());
    if (rc) return rc;
    return da_emit(P, tab.as<uint64_t>(), n_slots, out);
}

/* Identity-initialised global slot table. */
int da_table_init(const GroupPlan& P, uint64_t* tab, int64_t n_slots) {
    int ib = stream_grid(ceil_div(n_slots * P.ncells, 256 * 8));
    LAUNCH("group_init", k_cells_init, ib, 256, 0, tab, n_slots, P);
    return TD_OK;
}

/* Direct-array group-by over rows selected by M with the slot plan in P. */
static int run_da(GroupPlan& P, const RowMask& M, int64_t n_slots, tdd_group_result_t* out) {
    size_t table_bytes = (size_t)n_slots * P.ncells * 8;
    SCRATCH(tab, table_bytes);
    if (table_bytes <= LDS_TABLE_BYTES) {
        int per_cu = (int)(160 * 1024 / (table_bytes + 1024));
        if (per_cu > 4) per_cu = 4;
        if (per_cu < 1) per_cu = 1;
        int grid = stream_grid(ceil_div((M.nrows + 63) / 64, GT / WAVE), per_cu);
        SCRATCH(parts, (size_t)grid * table_bytes);
        DATile A;
        if (da_tile_plan(P, A))
            LAUNCH("group_da_lds", k_da_tile<DA_R>, grid, GT, table_bytes, P, A, M, n_slots,
                   parts.as<uint64_t>());
        else
            LAUNCH("group_da_lds", k_da_lds, grid, GT, table_bytes, P, M, n_slots,
                   parts.as<uint64_t>());
        int rc = fold_partials(P, parts.as<uint64_t>(), grid, n_slots, tab.as<uint64_t>());
        if (rc) return rc;
        out->path = 1;
    } else {
        int ib = stream_grid(ceil_div(n_slots * P.ncells, 256 * 8));
        LAUNCH("group_init", k_cells_init, ib, 256, 0, tab.as<uint64_t>(), n_slots, P);
        int grid = stream_grid(ceil_div((M.nrows + 63) / 64, GT / WAVE), 4);
        LAUNCH("group_da_global", k_da_global, grid, GT, 0, P, M, tab.as<uint64_t>());
        out->path = 2;
    }
    return da_emit(P, tab.as<uint64_t>(), n_slots, out);
}

/* ------------------------------------------------------------------ hash -- */

#define HLL_P 12
#define HLL_M (1 << HLL_P)
#define CACHE_SLOTS 256
#define CACHE_PROBES 4

__global__ __launch_bounds__(256) void k_hll(GroupPlan P, RowMask M, unsigned int* regs) {
    __shared__ unsigned int r[HLL_M];
    for (int i = threadIdx.x; i < HLL_M; i += blockDim.x) r[i] = 0;
    __syncthreads();
    int64_t n_words = (M.nrows + 63) >> 6;
    int lane = lane_id();
    int64_t wave = (int64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
    int64_t n_waves = (int64_t)gridDim.x * (256 / WAVE);
    for (int64_t w = wave; w < n_words; w += n_waves) {
        uint64_t m = word_mask(M, w);
        if ((m >> lane) & 1) {
            uint64_t h = mix64(ht_key(P, w * 64 + lane));
            unsigned idx = (unsigned)(h >> (64 - HLL_P));
            uint64_t rest = (h << HLL_P) | (1ull << (HLL_P - 1));
            unsigned rank = (unsigned)__clzll((long long)rest) + 1;
            atomicMax(&r[idx], rank);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HLL_M; i += blockDim.x)
        if (r[i]) atomicMax(&regs[i], r[i]);
}

/* The same estimate for one raw key column with R consecutive rows per lane
 * (all loads of a tile issued first). */
__device__ __forceinline__ void hll_add(unsigned int* r, uint64_t key) {
    const uint64_t h = mix64(key);
    const unsigned idx = (unsigned)(h >> (64 - HLL_P));
    const uint64_t rest = (h << HLL_P) | (1ull << (HLL_P - 1));
    atomicMax(&r[idx], (unsigned)__clzll((long long)rest) + 1);
}

template <int R>
__global__ __launch_bounds__(256) void k_hll_tile(tdd_col_t key, bool vec, RowMask M,
                                                  unsigned int* regs) {
    __shared__ unsigned int r[HLL_M];
    for (int i = threadIdx.x; i < HLL_M; i += blockDim.x) r[i] = 0;
    __syncthreads();
    const int lane = lane_id();
    const int64_t n_tiles = (M.nrows + 64 * R - 1) / (64 * R);
    const int64_t wave = (int64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
    const int64_t n_waves = (int64_t)gridDim.x * (256 / WAVE);
    for (int64_t t = wave; t < n_tiles; t += n_waves) {
        const int64_t base = t * 64 * R + (int64_t)lane * R;
        int64_t v[R];
        load_rows<R>(key, base, M.nrows, vec, v);
        const uint32_t m = tile_live<R>(M.bits, base, M.nrows);
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((m >> j) & 1) hll_add(r, (uint64_t)v[j]);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HLL_M; i += blockDim.x)
        if (r[i]) atomicMax(&regs[i], r[i]);
}


__global__ __launch_bounds__(GT) void k_ht_agg(GroupPlan P, RowMask M, unsigned long long* keys,
                                               uint64_t mask, uint64_t* cells, int* overflow) {
    extern __shared__ uint64_t sm[];
    unsigned long long* ckey = (unsigned long long*)sm;
    uint64_t* ccell = sm + CACHE_SLOTS;
    const int nc = P.ncells;
    for (int i = threadIdx.x; i < CACHE_SLOTS; i += blockDim.x) ckey[i] = EMPTY_KEY;
    for (int i = threadIdx.x; i < CACHE_SLOTS * nc; i += blockDim.x)
        ccell[i] = cell_identity(P.kind[i % nc]);
    __syncthreads();
    int64_t n_words = (M.nrows + 63) >> 6;
    int lane = lane_id();
    int64_t wave = (int64_t)blockIdx.x * (GT / WAVE) + threadIdx.x / WAVE;
    int64_t n_waves = (int64_t)gridDim.x * (GT / WAVE);
    for (int64_t w = wave; w < n_words; w += n_waves) {
        uint64_t m = word_mask(M, w);
        if (!((m >> lane) & 1)) continue;
        int64_t row = w * 64 + lane;
        uint64_t key = ht_key(P, row);
        int c = -1;
        if (key != EMPTY_KEY) {
            uint32_t ci = (uint32_t)mix64(key) & (CACHE_SLOTS - 1);
            for (int p = 0; p < CACHE_PROBES; p++) {
                unsigned long long k = ckey[ci];
                if (k == key) { c = (int)ci; break; }
                if (k == EMPTY_KEY) {
                    unsigned long long prev = atomicCAS(&ckey[ci], EMPTY_KEY, (unsigned long long)key);
                    if (prev == EMPTY_KEY || prev == key) { c = (int)ci; break; }
                }
                ci = (ci + 1) & (CACHE_SLOTS - 1);
            }
        }
        if (c >= 0) {
            cells_accum(&ccell[c * nc], P, row);
        } else {
            int64_t s = ht_find(keys, mask, key);
            if (s < 0) { *overflow = 1; continue; }
            cells_accum<false>(&cells[s * nc], P, row);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < CACHE_SLOTS; i += blockDim.x) {
        if (ckey[i] == EMPTY_KEY || ccell[i * nc] == 0) continue;
        int64_t s = ht_find(keys, mask, ckey[i]);
        if (s < 0) { *overflow = 1; continue; }
        cells_merge(&cells[s * nc], &ccell[i * nc], P);
    }
}

__device__ __forceinline__ void ht_emit_key(const GroupPlan& P, uint64_t key, int64_t g,
                                            const tdd_group_result_t& out) {
    if (P.n_keys == 1 && P.kstride[0] == 0) {
        st_int(out.keys[0], g, P.keys[0].esz, (int64_t)key);
    } else {
        for (int k = 0; k < P.n_keys; k++) {
            int64_t v = P.kmin[k] + ((int64_t)key / P.kstride[k]) % P.krange[k];
            st_int(out.keys[k], g, P.keys[k].esz, v);
        }
    }
}

/* Groups of the list: entry i is output row i. */
/* The list's groups in list order; a block's cells (row-major, ncells words
 * a group) come in through LDS with consecutive lanes on consecutive words,
 * then each lane emits its group from there. */
__global__ __launch_bounds__(256) void k_ht_emit_list(GroupPlan P, const unsigned long long* keys,
                                                      const uint64_t* cells, int64_t n,
                                                      tdd_group_result_t out) {
    extern __shared__ uint64_t lc[];              /* 256 x ncells */
    const int64_t g0 = (int64_t)blockIdx.x * 256;
    const int64_t ng = n - g0 < 256 ? n - g0 : 256;
    const int nc = P.ncells;
    for (int64_t w = threadIdx.x; w < ng * nc; w += 256) lc[w] = cells[g0 * nc + w];
    __syncthreads();
    const int64_t i = g0 + threadIdx.x;
    if (i >= n) return;
    ht_emit_key(P, keys[i], i, out);
    for (int a = 0; a < P.n_aggs; a++) cells_emit(lc + (int64_t)threadIdx.x * nc, P, a, out.vals[a], i);
}

__global__ __launch_bounds__(256) void k_ht_emit(GroupPlan P, const unsigned long long* keys,
                                                 const uint64_t* cells, int64_t n_slots,
                                                 const int64_t* pos, int64_t off,
                                                 tdd_group_result_t out) {
    extern __shared__ uint64_t lc[];              /* 256 x ncells */
    const uint64_t* c = emit_stage(lc, cells, n_slots, P.ncells);
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    if (c[0] == 0) return;
    int64_t g = off + pos[s];
    uint64_t key = s == n_slots - 1 ? EMPTY_KEY : keys[s];
    ht_emit_key(P, key, g, out);
    for (int a = 0; a < P.n_aggs; a++) cells_emit(c, P, a, out.vals[a], g);
}

static double hll_estimate(const unsigned int* regs) {
    double sum = 0.0;
    int zeros = 0;
    for (int i = 0; i < HLL_M; i++) {
        sum += ldexp(1.0, -(int)regs[i]);
        zeros += regs[i] == 0;
    }
    double m = HLL_M;
    double e = 0.7213 / (1.0 + 1.079 / m) * m * m / sum;
    if (e <= 2.5 * m && zeros) e = m * log(m / zeros);     /* linear counting */
    return e;
}

struct tdd_pht;
struct tdd_hot;
int hot_split(const GroupPlan& P, const uint64_t* mask, int64_t nrows, bool forced,
              tdd_hot** out);
int hot_count(const tdd_hot* h);
const unsigned long long* hot_keys(const tdd_hot* h);
uint64_t* hot_cells(const tdd_hot* h);
int hot_merge(const GroupPlan& P, const tdd_hot* h, unsigned long long* gkeys, uint64_t gmask,
              uint64_t* gcells, int* overflow);
void hot_free(tdd_hot* h);
int pht_partition(const GroupPlan& P, const uint64_t* mask, int64_t nrows, double est,
                  int64_t lds_bytes, bool forced, const unsigned long long* hot, int nhot,
                  uint64_t* hot_cells, tdd_pht** out);
int pht_fold(tdd_pht* h, unsigned long long* gkeys, uint64_t gmask, uint64_t* gcells,
             unsigned long long* lkeys, uint64_t* lcells, unsigned long long* lcount,
             int64_t lcap_out, int* overflow);
void pht_free(tdd_pht* h);
int64_t pht_global_keys(const tdd_pht* h);
int64_t pht_rows(const tdd_pht* h);

/* Hash group-by.  Many rows with many distinct keys are first split by the
 * key's hash until each part fits an LDS table (pht_*); otherwise rows go
 * straight to the global table behind a small LDS cache (k_ht_agg).
 * force 4: the split path with small LDS tables. */
static int run_ht(GroupPlan& P, const RowMask& M, int64_t n_rows_sel, int force,
                  tdd_group_result_t* out) {
    /* force 6: force 4 with a first attempt of 1,024-entry tables (the
     * overflow and regrow paths) */
    const bool tiny_first = force == 6;
    if (force == 6) force = 4;
    SCRATCH(regs, HLL_M * 4);
    int frc = tdd_fill(regs.p, 0, HLL_M * 4);
    if (frc) return frc;
    int hgrid = stream_grid(ceil_div((M.nrows + 63) / 64, 4 * 8), 4);
    if (P.n_keys == 1 && P.kstride[0] == 0 && P.keys[0].data)      /* the raw key: ht_key */
        LAUNCH("group_hll", k_hll_tile<4>, hgrid, 256, 0, P.keys[0], tile_vec_ok(P.keys[0]), M,
               regs.as<unsigned int>());
    else
        LAUNCH("group_hll", k_hll, hgrid, 256, 0, P, M, regs.as<unsigned int>());
    unsigned int h[HLL_M];
    HC(hipMemcpyAsync(h, regs.p, sizeof(h), hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    double est = hll_estimate(h) * 1.1 + 64;
    if (est > (double)n_rows_sel) est = (double)n_rows_sel + 1;
    uint64_t cap = 1024;
    while ((double)cap < 2.0 * est) cap <<= 1;
    tdd_pht* ph = nullptr;
    struct PhFree { tdd_pht** h; ~PhFree() { pht_free(*h); } } phf{&ph};
    tdd_hot* hot = nullptr;
    struct HotFree { tdd_hot** h; ~HotFree() { if (*h) hot_free(*h); } } hf{&hot};
    if (force == 4 || force == 5 ||
        (force == 0 && n_rows_sel >= ((int64_t)1 << 22) && est > 16384)) {
        /* skewed keys: the hot keys' rows are folded first, the partitions
         * carry the cold rows only */
        if (force != 4) {
            int hrc = hot_split(P, M.bits, M.nrows, force == 5, &hot);
            if (hrc && !(hrc == TD_ERR_OOM && force == 0)) return hrc;
        }
        int prc = pht_partition(P, M.bits, M.nrows, est,
                                force == 4 ? 4096 : TUNE_INT("TEIDE_PH_LDS_KB", 48) * 1024,
                                force >= 4, hot ? hot_keys(hot) : nullptr,
                                hot ? hot_count(hot) : 0, hot ? hot_cells(hot) : nullptr, &ph);
        /* no memory for the partitions: the global table alone needs less
         * (the reference's sequential fallback, exec.c:6486-6498) */
        if (prc && prc != TD_ERR_NYI && !(prc == TD_ERR_OOM && force == 0)) return prc;
        if (!ph && hot) { hot_free(hot); hot = nullptr; }   /* every row goes the other way */
    }
    /* partitioned: most groups leave through the list (cap entries); the
     * global table takes the split parts' keys, so it starts at their
     * bound and grows to cap if that overflows */
    uint64_t gcap = cap;
    if (ph) {
        const int64_t want = pht_global_keys(ph) + (hot ? hot_count(hot) : 0);
        uint64_t g = 1024;
        while ((double)g < 2.0 * (double)want) g <<= 1;
        if (g < gcap) gcap = g;
    }
    if (tiny_first) cap = gcap = 1024;
    int64_t ol[2] = {0, 0};
    for (int attempt = 0; attempt < 7; attempt++) {
        POLL_CANCEL();
        if (attempt > 0) {                                   /* the last try overflowed */
            if ((int32_t)ol[0]) {                            /* the global table: grow it */
                gcap = gcap < cap ? cap : gcap << 1;
                if (!ph && cap < gcap) cap = gcap;
            }
            if ((int32_t)(ol[0] >> 32)) {                    /* the list: room for every entry */
                while ((double)cap < 1.5 * (double)ol[1]) cap <<= 1;
            }
        }
        int64_t n_slots = (int64_t)gcap + 1;                 /* + the EMPTY_KEY slot */
        SCRATCH(keys, (size_t)gcap * 8);
        SCRATCH(cells, (size_t)n_slots * P.ncells * 8);
        SCRATCH(ovf, 16);
        int rc = tdd_fill(keys.p, EMPTY_KEY, (size_t)gcap * 8);
        if (rc) return rc;
        int ib = stream_grid(ceil_div(n_slots * P.ncells, 256 * 8));
        LAUNCH("group_init", k_cells_init, ib, 256, 0, cells.as<uint64_t>(), n_slots, P);
        rc = tdd_fill(ovf.p, 0, 16);          /* table flag, list flag, list count */
        if (rc) return rc;
        Scratch lkeys(ph ? (size_t)cap * 8 : 8), lcells(ph ? (size_t)cap * P.ncells * 8 : 8);
        if (!lkeys.ok() || !lcells.ok()) return TD_ERR_OOM;
        unsigned long long* lcount = (unsigned long long*)(ovf.as<uint8_t>() + 8);
        if (hot) {
            rc = hot_merge(P, hot, keys.as<unsigned long long>(), (uint64_t)(gcap - 1),
                           cells.as<uint64_t>(), ovf.as<int>());
            if (rc) return rc;
        }
        if (ph) {
            rc = pht_fold(ph, keys.as<unsigned long long>(), (uint64_t)(gcap - 1),
                          cells.as<uint64_t>(), lkeys.as<unsigned long long>(),
                          lcells.as<uint64_t>(), lcount, (int64_t)cap, ovf.as<int>());
            if (rc) return rc;
        } else {
            size_t lds = (size_t)CACHE_SLOTS * 8 * (1 + P.ncells);
            int grid = stream_grid(ceil_div((M.nrows + 63) / 64, GT / WAVE), 2);
            LAUNCH("group_ht", k_ht_agg, grid, GT, lds, P, M, keys.as<unsigned long long>(),
                   (uint64_t)(gcap - 1), cells.as<uint64_t>(), ovf.as<int>());
        }
        /* the table's occupied slots counted before the flags come back, so
         * both come back with one wait (an overflowed try discards them) */
        SCRATCH(flag, (size_t)n_slots * 4);
        SCRATCH(pos, (size_t)(n_slots + 1) * 8);
        int blocks = (int)ceil_div(n_slots, 256);
        LAUNCH("group_emit", k_nonempty, blocks, 256, 0, (const uint64_t*)cells.as<uint64_t>(),
               n_slots, P.ncells, flag.as<uint32_t>());
        rc = tdd::scan_u32_i64(flag.as<uint32_t>(), n_slots, pos.as<int64_t>());
        if (rc) return rc;
        int64_t nh = 0;
        HC(hipMemcpyAsync(ol, ovf.p, 16, hipMemcpyDeviceToHost, tdd::g_stream));
        HC(hipMemcpyAsync(&nh, pos.as<int64_t>() + n_slots, 8, hipMemcpyDeviceToHost, tdd::g_stream));
        HC(hipStreamSynchronize(tdd::g_stream));
        if (ol[0]) continue;                                 /* a table too small: retry */
        const int64_t n_list = ph ? ol[1] : 0;
        const int64_t ng = n_list + nh;
        rc = alloc_group_outputs(P, ng, out);
        if (rc) return rc;
        if (n_list > 0)
            LAUNCH("group_emit", k_ht_emit_list, (int)ceil_div(n_list, 256), 256,
                   (size_t)256 * P.ncells * 8, P,
                   (const unsigned long long*)lkeys.as<unsigned long long>(),
                   (const uint64_t*)lcells.as<uint64_t>(), n_list, *out);
        if (nh > 0)
            LAUNCH("group_emit", k_ht_emit, blocks, 256, (size_t)256 * P.ncells * 8, P,
                   (const unsigned long long*)keys.as<unsigned long long>(),
                   (const uint64_t*)cells.as<uint64_t>(), n_slots,
                   (const int64_t*)pos.as<int64_t>(), n_list, *out);
        out->path = ph ? (hot ? 6 : 5) : 3;
        out->tries = attempt + 1;
        if (ph) out->part_rows = pht_rows(ph) >= 0 ? pht_rows(ph) : n_rows_sel;
        return TD_OK;
    }
    return TD_ERR_OOM;
}

int run_pda(const GroupPlan& P, const uint64_t* mask, int64_t nrows, int64_t n_slots,
            int64_t leaf_bytes, tdd_group_result_t* out);
int group_tuples(const tdd_group_spec_t* spec, int64_t nsel, tdd_group_result_t* out);

/* --------------------------------------------------------------- driver -- */

static int64_t masked_rows(const tdd_sel_t* s, int64_t nrows) {
    if (!s) return nrows;
    int64_t t = 0;
    if (tdd_sel_total(s, &t) != TD_OK) return nrows;
    return t;
}

extern "C" int tdd_group(const tdd_group_spec_t* spec, tdd_group_result_t* out) {
    memset(out, 0, sizeof(*out));
    if (spec->n_keys < 0 || spec->n_keys > TDD_MAX_KEYS || spec->n_aggs > TDD_MAX_AGGS ||
        spec->n_aggs < 0)
        return TD_ERR_NYI;
    GroupPlan P;
    memset(&P, 0, sizeof(P));
    P.n_keys = spec->n_keys;
    P.n_aggs = spec->n_aggs;
    for (int k = 0; k < P.n_keys; k++) P.keys[k] = spec->keys[k];
    for (int a = 0; a < P.n_aggs; a++) { P.aggs[a] = spec->aggs[a]; P.ops[a] = spec->ops[a]; }
    int rc = plan_cells(P);
    if (rc) return rc;
    RowMask M = row_mask(spec->mask, spec->nrows);
    int64_t nsel = masked_rows(spec->mask, spec->nrows);
    if (P.n_keys == 0) {
        /* no keys: one group of every selected row (none when no row is
         * selected), the reference's scalar aggregate (exec.c:5766-5940) */
        rc = run_da(P, M, 1, out);
        if (rc) return rc;
        HC(hipStreamSynchronize(tdd::g_stream));
        return TD_OK;
    }

    /* Integer keys: masked ranges decide between direct array and hash. */
    bool all_int = true;
    for (int k = 0; k < P.n_keys; k++) all_int = all_int && is_int_type(P.keys[k].type);
    bool da = false, composite = false;
    if (all_int && spec->force_path != 2) {
        double slots = 1.0;
        composite = true;
        for (int k = 0; k < P.n_keys; k++) {
            int64_t mn, mx;
            rc = tdd_minmax(&P.keys[k], spec->nrows, spec->mask, &mn, &mx);
            if (rc) return rc;
            if (mn > mx) { mn = 0; mx = 0; }               /* no selected rows */
            P.kmin[k] = mn;
            P.krange[k] = (int64_t)((uint64_t)mx - (uint64_t)mn + 1);
            if (P.krange[k] <= 0) { composite = false; slots = 1e30; break; }
            slots *= (double)P.krange[k];
        }
        if (slots > 9.0e18) composite = false;
        if (composite) {
            int64_t st = 1;
            for (int k = P.n_keys - 1; k >= 0; k--) { P.kstride[k] = st; st *= P.krange[k]; }
        }
        /* many rows into a table beyond LDS: radix-partitioned direct array
         * (force 3 takes it at any size, with small leaf tables) */
        const bool pda_ok = composite && slots <= (double)((int64_t)1 << 26) &&
                            slots * P.ncells * 8 <= 8.0e9 && slots <= 4.0 * (double)(nsel + 1024);
        if (pda_ok && spec->force_path < 4 && (spec->force_path == 3 ||
                       (spec->force_path == 0 && nsel >= ((int64_t)1 << 22) &&
                        slots * P.ncells * 8 > LDS_TABLE_BYTES))) {
            rc = run_pda(P, M.bits, M.nrows, (int64_t)slots,
                         spec->force_path == 3 ? 64 : 96 * 1024, out);
            const bool degrade = rc == TD_ERR_OOM && spec->force_path == 0;   /* global DA / hash */
            if (rc != TD_ERR_NYI && !degrade) {
                if (rc) return rc;
                HC(hipStreamSynchronize(tdd::g_stream));
                return TD_OK;
            }
            memset(out, 0, sizeof(*out));
        }
        da = composite && (slots <= (double)g_da_max_slots || spec->force_path == 1) &&
             slots * P.ncells * 8 <= 8.0e9;
        if (da && spec->force_path != 1 && slots > 4.0 * (double)(nsel + 1024)) da = false;
        if (spec->force_path >= 4) da = false;
        if (da) {
            rc = run_da(P, M, (int64_t)slots, out);
            if (rc) return rc;
            HC(hipStreamSynchronize(tdd::g_stream));
            return TD_OK;
        }
    }
    if (P.n_keys > 1 && !composite) return group_tuples(spec, nsel, out);
    if (P.n_keys == 1 && !composite) P.kstride[0] = 0;      /* raw 64-bit key */
    if (P.n_keys == 1 && composite && P.keys[0].type != TD_F64) {
        P.kstride[0] = 0;                                    /* integer key: use its value */
    }
    rc = run_ht(P, M, nsel, spec->force_path, out);
    if (rc) return rc;
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}
