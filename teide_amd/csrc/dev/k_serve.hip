/*
 * k_serve.hip -- shared-scan kernels of the request server.
 *
 * Served family:
 *   GROUP(dim.key; SUM / AVG / COUNT of fact columns)
 *     over JOIN(FILTER(fact, fact.pcol CMP t), fact.fk = dim.pk)
 * with a per-request constant t, comparison and group key.  The filtered
 * column has a small integer domain [qlo, qlo + Q) (the resident column's
 * statistics), so a request's predicate is a range of that axis.  A step
 * folds one chunk of fact rows into a cube per group key -- cell (v, g) holds
 * the row count and one sum per value column of the rows with
 * pcol = qlo + v joining a dim row of group g -- and turns it into prefix and
 * suffix sums over v.  Each active request then adds, per group, the sums of
 * its range (one cell) into its own direct-array table, whose layout is the
 * group-by's (group_plan.h), so the emit is the group-by's emit.
 *
 *   k_srv_map / k_srv_jmap   dim: direct map pk -> group of each key, and
 *                  the joint map (every key's group per pk) (once, at open)
 *   k_srv_ccount / k_srv_cput / k_srv_cfold   fact chunk -> cubes by slices:
 *                  cells counted, records placed in slice order, each slice
 *                  folded in LDS and written with its prefix / suffix sums
 *   k_srv_cube + k_srv_cumul   the same with global atomics per row and key,
 *                  then the sums over v (cubes of more than 8,192 slices)
 *   k_srv_accum    per (request, g): table += the request's cube cell
 *   k_srv_flags / k_srv_emit   finished requests' groups, ascending keys
 *
 * Per row of a step the scan reads fk, pcol and the value columns once
 * (3 x 8 bytes for the fact layout fk i64, qty i64, price f64) for every
 * request in the mix, so the HBM cost of the scan is shared.
 *
 * Requests whose filter is anything else (a comparison on an f64 or
 * wide-domain column, several comparisons) are direct requests: they share
 * the same chunk with their own predicate program.
 *
 *   k_srv_bhist / k_srv_bscatter   once per step and group key: the
 *                  chunk's rows with a dim row, counting-sorted by group
 *                  slice (SG consecutive groups) into records holding the
 *                  group and every column the step's direct requests read
 *   k_srv_bfold    (slice, request group): each workgroup reads one slice's
 *                  records once for up to 16 requests, evaluates each
 *                  request's predicate on them and folds the passing records
 *                  into LDS slices of the requests' tables, then adds the
 *                  slices into the tables (one workgroup per slice and
 *                  request: no atomics on the tables).  A group of up to
 *                  77 requests `col CMP t` (< <= > >=) on one column with
 *                  one cell layout folds each record once, into the bucket
 *                  of the thresholds below its value; running sums over the
 *                  buckets give each request the sum of its buckets
 *   k_srv_direct   requests with more groups than 4,096 slices cover: every
 *                  row of the chunk folded with global atomics
 */
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "dev_common.h"
#include "group_plan.h"
#include "pred_eval.h"
#include "tile.h"

#define SRV_MAX_Q     4096                    /* pcol domain */
#define SRV_MAX_CELLS ((int64_t)1 << 26)      /* Q x G per key */
#define SRV_R         4                       /* rows per lane of a scan tile */
#define SRV_MAX_Y     65535                   /* grid.y limit */
#define SRV_DIR_LDS   (64 * 1024)             /* LDS table slices of one request group */
#define SRV_DIR_GRP   16                      /* direct requests per group */
#define SRV_SLICE_B   4096                    /* LDS bytes of one request's table slice */
#define SRV_MAX_S     4096                    /* group slices per key */
#define SRV_REC_C     8                       /* record columns */
#define SRV_TILE      4096                    /* rows per bucket-sort tile */

enum { M_NONE = 0, M_RAW = 1, M_SFX = 2, M_PFX = 3 };

struct SrvCube {                  /* arrays [v][g], v-major */
    uint64_t* cnt[4];             /* by mode (index 0 unused) */
    int64_t*  sq[4];              /* sum of pcol (sfx / pfx) */
    uint64_t* val[4][TDD_SRV_MAX_VALS];
};

struct SrvKeyArgs {
    const int32_t* map;           /* (pk - kmin) -> g, -1: no dim row */
    uint64_t* cnt;                /* raw cube */
    uint64_t* val[TDD_SRV_MAX_VALS];
    int64_t   G;
    int32_t   on, vmask;
};

struct SrvScanArgs {
    tdd_col_t fk, pcol;
    tdd_col_t vals[TDD_SRV_MAX_VALS];
    int64_t   r0, r1, kmin, range, qlo;
    int32_t   Q, nk;
    uint32_t  vec;                /* bit 0 fk, 1 pcol, 2+j vals[j]: 16-byte loads */
    int32_t   pad;
    SrvKeyArgs key[TDD_SRV_MAX_KEYS];
};

struct SrvAccArgs {               /* one active request of a step */
    uint64_t* tab;                /* [G][ncells] */
    int64_t   G, idx;
    int32_t   key, mode, ncells, pad;
    uint8_t   kind[GP_MAX_CELLS];
    int8_t    src[GP_MAX_CELLS];  /* ADD cells: value column, -2 pcol */
};

struct tdd_srv {
    tdd_srv_spec_t spec;
    int64_t  kmin, range, qlo;
    int32_t  Q, nchunks;
    int64_t  chunk;
    int64_t  step;
    int64_t  bytes;
    int32_t* map[TDD_SRV_MAX_KEYS];
    uint8_t* cube_mem[TDD_SRV_MAX_KEYS];
    SrvCube  cube[TDD_SRV_MAX_KEYS];
    int64_t  G[TDD_SRV_MAX_KEYS], klo[TDD_SRV_MAX_KEYS];
    int32_t  ok[TDD_SRV_MAX_KEYS];
    SrvCube* d_cubes;             /* device copy of cube[] */
    /* the cube by slices: one block holding cells, records, counts, scan */
    uint8_t*  cs_mem;             /* nullptr: k_srv_cube only */
    uint32_t* cs_cells;           /* [key slot][chunk row] */
    uint16_t* cs_roff;            /* [record] */
    uint64_t* cs_rval;            /* [value][record] */
    uint32_t* cs_C;
    int64_t*  cs_O;
    int32_t*  cs_jmap;            /* joint map: [pk - kmin][cs_js], slot order of the served keys */
    int64_t   cs_rows;            /* chunk rows the arrays hold per key slot */
    int32_t   cs_nk, cs_js;       /* key slots, ints per joint map row */
};

struct tdd_srv_acc {
    GroupPlan     P;
    tdd_srv_req_t req;
    SrvAccArgs    a;
    int32_t       direct;          /* 1: own predicate, k_srv_direct */
    tdd_pred_t    pred;
};

struct SrvDirReq {                 /* one direct request of a step */
    tdd_pred_t P;
    uint64_t*  tab;                /* [G][ncells] */
    int64_t    G;
    int32_t    key, ncells;
    int32_t    lds;                /* word offset of its LDS table, -1: the global table */
    int16_t    blo, bhi;           /* threshold group: the buckets it sums */
    uint8_t    kind[GP_MAX_CELLS];
    int8_t     src[GP_MAX_CELLS];  /* ADD cells: value column, -2 pcol */
};

/* A group of direct requests folded by one workgroup.  tmode 1: every
 * request is `col CMP cut` on one record column with one cell layout; a
 * record goes to bucket b = the cuts below it (strict: cut < x, else
 * cut <= x) and folds once into that bucket's table; at the end the
 * tables are summed over the buckets from the top (up: > and >=) or from
 * the bottom (< and <=) in place, and request j takes bucket blo (up) or
 * bhi: the sum of its buckets [blo, bhi]. */
#define SRV_TCUTS 127
#define SRV_DIR_TLDS (78 * 1024)       /* LDS of a threshold group: two workgroups a CU */
struct SrvDirGroup {
    int32_t first, n, words, tmode;
    int32_t m, strict, as_f64, up;     /* up: every request passes above its cut */
    int64_t cut[SRV_TCUTS];        /* ascending, distinct (f64: the bits) */
};

/* Request descriptors are read-only during a launch and uniform across a
 * workgroup: read through the constant address space they are scalar loads
 * (through a plain global pointer every byte of a predicate program would
 * be a vector load per lane). */
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) SrvDirReq ConstReq;
#else
typedef const SrvDirReq ConstReq;        /* the host pass only parses the kernels */
#endif

struct SrvBucket {                 /* one key's bucket sort of a chunk */
    tdd_col_t      fk;
    const int32_t* map;
    int64_t        r0, r1, kmin, range, ntiles;
    int32_t        sg, S;          /* groups per slice, slices */
    int32_t        nc, pad;
    tdd_col_t      src[SRV_REC_C]; /* fact columns carried into the records */
    uint32_t*      hist;           /* [S][ntiles] */
    const int64_t* off;            /* exclusive scan of hist, + total */
    uint32_t*      rg;             /* record group */
    int32_t*       rowg;           /* per chunk row: its group (-1: none), k_srv_bhist -> bscatter */
    void*          rc[SRV_REC_C];  /* record columns (the source's element size) */
};

struct SrvFoldArgs {
    const uint32_t*    rg;
    const int64_t*     off;
    int64_t            ntiles, G;
    int32_t            sg, ngroups;
    int32_t            parts, pad;     /* workgroups per (slice, request group) */
    const SrvDirReq*   reqs;       /* P over record columns, src = record column */
    const SrvDirGroup* groups;
};

struct SrvDirArgs {
    tdd_col_t          fk, pcol;   /* pcol.data NULL: no request sums it */
    tdd_col_t          vals[TDD_SRV_MAX_VALS];   /* data NULL: unused */
    const int32_t*     map[TDD_SRV_MAX_KEYS];
    int64_t            r0, r1, kmin, range, span;
    int32_t            ngroups, pad;
    const SrvDirReq*   reqs;
    const SrvDirGroup* groups;
};

/* --------------------------------------------------------------- kernels -- */

/* map[pk - kmin] = key - klo for every dim row; a second row with the same
 * pk raises *dup. */
__global__ void k_srv_map(tdd_col_t pk, tdd_col_t key, int64_t nr, int64_t kmin, int64_t klo,
                          int32_t* map, uint32_t* dup) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nr;
         r += (int64_t)gridDim.x * blockDim.x) {
        int32_t g = (int32_t)(ld_i64(key, r) - klo);
        int32_t old = atomicCAS(&map[ld_i64(pk, r) - kmin], -1, g);
        if (old != -1) *dup = 1u;
    }
}

__global__ __launch_bounds__(256) void k_srv_cube(SrvScanArgs S) {
    const int64_t rows = S.r1 - S.r0;
    const int64_t n_tiles = (rows + 64 * SRV_R - 1) / (64 * SRV_R);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t n_waves = (int64_t)gridDim.x * blockDim.x / WAVE;
    const int lane = lane_id();
    for (int64_t t = wave; t < n_tiles; t += n_waves) {
        const int64_t base = S.r0 + t * 64 * SRV_R + (int64_t)lane * SRV_R;
        int64_t fk[SRV_R], q[SRV_R], v[TDD_SRV_MAX_VALS][SRV_R];
        load_rows<SRV_R>(S.fk, base, S.r1, S.vec & 1u, fk);
        load_rows<SRV_R>(S.pcol, base, S.r1, (S.vec >> 1) & 1u, q);
#pragma unroll
        for (int j = 0; j < TDD_SRV_MAX_VALS; j++) {
            if (S.vals[j].data) {
                load_rows<SRV_R>(S.vals[j], base, S.r1, (S.vec >> (2 + j)) & 1u, v[j]);
            } else {
#pragma unroll
                for (int i = 0; i < SRV_R; i++) v[j][i] = 0;
            }
        }
        int64_t slot[SRV_R];
#pragma unroll
        for (int i = 0; i < SRV_R; i++) {
            const uint64_t k = (uint64_t)(fk[i] - S.kmin);
            const uint64_t c = (uint64_t)(q[i] - S.qlo);
            const bool live = base + i < S.r1 && k < (uint64_t)S.range && c < (uint64_t)S.Q;
            slot[i] = live ? (int64_t)k : -1;
            q[i] = (int64_t)c;
        }
        for (int kk = 0; kk < S.nk; kk++) {
            const SrvKeyArgs& K = S.key[kk];
            if (!K.on) continue;
            int32_t g[SRV_R];
#pragma unroll
            for (int i = 0; i < SRV_R; i++) g[i] = slot[i] >= 0 ? K.map[slot[i]] : -1;
#pragma unroll
            for (int i = 0; i < SRV_R; i++) {
                if (g[i] < 0) continue;
                const int64_t cell = q[i] * K.G + g[i];
                atomicAdd((unsigned long long*)&K.cnt[cell], 1ull);
#pragma unroll
                for (int j = 0; j < TDD_SRV_MAX_VALS; j++) {
                    if (!((K.vmask >> j) & 1)) continue;
                    if (S.vals[j].type == TD_F64)
                        atomicAdd((double*)&K.val[j][cell], __longlong_as_double(v[j][i]));
                    else
                        atomicAdd((unsigned long long*)&K.val[j][cell], (unsigned long long)v[j][i]);
                }
            }
        }
    }
}

/* The cube by slices: instead of two global atomics per row and key (they
 * bound k_srv_cube at ~23 G/s), a step counts the chunk's rows per cube
 * slice, places them as records in slice order and folds each slice in
 * LDS, writing every cell of it -- no atomics outside LDS, no fills. A
 * slice is gw consecutive groups x every pcol value, so the fold also
 * writes the prefix / suffix sums over the values (what k_srv_cumul does
 * after k_srv_cube) and the raw cube only when a request reads it.
 *
 *   k_srv_ccount    rows of workgroup b's span: one gather of the joint map
 *                   row (every served key's group of a dim row), each active
 *                   key's cell (g * Q + v) kept per row, counts per (slice, b)
 *   scan            counts -> record positions
 *   k_srv_cput      the same rows again: records (cell within its slice,
 *                   the key's value words) at the positions, each tile's
 *                   records counting-sorted by slice in LDS and written
 *                   slice run by slice run
 *   k_srv_cfold     one workgroup per slice: its records (one segment per
 *                   row-pass workgroup) folded into an LDS table of the
 *                   slice's cells, then per group the sums over v out
 *
 * Record positions: R consecutive slices form a run group and the scan
 * runs over (run group, XCD of b, b, slice in the group), so workgroup b
 * writes one run per run group (~256 per key) and neighbouring runs come
 * from workgroups of one XCD. */
#define SRV_CS_MAXS  8192                     /* slices of a step, all keys */
#define SRV_CS_WGMAX 512                      /* workgroups of the two row passes */
#define SRV_CS_TS    256                      /* target run groups per key */
#define SRV_CS_NONE  0xffffffffu

struct SrvCubePlan {
    int32_t   nk, ns, nwg, js;                /* js: ints per joint map row (1, 2, 4) */
    int64_t   span;                           /* rows per workgroup (a multiple of 4,096) */
    const int32_t* jmap;                      /* [pk - kmin][js] */
    int64_t   qlo;
    int32_t   Q;
    uint32_t  ftype;                          /* bit j: value column j is f64 */
    int32_t   on[TDD_SRV_MAX_KEYS], vmask[TDD_SRV_MAX_KEYS], kslot[TDD_SRV_MAX_KEYS];
    int32_t   modes[TDD_SRV_MAX_KEYS];        /* bit M_RAW / M_SFX / M_PFX: cubes to write */
    int32_t   R[TDD_SRV_MAX_KEYS];            /* slices per run group */
    int32_t   gw[TDD_SRV_MAX_KEYS];           /* groups per slice */
    int32_t   sbase[TDD_SRV_MAX_KEYS + 1];    /* first slice of key k (run groups whole) */
    int64_t   G[TDD_SRV_MAX_KEYS];
    uint32_t* cells[TDD_SRV_MAX_KEYS];        /* per row of the chunk: cell or NONE */
    uint32_t* C;                              /* counts, srv_cs_at order */
    const int64_t* O;                         /* their exclusive scan */
    uint16_t* roff;                           /* records: cell within the slice */
    uint64_t* rval[TDD_SRV_MAX_VALS];         /*          value words */
    SrvCube   cube[TDD_SRV_MAX_KEYS];
};

__device__ __forceinline__ int srv_cs_key(const SrvCubePlan& C, int s) {
    int kk = 0;
    while (kk + 1 < C.nk && s >= C.sbase[kk + 1]) kk++;
    return kk;
}

/* Scan position of the count (slice s of key kk, row-pass workgroup b):
 * key, run group, b's XCD (b mod 8), b's order on it, slice in the group.
 * The grid is a multiple of 8. */
__device__ __forceinline__ int64_t srv_cs_at(const SrvCubePlan& C, int kk, int s, int64_t b) {
    const int64_t l = s - C.sbase[kk], R = C.R[kk];
    return (int64_t)C.sbase[kk] * C.nwg +
           (((l / R) * 8 + (b & 7)) * (C.nwg >> 3) + (b >> 3)) * R + l % R;
}

/* Dim slot -> the joint map row (ints beyond the served keys hold -1). */
__global__ void k_srv_jmap(const int32_t* m0, const int32_t* m1, const int32_t* m2,
                           const int32_t* m3, int64_t range, int js, int32_t* jmap) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < range;
         r += (int64_t)gridDim.x * blockDim.x) {
        jmap[r * js] = m0[r];
        if (js > 1) jmap[r * js + 1] = m1 ? m1[r] : -1;
        if (js > 2) {
            jmap[r * js + 2] = m2 ? m2[r] : -1;
            jmap[r * js + 3] = m3 ? m3[r] : -1;
        }
    }
}

__device__ __forceinline__ int32_t srv_pick(int4 v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

__global__ __launch_bounds__(1024) void k_srv_ccount(SrvScanArgs S, SrvCubePlan C) {
    extern __shared__ uint32_t lh[];          /* per slice: the workgroup's counts */
    const int64_t b = blockIdx.x;
    const int64_t s0 = S.r0 + b * C.span;
    const int64_t s1 = s0 + C.span < S.r1 ? s0 + C.span : S.r1;
    for (int i = threadIdx.x; i < C.ns; i += 1024) lh[i] = 0u;
    __syncthreads();
    for (int64_t b0 = s0; b0 < s1; b0 += 1024 * SRV_R) {
        const int64_t base = b0 + (int64_t)threadIdx.x * SRV_R;
        {
            int64_t fk[SRV_R], q[SRV_R];
            load_rows<SRV_R>(S.fk, base, s1, S.vec & 1u, fk);
            load_rows<SRV_R>(S.pcol, base, s1, (S.vec >> 1) & 1u, q);
            int4 jr[SRV_R];
#pragma unroll
            for (int i = 0; i < SRV_R; i++) {
                const uint64_t k = (uint64_t)(fk[i] - S.kmin);
                const uint64_t c = (uint64_t)(q[i] - S.qlo);
                const bool live = base + i < s1 && k < (uint64_t)S.range && c < (uint64_t)S.Q;
                q[i] = live ? (int64_t)c : -1;
                jr[i] = make_int4(-1, -1, -1, -1);
                if (live) {
                    if (C.js == 4) {
                        jr[i] = ((const int4*)C.jmap)[k];
                    } else if (C.js == 2) {
                        const int2 t = ((const int2*)C.jmap)[k];
                        jr[i].x = t.x;
                        jr[i].y = t.y;
                    } else {
                        jr[i].x = C.jmap[k];
                    }
                }
            }
            for (int kk = 0; kk < C.nk; kk++) {
                if (!C.on[kk]) continue;
                const int ks = C.kslot[kk];
                const int64_t sl = (int64_t)C.gw[kk] * C.Q;
#pragma unroll
                for (int i = 0; i < SRV_R; i++) {
                    if (base + i >= s1) continue;
                    const int32_t g = srv_pick(jr[i], ks);
                    uint32_t cell = SRV_CS_NONE;
                    if (q[i] >= 0 && g >= 0) {
                        cell = (uint32_t)((int64_t)g * C.Q + q[i]);
                        atomicAdd(&lh[C.sbase[kk] + (int64_t)cell / sl], 1u);
                    }
                    C.cells[kk][base + i - S.r0] = cell;
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < C.ns; i += 1024) C.C[srv_cs_at(C, srv_cs_key(C, i), i, b)] = lh[i];
}

/* Exclusive scan of a[0, n) in place (n <= 8 * 1024), a[n] = the total;
 * every thread of a 1,024-thread block calls it. */
__device__ __forceinline__ void srv_block_scan(uint32_t* a, int n, uint32_t* wsum) {
    const int per = (n + 1023) / 1024;
    const int i0 = threadIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
    uint32_t t = 0;
    for (int i = i0; i < i1; i++) t += a[i];
    const int lane = lane_id(), w = threadIdx.x / WAVE;
    uint32_t inc = t;
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
    }
    if (lane == WAVE - 1) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < w; k++) before += wsum[k];
    uint32_t run = before + inc - t;
    for (int i = i0; i < i1; i++) {
        const uint32_t c = a[i];
        a[i] = run;
        run += c;
    }
    if (threadIdx.x == 1023) a[n] = run;                  /* the last thread ends at the total */
    __syncthreads();
}

/* The records of the sliced cube, placed through LDS: per tile of 4,096
 * rows and per active key, the tile's records are counting-sorted by slice
 * in LDS and written slice by slice, so the records of one slice from one
 * tile go out as one run by consecutive lanes (for a cube of ~256 slices,
 * ~16 records a run) instead of one store per record. */
__global__ __launch_bounds__(1024) void k_srv_cput(SrvScanArgs S, SrvCubePlan C, int nskmax) {
    extern __shared__ uint32_t lds[];
    uint32_t* lh = lds;                                   /* ns: the workgroup's cursors */
    uint32_t* tof = lh + C.ns;                            /* nsk + 1: the tile's offsets */
    uint16_t* spos = (uint16_t*)(tof + nskmax + 1);       /* 4,096: slice of sorted record */
    uint16_t* sroff = spos + 4096;                        /* 4,096: its cell in the slice */
    uint64_t* sval = (uint64_t*)(((uintptr_t)(sroff + 4096) + 7) & ~(uintptr_t)7);
    __shared__ uint32_t wsum[1024 / WAVE];
    const int64_t b = blockIdx.x;
    const int64_t s0 = S.r0 + b * C.span;
    const int64_t s1 = s0 + C.span < S.r1 ? s0 + C.span : S.r1;
    for (int i = threadIdx.x; i < C.ns; i += 1024)
        lh[i] = (uint32_t)C.O[srv_cs_at(C, srv_cs_key(C, i), i, b)];
    __syncthreads();
    for (int64_t b0 = s0; b0 < s1; b0 += 1024 * SRV_R) {
        const int64_t base = b0 + (int64_t)threadIdx.x * SRV_R;
        int64_t v[TDD_SRV_MAX_VALS][SRV_R];
#pragma unroll
        for (int j = 0; j < TDD_SRV_MAX_VALS; j++) {
            if (S.vals[j].data) {
                load_rows<SRV_R>(S.vals[j], base, s1, (S.vec >> (2 + j)) & 1u, v[j]);
            } else {
#pragma unroll
                for (int i = 0; i < SRV_R; i++) v[j][i] = 0;
            }
        }
        for (int kk = 0; kk < C.nk; kk++) {
            if (!C.on[kk]) continue;
            const int sb = C.sbase[kk], nsk = C.sbase[kk + 1] - sb;
            const int64_t sl = (int64_t)C.gw[kk] * C.Q;
            for (int i = threadIdx.x; i <= nsk; i += 1024) tof[i] = 0;
            __syncthreads();
            int s[SRV_R];
            uint32_t rank[SRV_R], off[SRV_R];
#pragma unroll
            for (int i = 0; i < SRV_R; i++) {
                s[i] = -1;
                if (base + i >= s1) continue;
                const uint32_t cell = C.cells[kk][base + i - S.r0];
                if (cell == SRV_CS_NONE) continue;
                s[i] = (int)((int64_t)cell / sl);
                off[i] = (uint32_t)((int64_t)cell % sl);
                rank[i] = atomicAdd(&tof[s[i]], 1u);
            }
            __syncthreads();
            srv_block_scan(tof, nsk, wsum);
#pragma unroll
            for (int i = 0; i < SRV_R; i++) {
                if (s[i] < 0) continue;
                rank[i] += tof[s[i]];                     /* sorted place in the tile */
                spos[rank[i]] = (uint16_t)s[i];
                sroff[rank[i]] = (uint16_t)off[i];
            }
            __syncthreads();
            const int n = (int)tof[nsk];
            for (int j = threadIdx.x; j < n; j += 1024) {
                const int sj = spos[j];
                C.roff[lh[sb + sj] + (uint32_t)j - tof[sj]] = sroff[j];
            }
#pragma unroll
            for (int q = 0; q < TDD_SRV_MAX_VALS; q++) {
                if (!((C.vmask[kk] >> q) & 1)) continue;
                __syncthreads();                          /* the previous word is out */
#pragma unroll
                for (int i = 0; i < SRV_R; i++)
                    if (s[i] >= 0) sval[rank[i]] = (uint64_t)v[q][i];
                __syncthreads();
                for (int j = threadIdx.x; j < n; j += 1024) {
                    const int sj = spos[j];
                    C.rval[q][lh[sb + sj] + (uint32_t)j - tof[sj]] = sval[j];
                }
            }
            __syncthreads();
            for (int i = threadIdx.x; i < nsk; i += 1024) lh[sb + i] += tof[i + 1] - tof[i];
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(512) void k_srv_cfold(SrvCubePlan C) {
    extern __shared__ uint64_t lt[];          /* count, then each value word: [g][v] each */
    __shared__ int64_t st[SRV_CS_WGMAX];      /* segment b: first record */
    __shared__ uint32_t pre[SRV_CS_WGMAX + 1]; /* records before segment b */
    const int s = blockIdx.x;
    const int kk = srv_cs_key(C, s);
    const int64_t G = C.G[kk], Q = C.Q, gw = C.gw[kk], sl = gw * Q;
    const int64_t g0 = (int64_t)(s - C.sbase[kk]) * gw;
    if (g0 >= G) return;                      /* past the cube: a run group's tail */
    const int64_t ng = g0 + gw <= G ? gw : G - g0;
    const int vm = C.vmask[kk];
    int nw = 0;
    for (int j = 0; j < TDD_SRV_MAX_VALS; j++) nw += (vm >> j) & 1;
    const int nwg = C.nwg;
    uint32_t len = 0;
    if ((int)threadIdx.x < nwg) {
        const int64_t at = srv_cs_at(C, kk, s, threadIdx.x);
        st[threadIdx.x] = C.O[at];
        len = (uint32_t)(C.O[at + 1] - C.O[at]);
    }
    for (int64_t i = threadIdx.x; i < sl * (1 + nw); i += 512) lt[i] = 0;
    /* inclusive scan of len over the 512 threads (Hillis-Steele in pre) */
    pre[threadIdx.x + 1] = len;
    if (threadIdx.x == 0) pre[0] = 0;
    __syncthreads();
    for (int d = 1; d < 512; d <<= 1) {
        const uint32_t add = (int)threadIdx.x >= d ? pre[threadIdx.x + 1 - d] : 0u;
        __syncthreads();
        pre[threadIdx.x + 1] += add;
        __syncthreads();
    }
    const uint32_t total = pre[nwg];
    for (uint32_t r = threadIdx.x; r < total; r += 512) {
        int lo = 0, hi = nwg - 1;                 /* last segment with pre[b] <= r */
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= r) lo = mid; else hi = mid - 1;
        }
        const int64_t at = st[lo] + (r - pre[lo]);
        const int64_t o = C.roff[at];
        atomicAdd((unsigned long long*)&lt[o], 1ull);
        int w = 0;
        for (int j = 0; j < TDD_SRV_MAX_VALS; j++) {
            if (!((vm >> j) & 1)) continue;
            w++;
            const uint64_t x = C.rval[j][at];
            if ((C.ftype >> j) & 1)
                atomicAdd((double*)&lt[w * sl + o], __longlong_as_double((long long)x));
            else
                atomicAdd((unsigned long long*)&lt[w * sl + o], (unsigned long long)x);
        }
    }
    __syncthreads();
    /* per (group, word): the raw cells, prefix and suffix sums over v; a
     * count word also gives the pcol sums n * (qlo + v) */
    const SrvCube& K = C.cube[kk];
    const int modes = C.modes[kk];
    for (int64_t t = threadIdx.x; t < ng * (1 + nw); t += 512) {
        const int64_t gl = t % ng, word = t / ng, g = g0 + gl;
        int j = -1;                                /* value column of the word */
        for (int jj = 0, w = 0; jj < TDD_SRV_MAX_VALS; jj++) {
            if (!((vm >> jj) & 1)) continue;
            if (++w == word) j = jj;
        }
        const bool f64 = j >= 0 && ((C.ftype >> j) & 1);
        const uint64_t* col = lt + word * sl + gl * Q;
        for (int dir = 0; dir < 3; dir++) {
            const int mode = dir == 0 ? M_RAW : dir == 1 ? M_PFX : M_SFX;
            if (!((modes >> mode) & 1)) continue;
            uint64_t acc = 0;
            int64_t sq = 0;
            for (int64_t i = 0; i < Q; i++) {
                const int64_t v = mode == M_SFX ? Q - 1 - i : i;
                const uint64_t x = col[v];
                const int64_t at = v * G + g;
                if (mode == M_RAW) {
                    if (j < 0) K.cnt[M_RAW][at] = x; else K.val[M_RAW][j][at] = x;
                    continue;
                }
                if (f64)
                    acc = (uint64_t)__double_as_longlong(__longlong_as_double((long long)acc) +
                                                         __longlong_as_double((long long)x));
                else
                    acc += x;
                if (j < 0) {
                    sq += (int64_t)x * (C.qlo + v);
                    K.cnt[mode][at] = acc;
                    K.sq[mode][at] = sq;
                } else {
                    K.val[mode][j][at] = acc;
                }
            }
        }
    }
}

/* Row I of a lane's four: fold it into request R's table when its
 * predicate bit is set and its key has a dim row.  I is a compile-time
 * index so the row values stay in registers. */
template <int I>
__device__ __forceinline__ void dir_fold(ConstReq& R, const SrvDirArgs& A, uint32_t pm,
                                         bool in, const int64_t (&fk)[4], const int64_t (&q)[4],
                                         const int64_t (&v)[TDD_SRV_MAX_VALS][4], uint64_t* lt) {
    const uint64_t k = (uint64_t)fk[I] - (uint64_t)A.kmin;
    if (!in || !((pm >> I) & 1) || k >= (uint64_t)A.range) return;
    const int32_t g = A.map[R.key][k];
    if (g < 0) return;
    uint64_t* c = R.lds >= 0 ? lt + R.lds + (int64_t)g * R.ncells : R.tab + (int64_t)g * R.ncells;
    atomicAdd((unsigned long long*)&c[0], 1ull);
    for (int cc = 1; cc < R.ncells; cc++) {
        uint64_t x;
        switch (R.src[cc]) {                       /* value column, or pcol (-2) */
        case 0: x = (uint64_t)v[0][I]; break;
        case 1: x = (uint64_t)v[1][I]; break;
        case 2: x = (uint64_t)v[2][I]; break;
        case 3: x = (uint64_t)v[3][I]; break;
        default: x = (uint64_t)q[I]; break;
        }
        if (R.kind[cc] == C_ADD_F64)
            atomicAdd((double*)&c[cc], __longlong_as_double((long long)x));
        else
            atomicAdd((unsigned long long*)&c[cc], (unsigned long long)x);
    }
}

/* Group slice of row r, -1 without a dim row. */
__device__ __forceinline__ int32_t bucket_slice(const SrvBucket& B, int64_t r, int32_t* g) {
    const uint64_t k = (uint64_t)ld_i64(B.fk, r) - (uint64_t)B.kmin;
    *g = k < (uint64_t)B.range ? B.map[k] : -1;
    return *g >= 0 ? *g / B.sg : -1;
}

__global__ __launch_bounds__(256) void k_srv_bhist(SrvBucket B) {
    extern __shared__ uint32_t h[];
    for (int i = threadIdx.x; i < B.S; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int64_t t = blockIdx.x, lo = B.r0 + t * SRV_TILE;
    const int64_t hi = lo + SRV_TILE < B.r1 ? lo + SRV_TILE : B.r1;
    for (int64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) {
        int32_t g;
        const int32_t sl = bucket_slice(B, r, &g);
        B.rowg[r - B.r0] = g;                             /* the scatter reads it back */
        if (sl >= 0) atomicAdd(&h[sl], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < B.S; i += blockDim.x) B.hist[(int64_t)i * B.ntiles + t] = h[i];
}

/* Records of a tile at their slice's run (a slice's records in no
 * particular order within the tile: the folds are sums).  The tile's rows
 * are counting-sorted by slice in LDS and each record array goes out
 * through an LDS stage, so a slice's records from the tile are written as
 * one run by consecutive lanes. */
#define SRV_BR (SRV_TILE / 256)                   /* rows per thread of a tile */
__device__ __forceinline__ uint64_t rec_get(const tdd_col_t& c, int64_t r) {
    switch (c.esz) {
    case 8: return ((const uint64_t*)c.data)[r];
    case 4: return ((const uint32_t*)c.data)[r];
    case 2: return ((const uint16_t*)c.data)[r];
    default: return ((const uint8_t*)c.data)[r];
    }
}

__device__ __forceinline__ void rec_set(void* d, int esz, int64_t at, uint64_t v) {
    switch (esz) {
    case 8: ((uint64_t*)d)[at] = v; break;
    case 4: ((uint32_t*)d)[at] = (uint32_t)v; break;
    case 2: ((uint16_t*)d)[at] = (uint16_t)v; break;
    default: ((uint8_t*)d)[at] = (uint8_t)v; break;
    }
}

__global__ __launch_bounds__(256) void k_srv_bscatter(SrvBucket B) {
    extern __shared__ uint64_t bsm[];
    uint64_t* stage = bsm;                                /* SRV_TILE */
    uint32_t* toff = (uint32_t*)(stage + SRV_TILE);       /* S + 1: tile offsets */
    uint16_t* spos = (uint16_t*)(toff + B.S + 1);         /* SRV_TILE: slice of record i */
    __shared__ uint32_t wsum[256 / WAVE];
    const int64_t t = blockIdx.x, lo = B.r0 + t * SRV_TILE;
    const int64_t hi = lo + SRV_TILE < B.r1 ? lo + SRV_TILE : B.r1;
    for (int i = threadIdx.x; i <= B.S; i += blockDim.x) toff[i] = 0;
    __syncthreads();
    int32_t sl[SRV_BR], g[SRV_BR];
    uint32_t rk[SRV_BR];
#pragma unroll
    for (int k = 0; k < SRV_BR; k++) {
        const int64_t r = lo + threadIdx.x + (int64_t)k * 256;
        sl[k] = -1;
        if (r < hi) {
            g[k] = B.rowg[r - B.r0];
            sl[k] = g[k] >= 0 ? g[k] / B.sg : -1;
        }
        rk[k] = sl[k] >= 0 ? atomicAdd(&toff[sl[k]], 1u) : 0u;
    }
    __syncthreads();
    /* exclusive scan of the tile's counts (S <= 4,096: 16 a thread) */
    {
        const int per = (B.S + 255) / 256;
        const int i0 = threadIdx.x * per, i1 = i0 + per < B.S ? i0 + per : B.S;
        uint32_t tsum = 0;
        for (int i = i0; i < i1; i++) tsum += toff[i];
        const int lane = lane_id(), w = threadIdx.x / WAVE;
        uint32_t inc = tsum;
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        if (lane == WAVE - 1) wsum[w] = inc;
        __syncthreads();
        uint32_t run = inc - tsum;
        for (int k = 0; k < w; k++) run += wsum[k];
        for (int i = i0; i < i1; i++) {
            const uint32_t c = toff[i];
            toff[i] = run;
            run += c;
        }
        __syncthreads();
    }
    uint32_t n = 0;
    for (int k = 0; k < 256 / WAVE; k++) n += wsum[k];
#pragma unroll
    for (int k = 0; k < SRV_BR; k++)
        if (sl[k] >= 0) {
            rk[k] += toff[sl[k]];                         /* sorted place in the tile */
            spos[rk[k]] = (uint16_t)sl[k];
        }
    __syncthreads();
    int64_t at[SRV_BR];                                   /* sorted record i = tid + 256 j */
#pragma unroll
    for (int j = 0; j < SRV_BR; j++) {
        const uint32_t i = threadIdx.x + 256u * j;
        at[j] = -1;
        if (i < n) {
            const int s = spos[i];
            at[j] = B.off[(int64_t)s * B.ntiles + t] + (i - toff[s]);
        }
    }
    for (int c = -1; c < B.nc; c++) {                     /* the group, then each column */
        __syncthreads();                                  /* the previous array is out */
#pragma unroll
        for (int k = 0; k < SRV_BR; k++)
            if (sl[k] >= 0)
                stage[rk[k]] = c < 0 ? (uint64_t)(uint32_t)g[k]
                                     : rec_get(B.src[c], lo + threadIdx.x + (int64_t)k * 256);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < SRV_BR; j++) {
            if (at[j] < 0) continue;
            const uint64_t v = stage[threadIdx.x + 256u * j];
            if (c < 0) B.rg[at[j]] = (uint32_t)v;
            else rec_set(B.rc[c], B.src[c].esz, at[j], v);
        }
    }
}

/* Record I of a lane's four into request R's LDS slice table. */
template <int I>
__device__ __forceinline__ void bfold_rec(ConstReq& R, uint32_t pm, bool in,
                                          const uint32_t (&g)[4], int64_t gbase, int64_t rec,
                                          uint64_t* lt) {
    if (!in || !((pm >> I) & 1)) return;
    uint64_t* c = lt + R.lds + ((int64_t)g[I] - gbase) * R.ncells;
    atomicAdd((unsigned long long*)&c[0], 1ull);
    for (int cc = 1; cc < R.ncells; cc++) {
        const tdd_col_t col = R.P.cols[R.src[cc]];
        if (R.kind[cc] == C_ADD_F64)
            atomicAdd((double*)&c[cc], ld_f64(col, rec));
        else
            atomicAdd((unsigned long long*)&c[cc], (unsigned long long)ld_i64(col, rec));
    }
}

__global__ __launch_bounds__(256) void k_srv_bfold(SrvFoldArgs F) {
    extern __shared__ uint64_t lt[];
    /* block -> (slice, part, request group), the group varying fastest so
     * that workgroups running together read the same records */
    const int q = (int)(blockIdx.x % (unsigned)F.ngroups);
    const int64_t sp = blockIdx.x / (unsigned)F.ngroups;
    const int64_t sl = sp / F.parts, part = sp % F.parts;
    const SrvDirGroup grp = F.groups[q];
    for (int i = threadIdx.x; i < grp.words; i += blockDim.x) lt[i] = 0;
    __syncthreads();
    const int64_t b0 = F.off[sl * F.ntiles], b1 = F.off[(sl + 1) * F.ntiles];
    const int64_t per = ((b1 - b0 + F.parts - 1) / F.parts + 255) / 256 * 256;
    const int64_t beg = b0 + part * per < b1 ? b0 + part * per : b1;
    const int64_t end = beg + per < b1 ? beg + per : b1;
    const int64_t gbase = sl * F.sg;
    const int w = threadIdx.x / WAVE, lane = lane_id();
    if (grp.tmode) {                              /* threshold buckets */
        __shared__ int64_t scut[SRV_TCUTS];
        ConstReq& R0 = ((ConstReq*)F.reqs)[grp.first];
        const tdd_col_t xc = R0.P.cols[R0.P.ins[0].a];
        const int nc = R0.ncells;
        const int64_t bw = (int64_t)F.sg * nc;   /* words of one bucket's table */
        if ((int)threadIdx.x < grp.m) scut[threadIdx.x] = F.groups[q].cut[threadIdx.x];
        __syncthreads();
        for (int64_t r = beg + threadIdx.x; r < end; r += blockDim.x) {
            int lo = 0, hi = grp.m;              /* b = the cuts below x */
            if (grp.as_f64) {
                const double x = ld_f64(xc, r);
                if (x != x) continue;             /* NaN: no comparison holds */
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    const double c = __longlong_as_double((long long)scut[mid]);
                    if (grp.strict ? c < x : c <= x) lo = mid + 1; else hi = mid;
                }
            } else {
                const int64_t x = ld_i64(xc, r);
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    const int64_t c = scut[mid];
                    if (grp.strict ? c < x : c <= x) lo = mid + 1; else hi = mid;
                }
            }
            const int b = lo;
            uint64_t* c = lt + b * bw + ((int64_t)F.rg[r] - gbase) * nc;
            atomicAdd((unsigned long long*)&c[0], 1ull);
            for (int cc = 1; cc < nc; cc++) {
                const tdd_col_t col = R0.P.cols[R0.src[cc]];
                if (R0.kind[cc] == C_ADD_F64) atomicAdd((double*)&c[cc], ld_f64(col, r));
                else atomicAdd((unsigned long long*)&c[cc], (unsigned long long)ld_i64(col, r));
            }
        }
        __syncthreads();
        /* running sums over the buckets, in place: from the top when the
         * requests pass above their cuts, else from the bottom */
        const int64_t ns = gbase + F.sg <= R0.G ? F.sg : R0.G - gbase;
        for (int64_t i = threadIdx.x; i < ns * nc; i += blockDim.x) {
            const bool f = R0.kind[i % nc] == C_ADD_F64;
            uint64_t a = 0;
            double af = 0.0;
            for (int k = 0; k <= grp.m; k++) {
                uint64_t* p = lt + (grp.up ? grp.m - k : k) * bw + i;
                if (f) {
                    af += __longlong_as_double((long long)*p);
                    *p = (uint64_t)__double_as_longlong(af);
                } else {
                    a += *p;
                    *p = a;
                }
            }
        }
        __syncthreads();
        for (int j = 0; j < grp.n; j++) {         /* request j: its buckets' sum */
            ConstReq& R = ((ConstReq*)F.reqs)[grp.first + j];
            const int64_t n = gbase + F.sg <= R.G ? F.sg : R.G - gbase;
            const int64_t b0 = (int64_t)(grp.up ? R.blo : R.bhi) * bw;
            for (int64_t i = threadIdx.x; i < n * nc; i += blockDim.x) {
                const bool f = R.kind[i % nc] == C_ADD_F64;
                uint64_t x = 0;
                double xf = 0.0;
                if (grp.up ? R.blo <= grp.m : true) {
                    x = lt[b0 + i];
                    xf = __longlong_as_double((long long)x);
                }
                if (!x) continue;
                uint64_t* d = R.tab + gbase * nc + i;
                if (F.parts > 1) {
                    if (f) atomicAdd((double*)d, xf);
                    else atomicAdd((unsigned long long*)d, (unsigned long long)x);
                } else if (f) {
                    *d = (uint64_t)__double_as_longlong(__longlong_as_double((long long)*d) + xf);
                } else {
                    *d += x;
                }
            }
        }
        return;
    }
    for (int64_t t = beg + (int64_t)w * 256; t < end; t += 1024) {
        const int64_t r0 = t + lane;
        uint32_t g[4];
#pragma unroll
        for (int i = 0; i < 4; i++) g[i] = r0 + 64 * i < end ? F.rg[r0 + 64 * i] : 0u;
        for (int j = 0; j < grp.n; j++) {
            ConstReq& R = ((ConstReq*)F.reqs)[grp.first + j];
            const uint32_t pm = pred_eval4(R.P, r0, end);
            bfold_rec<0>(R, pm, r0 < end, g, gbase, r0, lt);
            bfold_rec<1>(R, pm, r0 + 64 < end, g, gbase, r0 + 64, lt);
            bfold_rec<2>(R, pm, r0 + 128 < end, g, gbase, r0 + 128, lt);
            bfold_rec<3>(R, pm, r0 + 192 < end, g, gbase, r0 + 192, lt);
        }
    }
    __syncthreads();
    /* this slice of these requests' tables: added by this workgroup alone
     * when the slice has one part, else with atomics */
    for (int j = 0; j < grp.n; j++) {
        ConstReq& R = ((ConstReq*)F.reqs)[grp.first + j];
        const int64_t n = gbase + F.sg <= R.G ? F.sg : R.G - gbase;
        for (int64_t i = threadIdx.x; i < n * R.ncells; i += blockDim.x) {
            const uint64_t x = lt[R.lds + i];
            if (!x) continue;
            uint64_t* d = R.tab + gbase * R.ncells + i;
            const bool f = R.kind[i % R.ncells] == C_ADD_F64;
            if (F.parts > 1) {
                if (f) atomicAdd((double*)d, __longlong_as_double((long long)x));
                else atomicAdd((unsigned long long*)d, (unsigned long long)x);
            } else if (f) {
                *d = (uint64_t)__double_as_longlong(__longlong_as_double((long long)*d) +
                                                    __longlong_as_double((long long)x));
            } else {
                *d += x;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_srv_direct(SrvDirArgs A) {
    extern __shared__ uint64_t lt[];
    const int gi = (int)(blockIdx.x % (unsigned)A.ngroups);
    const int64_t rb = blockIdx.x / (unsigned)A.ngroups;
    const SrvDirGroup grp = A.groups[gi];
    for (int i = threadIdx.x; i < grp.words; i += blockDim.x) lt[i] = 0;  /* COUNT / ADD: 0 */
    __syncthreads();
    const int64_t lo = A.r0 + rb * A.span;
    const int64_t hi = lo + A.span < A.r1 ? lo + A.span : A.r1;
    const int w = threadIdx.x / WAVE, lane = lane_id();
    for (int64_t t = lo + (int64_t)w * 256; t < hi; t += 1024) {
        const int64_t r0 = t + lane;               /* rows r0 + 64 i, i = 0..3 */
        int64_t fk[4], q[4] = {0, 0, 0, 0}, v[TDD_SRV_MAX_VALS][4];
        ld4_i64(A.fk, r0, hi, fk);
        if (A.pcol.data) ld4_i64(A.pcol, r0, hi, q);
#pragma unroll
        for (int j = 0; j < TDD_SRV_MAX_VALS; j++) {
            if (A.vals[j].data) ld4_i64(A.vals[j], r0, hi, v[j]);
            else for (int i = 0; i < 4; i++) v[j][i] = 0;
        }
        for (int j = 0; j < grp.n; j++) {
            ConstReq& R = ((ConstReq*)A.reqs)[grp.first + j];
            const uint32_t pm = pred_eval4(R.P, r0, hi);
            dir_fold<0>(R, A, pm, r0 < hi, fk, q, v, lt);
            dir_fold<1>(R, A, pm, r0 + 64 < hi, fk, q, v, lt);
            dir_fold<2>(R, A, pm, r0 + 128 < hi, fk, q, v, lt);
            dir_fold<3>(R, A, pm, r0 + 192 < hi, fk, q, v, lt);
        }
    }
    __syncthreads();
    for (int j = 0; j < grp.n; j++) {             /* LDS tables into the requests' tables */
        ConstReq& R = ((ConstReq*)A.reqs)[grp.first + j];
        if (R.lds < 0) continue;
        for (int64_t g = threadIdx.x; g < R.G; g += blockDim.x) {
            const uint64_t* c = lt + R.lds + g * R.ncells;
            if (!c[0]) continue;
            uint64_t* d = R.tab + g * R.ncells;
            atomicAdd((unsigned long long*)&d[0], (unsigned long long)c[0]);
            for (int cc = 1; cc < R.ncells; cc++) {
                if (R.kind[cc] == C_ADD_F64)
                    atomicAdd((double*)&d[cc], __longlong_as_double((long long)c[cc]));
                else
                    atomicAdd((unsigned long long*)&d[cc], (unsigned long long)c[cc]);
            }
        }
    }
}

/* Prefix (v' <= v) and suffix (v' >= v) sums of one key's raw cube; a
 * thread owns one group and walks the pcol axis (coalesced across g). */
__global__ void k_srv_cumul(SrvCube C, int64_t G, int Q, int64_t qlo, int nv, uint32_t ftype,
                            int vmask, int want_pfx, int want_sfx) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    for (int dir = 0; dir < 2; dir++) {
        const int mode = dir ? M_SFX : M_PFX;
        if (!(dir ? want_sfx : want_pfx)) continue;
        uint64_t c = 0;
        int64_t s = 0;
        uint64_t acc[TDD_SRV_MAX_VALS] = {0, 0, 0, 0};
        for (int i = 0; i < Q; i++) {
            const int v = dir ? Q - 1 - i : i;
            const int64_t at = (int64_t)v * G + g;
            const uint64_t n = C.cnt[M_RAW][at];
            c += n;
            s += (int64_t)n * (qlo + v);
            C.cnt[mode][at] = c;
            C.sq[mode][at] = s;
            for (int j = 0; j < nv; j++) {
                if (!((vmask >> j) & 1)) continue;
                const uint64_t x = C.val[M_RAW][j][at];
                if ((ftype >> j) & 1)
                    acc[j] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)acc[j]) +
                                                            __longlong_as_double((long long)x));
                else
                    acc[j] += x;
                C.val[mode][j][at] = acc[j];
            }
        }
    }
}

__global__ void k_srv_accum(const SrvAccArgs* reqs, const SrvCube* cubes, int64_t qlo) {
    const SrvAccArgs& A = reqs[blockIdx.y];
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.G || A.mode == M_NONE) return;
    const SrvCube& C = cubes[A.key];
    const int64_t at = A.idx * A.G + g;
    const uint64_t n = C.cnt[A.mode][at];
    if (!n) return;
    uint64_t* c = A.tab + g * A.ncells;
    c[0] += n;
    for (int i = 1; i < A.ncells; i++) {
        const int s = A.src[i];
        uint64_t x;
        if (s == -2)
            x = A.mode == M_RAW ? (uint64_t)((int64_t)n * (qlo + A.idx)) : (uint64_t)C.sq[A.mode][at];
        else
            x = C.val[A.mode][s][at];
        if (A.kind[i] == C_ADD_F64)
            c[i] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)c[i]) +
                                                  __longlong_as_double((long long)x));
        else
            c[i] += x;
    }
}

struct SrvEmitArgs {
    GroupPlan          P;
    const uint64_t*    tab;
    int64_t            G;
    tdd_group_result_t out;
};

__global__ void k_srv_flags(const SrvEmitArgs* E, int64_t gmax, uint32_t* flag) {
    const int r = blockIdx.y;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= gmax) return;
    const SrvEmitArgs& A = E[r];
    flag[(int64_t)r * gmax + g] = g < A.G && A.tab[g * A.P.ncells] != 0;
}

/* first[r] = pos[r * gmax]: the flagged rows before request r. */
__global__ void k_srv_firsts(const int64_t* pos, int64_t gmax, int n, int64_t* first) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= n) first[r] = pos[(int64_t)r * gmax];
}

__global__ void k_srv_emit(const SrvEmitArgs* E, int64_t gmax, const int64_t* pos) {
    const int r = blockIdx.y;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const SrvEmitArgs& A = E[r];
    if (g >= A.G) return;
    const uint64_t* c = A.tab + g * A.P.ncells;
    if (c[0] == 0) return;
    const int64_t o = pos[(int64_t)r * gmax + g] - pos[(int64_t)r * gmax];
    st_int(A.out.keys[0], o, A.P.keys[0].esz, A.P.kmin[0] + g);
    for (int a = 0; a < A.P.n_aggs; a++) cells_emit(c, A.P, a, A.out.vals[a], o);
}

/* ------------------------------------------------------------ host side -- */

static bool srv_int(const tdd_col_t& c) {
    return c.data && c.type != TD_F64 && c.esz >= 1 && c.esz <= 8;
}

static int col_range(const tdd_col_t& c, int64_t n, int64_t* lo, int64_t* hi) {
    if (c.ranged) {
        *lo = c.lo;
        *hi = c.hi;
        return TD_OK;
    }
    return tdd_minmax(&c, n, nullptr, lo, hi);
}

extern "C" int tdd_srv_open(const tdd_srv_spec_t* spec, tdd_srv_t** out) {
    *out = nullptr;
    const tdd_srv_spec_t& S = *spec;
    if (S.n_keys < 1 || S.n_keys > TDD_SRV_MAX_KEYS || S.n_vals < 0 ||
        S.n_vals > TDD_SRV_MAX_VALS || S.nf < 0 || S.nd < 1)
        return TD_ERR_NYI;
    if (!srv_int(S.fk) || !srv_int(S.pk) || !srv_int(S.pcol)) return TD_ERR_NYI;
    for (int j = 0; j < S.n_vals; j++)
        if (!S.vals[j].data || (S.vals[j].type != TD_F64 && !srv_int(S.vals[j]))) return TD_ERR_NYI;
    tdd_srv_t* s = (tdd_srv_t*)calloc(1, sizeof(tdd_srv_t));
    if (!s) return TD_ERR_OOM;
    s->spec = S;
    int64_t plo = 0, phi = -1, qlo = 0, qhi = 0;
    int rc = col_range(S.pk, S.nd, &plo, &phi);
    if (!rc && S.nf > 0) rc = col_range(S.pcol, S.nf, &qlo, &qhi);
    if (rc) { free(s); return rc; }
    if (phi < plo || phi - plo + 1 > ((int64_t)1 << 31) || phi - plo + 1 > 16 * S.nd + 4096 ||
        qhi < qlo || qhi - qlo + 1 > SRV_MAX_Q) {
        free(s);
        return TD_ERR_NYI;
    }
    s->kmin = plo;
    s->range = phi - plo + 1;
    s->qlo = qlo;
    s->Q = (int32_t)(qhi - qlo + 1);
    int64_t chunk = S.chunk_rows > 0 ? S.chunk_rows : ceil_div(S.nf > 0 ? S.nf : 1, 8);
    chunk = ceil_div(chunk, 4096) * 4096;
    s->chunk = chunk;
    s->nchunks = (int32_t)(S.nf > 0 ? ceil_div(S.nf, chunk) : 1);
    Scratch dup(8);
    if (!dup.ok()) { free(s); return TD_ERR_OOM; }
    const int nv = S.n_vals;
    int any = 0;
    for (int k = 0; k < S.n_keys; k++) {
        int64_t klo = 0, khi = -1;
        if (!srv_int(S.keys[k]) || col_range(S.keys[k], S.nd, &klo, &khi) || khi < klo ||
            khi - klo + 1 > ((int64_t)1 << 24) || (khi - klo + 1) * s->Q > SRV_MAX_CELLS)
            continue;                                   /* requests on this key run alone */
        const int64_t G = khi - klo + 1, QG = G * s->Q;
        s->map[k] = (int32_t*)td__arena_alloc((size_t)s->range * 4);
        const size_t per = (size_t)QG * 8;
        const size_t cube_bytes = per * (size_t)(1 + nv + 2 * (2 + nv));
        s->cube_mem[k] = (uint8_t*)td__arena_alloc(cube_bytes);
        if (!s->map[k] || !s->cube_mem[k]) { tdd_srv_close(s); return TD_ERR_OOM; }
        s->bytes += (int64_t)s->range * 4 + (int64_t)cube_bytes;
        uint8_t* p = s->cube_mem[k];
        SrvCube& C = s->cube[k];
        C.cnt[M_RAW] = (uint64_t*)p; p += per;
        for (int j = 0; j < nv; j++) { C.val[M_RAW][j] = (uint64_t*)p; p += per; }
        for (int m = M_SFX; m <= M_PFX; m++) {
            C.cnt[m] = (uint64_t*)p; p += per;
            C.sq[m] = (int64_t*)p; p += per;
            for (int j = 0; j < nv; j++) { C.val[m][j] = (uint64_t*)p; p += per; }
        }
        rc = tdd_fill(s->map[k], ~0ull, (size_t)s->range * 4);
        if (!rc) rc = tdd_fill(dup.p, 0, 8);
        if (rc) { tdd_srv_close(s); return rc; }
        LAUNCH("serve_map", k_srv_map, stream_grid(ceil_div(S.nd, 256 * 4)), 256, 0, S.pk,
               S.keys[k], S.nd, s->kmin, klo, s->map[k], dup.as<uint32_t>());
        uint32_t d = 0;
        HC(hipMemcpyAsync(&d, dup.p, 4, hipMemcpyDeviceToHost, tdd::g_stream));
        HC(hipStreamSynchronize(tdd::g_stream));
        if (d) { tdd_srv_close(s); return TD_ERR_NYI; }   /* duplicate build keys */
        s->G[k] = G;
        s->klo[k] = klo;
        s->ok[k] = 1;
        any = 1;
    }
    if (!any) { tdd_srv_close(s); return TD_ERR_NYI; }
    {   /* arrays of the sliced cube; without them every step takes k_srv_cube */
        int nok = 0;
        for (int k = 0; k < S.n_keys; k++) nok += s->ok[k];
        const int64_t rows = S.nf < chunk ? S.nf : chunk;
        const int64_t nrec = rows * nok;
        auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };   /* sections 256-aligned */
        const size_t b_cells = al((size_t)nrec * 4), b_roff = al((size_t)nrec * 2);
        const size_t b_rval = (size_t)nrec * 8 * (size_t)nv;
        const size_t b_C = (size_t)SRV_CS_MAXS * SRV_CS_WGMAX * 4;
        const size_t b_O = ((size_t)SRV_CS_MAXS * SRV_CS_WGMAX + 1) * 8;
        const int js = nok == 1 ? 1 : nok == 2 ? 2 : 4;
        const size_t b_J = (size_t)s->range * 4 * (size_t)js;
        const size_t tot = b_cells + b_roff + b_rval + b_C + b_O + b_J;
        if (rows > 0 && nrec < ((int64_t)1 << 32)) {
            s->cs_mem = (uint8_t*)td__arena_alloc(tot);
            if (s->cs_mem) {
                uint8_t* p = s->cs_mem;
                s->cs_cells = (uint32_t*)p; p += b_cells;
                s->cs_roff = (uint16_t*)p; p += b_roff;
                s->cs_rval = (uint64_t*)p; p += b_rval;
                s->cs_C = (uint32_t*)p; p += b_C;
                s->cs_O = (int64_t*)p; p += b_O;
                s->cs_jmap = (int32_t*)p;
                s->cs_rows = rows;
                s->cs_nk = nok;
                s->cs_js = js;
                s->bytes += (int64_t)tot;
                const int32_t* m[4] = {nullptr, nullptr, nullptr, nullptr};
                for (int k = 0, sl = 0; k < S.n_keys; k++) if (s->ok[k]) m[sl++] = s->map[k];
                LAUNCH("serve_map", k_srv_jmap, stream_grid(ceil_div(s->range, 256 * 4)), 256, 0,
                       m[0], m[1], m[2], m[3], s->range, js, s->cs_jmap);
            }
        }
    }
    s->d_cubes = (SrvCube*)td__arena_alloc(sizeof(SrvCube) * TDD_SRV_MAX_KEYS);
    if (!s->d_cubes) { tdd_srv_close(s); return TD_ERR_OOM; }
    HC(hipMemcpyAsync(s->d_cubes, s->cube, sizeof(s->cube), hipMemcpyHostToDevice, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    *out = s;
    return TD_OK;
}

extern "C" int64_t tdd_srv_chunks(const tdd_srv_t* s) { return s ? s->nchunks : 0; }
extern "C" int64_t tdd_srv_bytes(const tdd_srv_t* s) { return s ? s->bytes : 0; }

extern "C" void tdd_srv_close(tdd_srv_t* s) {
    if (!s) return;
    (void)hipStreamSynchronize(tdd::g_stream);
    for (int k = 0; k < TDD_SRV_MAX_KEYS; k++) {
        if (s->map[k]) td__arena_free(s->map[k]);
        if (s->cube_mem[k]) td__arena_free(s->cube_mem[k]);
    }
    if (s->d_cubes) td__arena_free(s->d_cubes);
    if (s->cs_mem) td__arena_free(s->cs_mem);
    free(s);
}

/* Cube mode and cell of a request's predicate over [qlo, qlo + Q). */
static void req_cell(const tdd_srv_t* s, int cmp, int64_t t, int32_t* mode, int64_t* idx) {
    const int64_t lo = s->qlo, hi = s->qlo + s->Q - 1;
    *mode = M_NONE;
    *idx = 0;
    switch (cmp) {
    case OP_GT: case OP_GE: {                       /* v >= first */
        int64_t first;
        if (cmp == OP_GT) { if (t >= hi) return; first = t < lo ? lo : t + 1; }
        else { if (t > hi) return; first = t < lo ? lo : t; }
        *mode = M_SFX;
        *idx = first - lo;
        return;
    }
    case OP_LT: case OP_LE: {                       /* v <= last */
        int64_t last;
        if (cmp == OP_LT) { if (t <= lo) return; last = t > hi ? hi : t - 1; }
        else { if (t < lo) return; last = t > hi ? hi : t; }
        *mode = M_PFX;
        *idx = last - lo;
        return;
    }
    case OP_EQ:
        if (t < lo || t > hi) return;
        *mode = M_RAW;
        *idx = t - lo;
        return;
    default:
        return;
    }
}

/* The request's group plan and cube sources; TD_ERR_NYI if not servable. */
static int req_plan(const tdd_srv_t* s, const tdd_srv_req_t* r, tdd_srv_acc_t* A) {
    memset(A, 0, sizeof(*A));
    const tdd_srv_spec_t& S = s->spec;
    if (r->key < 0 || r->key >= S.n_keys || !s->ok[r->key] || r->n_aggs < 0 ||
        r->n_aggs > TDD_MAX_AGGS)
        return TD_ERR_NYI;
    if (r->pred) {                                  /* direct: its own predicate program */
        if (r->pred->n_ins < 1 || r->pred->n_ins > TDD_PRED_MAX_INS) return TD_ERR_NYI;
        for (int i = 0; i < r->pred->n_ins; i++) {
            const tdd_pred_ins_t& in = r->pred->ins[i];
            const bool cols = in.kind == TDD_P_CMP_CONST || in.kind == TDD_P_CMP_COL ||
                              in.kind == TDD_P_BOOLCOL;
            if (cols && (in.a >= r->pred->n_cols || !r->pred->cols[in.a].data)) return TD_ERR_NYI;
            if (in.kind == TDD_P_CMP_COL &&
                (in.b >= r->pred->n_cols || !r->pred->cols[in.b].data))
                return TD_ERR_NYI;
        }
    } else if (r->cmp != OP_EQ && r->cmp != OP_LT && r->cmp != OP_LE && r->cmp != OP_GT &&
               r->cmp != OP_GE) {
        return TD_ERR_NYI;
    }
    GroupPlan& P = A->P;
    P.n_keys = 1;
    P.n_aggs = r->n_aggs;
    P.keys[0] = S.keys[r->key];
    P.key_side[0] = 1;
    P.kmin[0] = s->klo[r->key];
    P.kstride[0] = 1;
    P.krange[0] = s->G[r->key];
    for (int a = 0; a < r->n_aggs; a++) {
        const uint16_t op = r->ops[a];
        const int src = r->src[a];
        if (op != OP_SUM && op != OP_AVG && op != OP_COUNT) return TD_ERR_NYI;
        if (src >= S.n_vals || src < -2 || (op != OP_COUNT && src == -1)) return TD_ERR_NYI;
        P.ops[a] = op;
        if (src == -2) P.aggs[a] = S.pcol;
        else if (src >= 0) P.aggs[a] = S.vals[src];
        else { P.aggs[a].type = TD_I64; P.aggs[a].esz = 8; }
    }
    int rc = plan_cells(P);
    if (rc) return rc == TD_ERR_TYPE ? TD_ERR_NYI : rc;
    A->req = *r;
    SrvAccArgs& a = A->a;
    a.G = s->G[r->key];
    a.key = r->key;
    a.ncells = P.ncells;
    for (int i = 0; i < P.ncells; i++) {
        a.kind[i] = P.kind[i];
        a.src[i] = i ? (int8_t)r->src[P.col[i]] : -1;
        if (i && P.kind[i] != C_ADD_F64 && P.kind[i] != C_ADD_I64) return TD_ERR_NYI;
    }
    if (r->pred) {
        A->direct = 1;
        A->pred = *r->pred;
        A->req.pred = nullptr;                      /* the caller's program is not kept */
        a.mode = M_NONE;                            /* not folded from the cube */
    } else {
        req_cell(s, r->cmp, r->t, &a.mode, &a.idx);
    }
    return TD_OK;
}

extern "C" int tdd_srv_req_bytes(const tdd_srv_t* s, const tdd_srv_req_t* r, int64_t* bytes) {
    tdd_srv_acc_t A;
    int rc = req_plan(s, r, &A);
    if (rc) return rc;
    const int64_t G = A.a.G;
    int64_t b = G * A.P.ncells * 8 + G * A.P.keys[0].esz;
    for (int a = 0; a < A.P.n_aggs; a++) b += G * 8;
    *bytes = b;
    return TD_OK;
}

extern "C" int tdd_srv_begin(tdd_srv_t* s, const tdd_srv_req_t* r, tdd_srv_acc_t** out) {
    *out = nullptr;
    tdd_srv_acc_t* A = (tdd_srv_acc_t*)malloc(sizeof(tdd_srv_acc_t));
    if (!A) return TD_ERR_OOM;
    int rc = req_plan(s, r, A);
    if (rc) { free(A); return rc; }
    const size_t tb = (size_t)A->a.G * A->P.ncells * 8;
    A->a.tab = (uint64_t*)td__arena_alloc(tb);
    if (!A->a.tab) { free(A); return TD_ERR_OOM; }
    rc = tdd_fill(A->a.tab, 0, tb);                 /* COUNT / ADD identity */
    if (rc) { tdd_srv_acc_free(A); return rc; }
    *out = A;
    return TD_OK;
}

extern "C" void tdd_srv_acc_free(tdd_srv_acc_t* A) {
    if (!A) return;
    if (A->a.tab) td__arena_free(A->a.tab);
    free(A);
}

/* Distinct fact columns a direct request reads (its predicate's, then its
 * sums' sources); -1 when its predicate program plus its sums would pass
 * TDD_PRED_MAX_COLS columns. */
static int req_cols(const tdd_srv_t* s, const tdd_srv_acc_t* A, tdd_col_t* out) {
    int n = 0;
    auto add = [&](const tdd_col_t& c) {
        for (int i = 0; i < n; i++) if (out[i].data == c.data && out[i].type == c.type) return;
        out[n++] = c;
    };
    const tdd_pred_t& P = A->pred;
    for (int i = 0; i < P.n_ins; i++) {
        const tdd_pred_ins_t& in = P.ins[i];
        if (in.kind == TDD_P_CMP_CONST || in.kind == TDD_P_CMP_COL || in.kind == TDD_P_BOOLCOL)
            add(P.cols[in.a]);
        if (in.kind == TDD_P_CMP_COL) add(P.cols[in.b]);
    }
    for (int c = 1; c < A->a.ncells; c++)
        add(A->a.src[c] == -2 ? s->spec.pcol : s->spec.vals[A->a.src[c]]);
    return P.n_cols + A->a.ncells - 1 <= TDD_PRED_MAX_COLS ? n : -1;
}

static int64_t slice_groups(int ncells) {
    const int64_t sg = SRV_SLICE_B / (8 * ncells);
    return sg > 0 ? sg : 1;
}

/* One key's batch of direct requests over chunk rows [r0, r1): the rows
 * bucket-sorted by group slice into records of `cols`, then folded per
 * (slice, request group). */
static int srv_bucket(tdd_srv_t* s, int key, const std::vector<const tdd_srv_acc_t*>& batch,
                      const std::vector<tdd_col_t>& cols, int ncmax, int64_t r0, int64_t r1) {
    const tdd_srv_spec_t& S = s->spec;
    const int64_t G = s->G[key];
    /* requests `col CMP cut` (one comparison with a constant, < <= > >=)
     * that share the column, the comparison's strictness class and the
     * cell layout fold through threshold buckets, up to SRV_TCUTS a group */
    std::map<std::string, std::vector<const tdd_srv_acc_t*>> fam;
    std::vector<const tdd_srv_acc_t*> plain;
    for (const tdd_srv_acc_t* A : batch) {
        const tdd_pred_t& P = A->pred;
        const tdd_pred_ins_t& in = P.ins[0];
        const bool ok = !TUNE_INT("TEIDE_SRV_NO_TFOLD", 0) && P.n_ins == 1 &&
                        in.kind == TDD_P_CMP_CONST &&
                        (in.cmp == OP_GT || in.cmp == OP_GE || in.cmp == OP_LT || in.cmp == OP_LE) &&
                        !(in.as_f64 && in.cf != in.cf);
        if (!ok) { plain.push_back(A); continue; }
        const int strict = in.cmp == OP_GT || in.cmp == OP_LE;
        const int up = in.cmp == OP_GT || in.cmp == OP_GE;
        std::string k;
        auto put = [&](const void* p, size_t n) { k.append((const char*)p, n); };
        put(&P.cols[in.a].data, sizeof(void*));
        put(&P.cols[in.a].type, sizeof(P.cols[in.a].type));
        put(&P.cols[in.a].attrs, sizeof(P.cols[in.a].attrs));
        put(&in.as_f64, 1);
        put(&strict, sizeof(strict));
        put(&up, sizeof(up));
        put(&A->a.ncells, sizeof(A->a.ncells));
        put(A->a.kind, (size_t)A->a.ncells);
        put(A->a.src, (size_t)A->a.ncells);
        fam[k].push_back(A);
    }
    /* threshold groups hold one bucket table per threshold + 1 in LDS: 1 KB
     * slices of a request's table let a group take 77 thresholds (4 KB: 18;
     * 512 B slices, 153, measured no faster: twice the slices) */
    int64_t sg = slice_groups(ncmax);
    if (!fam.empty()) {
        const int64_t sg1 = sg / 4 > 0 ? sg / 4 : 1;
        if (ceil_div(G, sg1) <= SRV_MAX_S) sg = sg1;
    }
    const int64_t tcuts = (SRV_DIR_TLDS / (sg * ncmax * 8)) - 1 < SRV_TCUTS
                              ? (SRV_DIR_TLDS / (sg * ncmax * 8)) - 1 : SRV_TCUTS;
    const int64_t nsl = ceil_div(G, sg);
    const int64_t rows = r1 - r0, ntiles = ceil_div(rows, SRV_TILE);
    SCRATCH(hist, (size_t)(nsl * ntiles) * 4);
    SCRATCH(off, (size_t)(nsl * ntiles + 1) * 8);
    SCRATCH(rg, (size_t)rows * 4);
    SCRATCH(rowg, (size_t)rows * 4);
    SrvBucket B;
    memset(&B, 0, sizeof(B));
    B.fk = S.fk;
    B.map = s->map[key];
    B.r0 = r0;
    B.r1 = r1;
    B.kmin = s->kmin;
    B.range = s->range;
    B.ntiles = ntiles;
    B.sg = (int32_t)sg;
    B.S = (int32_t)nsl;
    B.nc = (int32_t)cols.size();
    B.hist = hist.as<uint32_t>();
    B.off = off.as<int64_t>();
    B.rg = rg.as<uint32_t>();
    B.rowg = rowg.as<int32_t>();
    Scratch rc[SRV_REC_C] = {Scratch(8), Scratch(8), Scratch(8), Scratch(8),
                             Scratch(8), Scratch(8), Scratch(8), Scratch(8)};
    for (int c = 0; c < B.nc; c++) {
        B.src[c] = cols[c];
        if (!rc[c].reset((size_t)rows * cols[c].esz)) return TD_ERR_OOM;
        B.rc[c] = rc[c].p;
    }
    LAUNCH("serve_bucket", k_srv_bhist, (int)ntiles, 256, (size_t)nsl * 4, B);
    int rcode = tdd::scan_u32_i64(hist.as<uint32_t>(), nsl * ntiles, off.as<int64_t>());
    if (rcode) return rcode;
    LAUNCH("serve_bucket", k_srv_bscatter, (int)ntiles, 256,
           (size_t)SRV_TILE * 8 + ((size_t)nsl + 1) * 4 + (size_t)SRV_TILE * 2, B);
    /* the requests' programs over the record columns; sums read record columns */
    std::vector<SrvDirReq> reqs;
    std::vector<SrvDirGroup> groups;
    auto rec_col = [&](const tdd_col_t& c) {
        for (int i = 0; i < B.nc; i++)
            if (cols[i].data == c.data && cols[i].type == c.type) {
                tdd_col_t r = c;
                r.data = B.rc[i];
                return r;
            }
        return c;                                    /* not reached: req_cols lists them */
    };
    auto make_req = [&](const tdd_srv_acc_t* A) {
        SrvDirReq R;
        memset(&R, 0, sizeof(R));
        R.P = A->pred;
        for (int i = 0; i < R.P.n_cols; i++)
            if (R.P.cols[i].data) R.P.cols[i] = rec_col(R.P.cols[i]);
        int nextc = R.P.n_cols;
        for (int c = 1; c < A->a.ncells; c++) {
            R.P.cols[nextc] = rec_col(A->a.src[c] == -2 ? S.pcol : S.vals[A->a.src[c]]);
            R.src[c] = (int8_t)nextc++;
        }
        R.tab = A->a.tab;
        R.G = G;
        R.key = key;
        R.ncells = A->a.ncells;
        memcpy(R.kind, A->a.kind, sizeof(R.kind));
        return R;
    };
    for (const auto& kv : fam) {
        const std::vector<const tdd_srv_acc_t*>& v = kv.second;
        for (size_t i0 = 0; i0 < v.size(); i0 += (size_t)tcuts) {
            const size_t n = v.size() - i0 < (size_t)tcuts ? v.size() - i0 : (size_t)tcuts;
            const tdd_pred_ins_t& in0 = v[i0]->pred.ins[0];
            SrvDirGroup gr;
            memset(&gr, 0, sizeof(gr));
            gr.first = (int32_t)reqs.size();
            gr.n = (int32_t)n;
            gr.tmode = 1;
            gr.as_f64 = in0.as_f64;
            gr.strict = in0.cmp == OP_GT || in0.cmp == OP_LE;
            gr.up = in0.cmp == OP_GT || in0.cmp == OP_GE;
            std::vector<int64_t> cuts;                /* ascending, distinct */
            for (size_t j = 0; j < n; j++) {
                const tdd_pred_ins_t& in = v[i0 + j]->pred.ins[0];
                int64_t c;
                if (gr.as_f64) memcpy(&c, &in.cf, 8);
                else c = in.ci;
                cuts.push_back(c);
            }
            auto less = [&](int64_t a, int64_t b) {
                if (!gr.as_f64) return a < b;
                double x, y;
                memcpy(&x, &a, 8);
                memcpy(&y, &b, 8);
                return x < y;
            };
            std::sort(cuts.begin(), cuts.end(), less);
            cuts.erase(std::unique(cuts.begin(), cuts.end(),
                                   [&](int64_t a, int64_t b) { return !less(a, b) && !less(b, a); }),
                       cuts.end());
            gr.m = (int32_t)cuts.size();
            for (int k = 0; k < gr.m; k++) gr.cut[k] = cuts[(size_t)k];
            gr.words = (int32_t)((gr.m + 1) * sg * v[i0]->a.ncells);
            for (size_t j = 0; j < n; j++) {
                const tdd_pred_ins_t& in = v[i0 + j]->pred.ins[0];
                int64_t c;
                if (gr.as_f64) memcpy(&c, &in.cf, 8);
                else c = in.ci;
                const int ci = (int)(std::lower_bound(cuts.begin(), cuts.end(), c, less) - cuts.begin());
                SrvDirReq R = make_req(v[i0 + j]);
                const bool up = in.cmp == OP_GT || in.cmp == OP_GE;   /* passes above the cut */
                R.blo = (int16_t)(up ? ci + 1 : 0);
                R.bhi = (int16_t)(up ? gr.m : ci);
                R.lds = 0;
                reqs.push_back(R);
            }
            groups.push_back(gr);
        }
    }
    for (const tdd_srv_acc_t* A : plain) {
        SrvDirReq R = make_req(A);
        const int32_t words = (int32_t)(sg * R.ncells);
        SrvDirGroup* gp = groups.empty() || groups.back().tmode ? nullptr : &groups.back();
        if (!gp || gp->n == SRV_DIR_GRP || (int64_t)(gp->words + words) * 8 > SRV_DIR_LDS) {
            SrvDirGroup ng;
            memset(&ng, 0, sizeof(ng));
            ng.first = (int32_t)reqs.size();
            groups.push_back(ng);
            gp = &groups.back();
        }
        R.lds = gp->words;
        gp->words += words;
        gp->n++;
        reqs.push_back(R);
    }
    int32_t maxw = 0;
    for (const auto& g : groups) if (g.words > maxw) maxw = g.words;
    SCRATCH(dreq, sizeof(SrvDirReq) * reqs.size());
    SCRATCH(dgrp, sizeof(SrvDirGroup) * groups.size());
    HC(hipMemcpyAsync(dreq.p, reqs.data(), sizeof(SrvDirReq) * reqs.size(),
                      hipMemcpyHostToDevice, tdd::g_stream));
    HC(hipMemcpyAsync(dgrp.p, groups.data(), sizeof(SrvDirGroup) * groups.size(),
                      hipMemcpyHostToDevice, tdd::g_stream));
    SrvFoldArgs F;
    memset(&F, 0, sizeof(F));
    F.rg = rg.as<uint32_t>();
    F.off = off.as<int64_t>();
    F.ntiles = ntiles;
    F.G = G;
    F.sg = (int32_t)sg;
    F.ngroups = (int32_t)groups.size();
    F.reqs = dreq.as<SrvDirReq>();
    F.groups = dgrp.as<SrvDirGroup>();
    /* parts per slice: about 8 workgroups per CU in all, a part at least
     * 4,096 records */
    const int64_t units = nsl * F.ngroups, want = (int64_t)tdd::g_num_cu * 8;
    int64_t parts = units < want ? ceil_div(want, units) : 1;
    const int64_t most = ceil_div(ceil_div(rows, nsl), 4096);
    if (parts > most) parts = most > 0 ? most : 1;
    F.parts = (int32_t)parts;
    LAUNCH("serve_fold", k_srv_bfold, (int)(units * parts), 256, (size_t)maxw * 8, F);
    HC(hipStreamSynchronize(tdd::g_stream));         /* host argument arrays go out of scope */
    return TD_OK;
}

/* Direct requests that no slice plan covers (more groups than SRV_MAX_S
 * slices, or more columns than a record holds): every row of the chunk
 * folded into their tables with global atomics. */
static int srv_global(tdd_srv_t* s, const std::vector<const tdd_srv_acc_t*>& list, int64_t r0,
                      int64_t r1) {
    const tdd_srv_spec_t& S = s->spec;
    std::vector<SrvDirReq> reqs;
    std::vector<SrvDirGroup> groups;
    int vused = 0, pused = 0;
    for (const tdd_srv_acc_t* A : list) {
        SrvDirReq R;
        memset(&R, 0, sizeof(R));
        R.P = A->pred;
        R.tab = A->a.tab;
        R.G = A->a.G;
        R.key = A->a.key;
        R.ncells = A->a.ncells;
        R.lds = -1;
        memcpy(R.kind, A->a.kind, sizeof(R.kind));
        memcpy(R.src, A->a.src, sizeof(R.src));
        for (int c = 1; c < R.ncells; c++) {
            if (R.src[c] == -2) pused = 1;
            else if (R.src[c] >= 0) vused |= 1 << R.src[c];
        }
        groups.push_back({(int32_t)reqs.size(), 1, 0, 0});
        reqs.push_back(R);
    }
    SrvDirArgs D;
    memset(&D, 0, sizeof(D));
    D.fk = S.fk;
    if (pused) D.pcol = S.pcol;
    for (int j = 0; j < S.n_vals; j++) if ((vused >> j) & 1) D.vals[j] = S.vals[j];
    for (int k = 0; k < S.n_keys; k++) D.map[k] = s->map[k];
    D.r0 = r0;
    D.r1 = r1;
    D.kmin = s->kmin;
    D.range = s->range;
    D.ngroups = (int32_t)groups.size();
    const int64_t rows = r1 - r0;
    int64_t nb = (int64_t)tdd::g_num_cu * 4 / D.ngroups;
    if (nb < 1) nb = 1;
    if (nb > ceil_div(rows, 4096)) nb = ceil_div(rows, 4096);
    D.span = ceil_div(ceil_div(rows, nb), 1024) * 1024;
    nb = ceil_div(rows, D.span);
    if (nb * D.ngroups > ((int64_t)1 << 31) - 1) return TD_ERR_NYI;
    SCRATCH(dreq, sizeof(SrvDirReq) * reqs.size());
    SCRATCH(dgrp, sizeof(SrvDirGroup) * groups.size());
    HC(hipMemcpyAsync(dreq.p, reqs.data(), sizeof(SrvDirReq) * reqs.size(),
                      hipMemcpyHostToDevice, tdd::g_stream));
    HC(hipMemcpyAsync(dgrp.p, groups.data(), sizeof(SrvDirGroup) * groups.size(),
                      hipMemcpyHostToDevice, tdd::g_stream));
    D.reqs = dreq.as<SrvDirReq>();
    D.groups = dgrp.as<SrvDirGroup>();
    LAUNCH("serve_direct", k_srv_direct, (int)(nb * D.ngroups), 256, 0, D);
    HC(hipStreamSynchronize(tdd::g_stream));         /* host argument arrays go out of scope */
    return TD_OK;
}

/* The step's direct requests over chunk rows [r0, r1): per group key,
 * batches whose columns fit one record go through the bucket sort and the
 * slice folds; the rest through global atomics. */
static int srv_direct(tdd_srv_t* s, tdd_srv_acc_t* const* active, int n, int64_t r0,
                      int64_t r1) {
    if (r1 <= r0) return TD_OK;
    std::vector<const tdd_srv_acc_t*> glob;
    for (int k = 0; k < s->spec.n_keys; k++) {
        std::vector<const tdd_srv_acc_t*> batch;
        std::vector<tdd_col_t> cols;
        int ncmax = 1;
        auto flush = [&]() -> int {
            int rc = batch.empty() ? TD_OK : srv_bucket(s, k, batch, cols, ncmax, r0, r1);
            batch.clear();
            cols.clear();
            ncmax = 1;
            return rc;
        };
        for (int i = 0; i < n; i++) {
            const tdd_srv_acc_t* A = active[i];
            if (!A->direct || A->a.key != k) continue;
            tdd_col_t need[2 * TDD_PRED_MAX_COLS + GP_MAX_CELLS];
            const int nn = req_cols(s, A, need);
            if (nn < 0 || nn > SRV_REC_C ||
                ceil_div(s->G[k], slice_groups(A->a.ncells)) > SRV_MAX_S) {
                glob.push_back(A);
                continue;
            }
            std::vector<tdd_col_t> merged = cols;
            for (int j = 0; j < nn; j++) {
                bool have = false;
                for (const auto& c : merged) have = have || (c.data == need[j].data &&
                                                            c.type == need[j].type);
                if (!have) merged.push_back(need[j]);
            }
            const int nc = A->a.ncells > ncmax ? A->a.ncells : ncmax;
            if (merged.size() > SRV_REC_C || ceil_div(s->G[k], slice_groups(nc)) > SRV_MAX_S) {
                int rc = flush();                   /* this batch is full: start another */
                if (rc) return rc;
                merged.assign(need, need + nn);
            }
            cols = merged;
            batch.push_back(A);
            if (A->a.ncells > ncmax) ncmax = A->a.ncells;
        }
        int rc = flush();
        if (rc) return rc;
    }
    return glob.empty() ? TD_OK : srv_global(s, glob, r0, r1);
}

/* The step's raw cubes by global atomics (cubes the sliced form cannot take). */
static int srv_cube_atomic(tdd_srv_t* s, const SrvScanArgs& A) {
    for (int k = 0; k < A.nk; k++) {
        const SrvKeyArgs& K = A.key[k];
        if (!K.on) continue;
        const size_t bytes = (size_t)(K.G * s->Q) * 8;
        int rc = tdd_fill(K.cnt, 0, bytes);
        for (int j = 0; j < s->spec.n_vals && !rc; j++)
            if ((K.vmask >> j) & 1) rc = tdd_fill(K.val[j], 0, bytes);
        if (rc) return rc;
    }
    const int64_t tiles = ceil_div(A.r1 - A.r0, 64 * SRV_R);
    LAUNCH("serve_cube", k_srv_cube, resident_grid(k_srv_cube, 256, 0, ceil_div(tiles, 4)), 256,
           0, A);
    return TD_OK;
}

/* The step's cubes by slices, prefix / suffix sums included: TD_ERR_NYI
 * when the server has no slice arrays, one group's pcol values do not fit
 * an LDS table or the active cubes need more than SRV_CS_MAXS slices.
 * modes[k]: bit M_RAW / M_PFX / M_SFX for the cubes key k's requests read. */
static int srv_cube_sliced(tdd_srv_t* s, const SrvScanArgs& A, uint32_t ftype,
                           const int* modes) {
    const int64_t rows = A.r1 - A.r0;
    if (!s->cs_mem || rows > s->cs_rows || TUNE_INT("TEIDE_SRV_CUBE_ATOMIC", 0)) return TD_ERR_NYI;
    SrvCubePlan P;
    memset(&P, 0, sizeof(P));
    P.nk = A.nk;
    P.ftype = ftype;
    P.jmap = s->cs_jmap;
    P.js = s->cs_js;
    P.qlo = s->qlo;
    P.Q = s->Q;
    int ns = 0, slot = 0;
    size_t fold_lds = 8;
    /* LDS of a slice's table, beside the fold's 6 KB of segment tables */
    int64_t lds = (int64_t)TUNE_INT("TEIDE_SRV_CS_LDS_KB", 40) * 1024;
    if (lds > tdd::g_lds_block - 8192) lds = tdd::g_lds_block - 8192;
    for (int k = 0; k < A.nk; k++) {
        P.sbase[k] = ns;
        const SrvKeyArgs& K = A.key[k];
        if (!s->ok[k]) continue;
        const int kslot = slot++;                  /* key slots follow the served keys */
        if (!K.on) continue;
        int nw = 0;
        for (int j = 0; j < TDD_SRV_MAX_VALS; j++) nw += (K.vmask >> j) & 1;
        int64_t gw = lds / ((1 + nw) * 8) / s->Q;             /* groups of an LDS table */
        if (gw < 1) return TD_ERR_NYI;
        if (gw >= 32) gw &= ~(int64_t)15;                      /* whole lines per value row */
        const int64_t gpar = ceil_div(K.G, 256);              /* >= 256 slices where G allows */
        if (gw > gpar) gw = gpar;
        const int64_t nsl = ceil_div(K.G, gw);
        const int64_t R = ceil_div(nsl, SRV_CS_TS);            /* slices per run group */
        const int64_t nsk = ceil_div(nsl, R) * R;
        if (ns + nsk > SRV_CS_MAXS) return TD_ERR_NYI;
        ns += (int)nsk;
        P.on[k] = 1;
        P.vmask[k] = K.vmask;
        P.kslot[k] = kslot;
        P.modes[k] = modes[k];
        P.R[k] = (int32_t)R;
        P.gw[k] = (int32_t)gw;
        P.G[k] = K.G;
        P.cells[k] = s->cs_cells + kslot * s->cs_rows;
        P.cube[k] = s->cube[k];
        const size_t tab = (size_t)(gw * s->Q) * (1 + nw) * 8;
        if (tab > fold_lds) fold_lds = tab;
    }
    P.sbase[A.nk] = ns;
    if (ns == 0) return TD_OK;
    /* k_srv_cput's LDS: cursors, one key's tile offsets, 4,096 sorted records */
    int nskmax = 0;
    for (int k = 0; k < A.nk; k++)
        if (P.sbase[k + 1] - P.sbase[k] > nskmax) nskmax = P.sbase[k + 1] - P.sbase[k];
    const size_t put_lds = (size_t)ns * 4 + ((size_t)nskmax + 1) * 4 + 4096 * 4 + 8 + 4096 * 8;
    if (put_lds + 256 > (size_t)tdd::g_lds_block) return TD_ERR_NYI;
    const int64_t wgmax = TUNE_INT("TEIDE_SRV_CS_WG", 256);
    int64_t nwg = ceil_div(ceil_div(rows, 4096), 8) * 8;
    if (nwg > wgmax) nwg = wgmax;
    if (nwg > SRV_CS_WGMAX) nwg = SRV_CS_WGMAX;
    nwg = nwg < 8 ? 8 : nwg & ~(int64_t)7;
    P.ns = ns;
    P.nwg = (int32_t)nwg;
    P.span = ceil_div(ceil_div(rows, nwg), 4096) * 4096;
    P.C = s->cs_C;
    P.O = s->cs_O;
    P.roff = s->cs_roff;
    const int64_t nrec = s->cs_rows * s->cs_nk;
    for (int j = 0; j < TDD_SRV_MAX_VALS; j++)
        P.rval[j] = j < s->spec.n_vals ? s->cs_rval + j * nrec : nullptr;
    LAUNCH("serve_cube", k_srv_ccount, (int)nwg, 1024, (size_t)ns * 4, A, P);
    int rc = tdd::scan_u32_i64(s->cs_C, (int64_t)ns * nwg, s->cs_O);
    if (rc) return rc;
    LAUNCH("serve_cube", k_srv_cput, (int)nwg, 1024, put_lds, A, P, nskmax);
    LAUNCH("serve_cube", k_srv_cfold, ns, 512, fold_lds, P);
    return TD_OK;
}

extern "C" int tdd_srv_step(tdd_srv_t* s, tdd_srv_acc_t* const* active, int n) {
    const tdd_srv_spec_t& S = s->spec;
    const int nv = S.n_vals;
    int on[TDD_SRV_MAX_KEYS] = {0}, vm[TDD_SRV_MAX_KEYS] = {0};
    int pfx[TDD_SRV_MAX_KEYS] = {0}, sfx[TDD_SRV_MAX_KEYS] = {0}, modes[TDD_SRV_MAX_KEYS] = {0};
    int live = 0, direct = 0;
    for (int i = 0; i < n; i++) {
        const SrvAccArgs& a = active[i]->a;
        direct |= active[i]->direct;
        if (a.mode == M_NONE) continue;
        live++;
        on[a.key] = 1;
        pfx[a.key] |= a.mode == M_PFX;
        sfx[a.key] |= a.mode == M_SFX;
        modes[a.key] |= 1 << a.mode;
        for (int c = 1; c < a.ncells; c++) if (a.src[c] >= 0) vm[a.key] |= 1 << a.src[c];
    }
    const int64_t c = s->step % s->nchunks;
    s->step++;
    if (direct && S.nf > 0) {
        const int64_t r0 = c * s->chunk, r1 = r0 + s->chunk < S.nf ? r0 + s->chunk : S.nf;
        int rc = srv_direct(s, active, n, r0, r1);
        if (rc) return rc;
    }
    if (!live || S.nf == 0) return TD_OK;
    SrvScanArgs A;
    memset(&A, 0, sizeof(A));
    A.fk = S.fk;
    A.pcol = S.pcol;
    A.r0 = c * s->chunk;
    A.r1 = A.r0 + s->chunk < S.nf ? A.r0 + s->chunk : S.nf;
    A.kmin = s->kmin;
    A.range = s->range;
    A.qlo = s->qlo;
    A.Q = s->Q;
    A.nk = S.n_keys;
    A.vec = (tile_vec_ok(S.fk) ? 1u : 0u) | (tile_vec_ok(S.pcol) ? 2u : 0u);
    int vall = 0;
    for (int k = 0; k < S.n_keys; k++) vall |= on[k] ? vm[k] : 0;
    for (int j = 0; j < nv; j++) {
        if (!((vall >> j) & 1)) continue;          /* unused value columns stay unread */
        A.vals[j] = S.vals[j];
        if (tile_vec_ok(S.vals[j])) A.vec |= 1u << (2 + j);
    }
    uint32_t ftype = 0;
    for (int j = 0; j < nv; j++) if (S.vals[j].type == TD_F64) ftype |= 1u << j;
    for (int k = 0; k < S.n_keys; k++) {
        if (!on[k]) continue;
        SrvKeyArgs& K = A.key[k];
        K.map = s->map[k];
        K.cnt = s->cube[k].cnt[M_RAW];
        K.G = s->G[k];
        K.on = 1;
        K.vmask = vm[k];
        for (int j = 0; j < nv; j++) K.val[j] = s->cube[k].val[M_RAW][j];
    }
    int rc = srv_cube_sliced(s, A, ftype, modes);
    if (rc == TD_ERR_NYI) {
        rc = srv_cube_atomic(s, A);
        if (rc) return rc;
        for (int k = 0; k < S.n_keys; k++) {
            if (!on[k]) continue;
            LAUNCH("serve_cumul", k_srv_cumul, (int)ceil_div(s->G[k], 256), 256, 0, s->cube[k],
                   s->G[k], s->Q, s->qlo, nv, ftype, vm[k], pfx[k], sfx[k]);
        }
    } else if (rc) {
        return rc;
    }
    std::vector<SrvAccArgs> h((size_t)n);
    int64_t gmax = 1;
    for (int i = 0; i < n; i++) {
        h[i] = active[i]->a;
        if (h[i].G > gmax) gmax = h[i].G;
    }
    SCRATCH(d, sizeof(SrvAccArgs) * (size_t)n);
    HC(hipMemcpyAsync(d.p, h.data(), sizeof(SrvAccArgs) * (size_t)n, hipMemcpyHostToDevice,
                      tdd::g_stream));
    for (int i0 = 0; i0 < n; i0 += SRV_MAX_Y) {
        const int m = n - i0 < SRV_MAX_Y ? n - i0 : SRV_MAX_Y;
        dim3 grid((unsigned)ceil_div(gmax, 256), (unsigned)m);
        LAUNCH("serve_accum", k_srv_accum, grid, 256, 0, d.as<SrvAccArgs>() + i0,
               (const SrvCube*)s->d_cubes, s->qlo);
    }
    HC(hipStreamSynchronize(tdd::g_stream));      /* step done: requests may finish */
    return TD_OK;
}

extern "C" int tdd_srv_finish(tdd_srv_t* s, tdd_srv_acc_t* const* acc, int n,
                              tdd_group_result_t* out) {
    (void)s;
    if (n <= 0) return TD_OK;
    std::vector<SrvEmitArgs> h((size_t)n);
    int64_t gmax = 1;
    for (int i = 0; i < n; i++) {
        memset(&out[i], 0, sizeof(out[i]));
        h[i].P = acc[i]->P;
        h[i].tab = acc[i]->a.tab;
        h[i].G = acc[i]->a.G;
        memset(&h[i].out, 0, sizeof(h[i].out));
        if (h[i].G > gmax) gmax = h[i].G;
    }
    const int64_t cells = (int64_t)n * gmax;
    SCRATCH(d, sizeof(SrvEmitArgs) * (size_t)n);
    SCRATCH(flag, (size_t)cells * 4);
    SCRATCH(pos, (size_t)(cells + 1) * 8);
    SCRATCH(first, (size_t)(n + 1) * 8);
    HC(hipMemcpyAsync(d.p, h.data(), sizeof(SrvEmitArgs) * (size_t)n, hipMemcpyHostToDevice,
                      tdd::g_stream));
    for (int i0 = 0; i0 < n; i0 += SRV_MAX_Y) {
        const int m = n - i0 < SRV_MAX_Y ? n - i0 : SRV_MAX_Y;
        dim3 grid((unsigned)ceil_div(gmax, 256), (unsigned)m);
        LAUNCH("serve_emit", k_srv_flags, grid, 256, 0, d.as<SrvEmitArgs>() + i0, gmax,
               flag.as<uint32_t>() + (int64_t)i0 * gmax);
    }
    int rc = tdd::scan_u32_i64(flag.as<uint32_t>(), cells, pos.as<int64_t>());
    if (rc) return rc;
    LAUNCH("serve_emit", k_srv_firsts, (int)ceil_div(n + 1, 256), 256, 0,
           (const int64_t*)pos.as<int64_t>(), gmax, n, first.as<int64_t>());
    std::vector<int64_t> f((size_t)n + 1);
    HC(hipMemcpyAsync(f.data(), first.p, (size_t)(n + 1) * 8, hipMemcpyDeviceToHost,
                      tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    for (int i = 0; i < n; i++) {
        const int64_t ng = f[i + 1] - f[i];
        rc = alloc_group_outputs(h[i].P, ng, &out[i]);
        if (rc) return rc;
        out[i].n_groups = ng;
        out[i].path = 1;
        for (int a = 0; a < h[i].P.n_aggs; a++) out[i].out_types[a] = h[i].P.out_types[a];
        h[i].out = out[i];
    }
    HC(hipMemcpyAsync(d.p, h.data(), sizeof(SrvEmitArgs) * (size_t)n, hipMemcpyHostToDevice,
                      tdd::g_stream));
    for (int i0 = 0; i0 < n; i0 += SRV_MAX_Y) {
        const int m = n - i0 < SRV_MAX_Y ? n - i0 : SRV_MAX_Y;
        dim3 grid((unsigned)ceil_div(gmax, 256), (unsigned)m);
        LAUNCH("serve_emit", k_srv_emit, grid, 256, 0, d.as<SrvEmitArgs>() + i0, gmax,
               (const int64_t*)pos.as<int64_t>() + (int64_t)i0 * gmax);
    }
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}
