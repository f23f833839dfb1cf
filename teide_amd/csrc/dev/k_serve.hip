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
 *   k_srv_map      dim: direct map pk -> group of each key (once, at open)
 *   k_srv_cube     fact chunk: map gather + cube atomics per row and key
 *   k_srv_cumul    per (key, g): prefix and suffix sums over v
 *   k_srv_accum    per (request, g): table += the request's cube cell
 *   k_srv_flags / k_srv_emit   finished requests' groups, ascending keys
 *
 * Per row of a step the scan reads fk, pcol and the value columns once
 * (3 x 8 bytes for the fact layout fk i64, qty i64, price f64) for every
 * request in the mix, so the HBM cost of the scan is shared; the atomics
 * (1 + value columns per key) bound it, not the bytes.
 */
#include <vector>

#include "dev_common.h"
#include "group_plan.h"
#include "tile.h"

#define SRV_MAX_Q     4096                    /* pcol domain */
#define SRV_MAX_CELLS ((int64_t)1 << 26)      /* Q x G per key */
#define SRV_R         4                       /* rows per lane of a scan tile */
#define SRV_MAX_Y     65535                   /* grid.y limit */

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
};

struct tdd_srv_acc {
    GroupPlan     P;
    tdd_srv_req_t req;
    SrvAccArgs    a;
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
    if (r->cmp != OP_EQ && r->cmp != OP_LT && r->cmp != OP_LE && r->cmp != OP_GT && r->cmp != OP_GE)
        return TD_ERR_NYI;
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
    req_cell(s, r->cmp, r->t, &a.mode, &a.idx);
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

extern "C" int tdd_srv_step(tdd_srv_t* s, tdd_srv_acc_t* const* active, int n) {
    const tdd_srv_spec_t& S = s->spec;
    const int nv = S.n_vals;
    int on[TDD_SRV_MAX_KEYS] = {0}, vm[TDD_SRV_MAX_KEYS] = {0};
    int pfx[TDD_SRV_MAX_KEYS] = {0}, sfx[TDD_SRV_MAX_KEYS] = {0};
    int live = 0;
    for (int i = 0; i < n; i++) {
        const SrvAccArgs& a = active[i]->a;
        if (a.mode == M_NONE) continue;
        live++;
        on[a.key] = 1;
        pfx[a.key] |= a.mode == M_PFX;
        sfx[a.key] |= a.mode == M_SFX;
        for (int c = 1; c < a.ncells; c++) if (a.src[c] >= 0) vm[a.key] |= 1 << a.src[c];
    }
    const int64_t c = s->step % s->nchunks;
    s->step++;
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
        const int64_t QG = s->G[k] * s->Q;
        SrvKeyArgs& K = A.key[k];
        K.map = s->map[k];
        K.cnt = s->cube[k].cnt[M_RAW];
        K.G = s->G[k];
        K.on = 1;
        K.vmask = vm[k];
        int rc = tdd_fill(K.cnt, 0, (size_t)QG * 8);
        for (int j = 0; j < nv && !rc; j++) {
            K.val[j] = s->cube[k].val[M_RAW][j];
            if ((vm[k] >> j) & 1) rc = tdd_fill(K.val[j], 0, (size_t)QG * 8);
        }
        if (rc) return rc;
    }
    const int64_t tiles = ceil_div(A.r1 - A.r0, 64 * SRV_R);
    LAUNCH("serve_cube", k_srv_cube, resident_grid(k_srv_cube, 256, 0, ceil_div(tiles, 4)), 256,
           0, A);
    for (int k = 0; k < S.n_keys; k++) {
        if (!on[k]) continue;
        LAUNCH("serve_cumul", k_srv_cumul, (int)ceil_div(s->G[k], 256), 256, 0, s->cube[k],
               s->G[k], s->Q, s->qlo, nv, ftype, vm[k], pfx[k], sfx[k]);
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
