/*
 * orc.c -- the oracle (see orc.h).  TEST INFRASTRUCTURE ONLY: a slow,
 * sequential restatement of the reference's semantics, written for clarity
 * rather than speed.  Never linked into the product.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* td.h tags and opcodes used here (values fixed by the ABI). */
enum { T_BOOL = 1, T_U8 = 2, T_I16 = 4, T_I32 = 5, T_I64 = 6, T_F64 = 7,
       T_DATE = 9, T_TIME = 10, T_TIMESTAMP = 11, T_SYM = 20 };
enum { O_EQ = 25, O_NE, O_LT, O_LE, O_GT, O_GE, O_AND, O_OR, O_NOT = 12 };
enum { O_SUM = 50, O_PROD, O_MIN, O_MAX, O_COUNT, O_AVG, O_FIRST, O_LAST,
       O_STDDEV = 59, O_STDDEV_POP = 73, O_VAR, O_VAR_POP };

#define SEG 1024

/* read_col_i64 (exec.c:90-109): every non-F64 type widened to int64. */
static int64_t rd_i64(orc_col_t c, int64_t r) {
    switch (c.type) {
    case T_I64: case T_TIMESTAMP: return ((const int64_t*)c.data)[r];
    case T_SYM:
        switch (c.attrs & 3) {
        case 0: return ((const uint8_t*)c.data)[r];
        case 1: return ((const uint16_t*)c.data)[r];
        case 2: return ((const uint32_t*)c.data)[r];
        default: return ((const int64_t*)c.data)[r];
        }
    case T_I32: case T_DATE: case T_TIME: return ((const int32_t*)c.data)[r];
    case T_I16: return ((const int16_t*)c.data)[r];
    default: return ((const uint8_t*)c.data)[r];
    }
}

static double rd_f64(orc_col_t c, int64_t r) {
    if (c.type == T_F64) return ((const double*)c.data)[r];
    return (double)rd_i64(c, r);
}

/* ------------------------------------------------------------ selection -- */

static int64_t al8(int64_t x) { return (x + 7) & ~(int64_t)7; }

int64_t orc_sel_payload_bytes(int64_t nrows) {
    int64_t nseg = (nrows + SEG - 1) / SEG, nw = (nrows + 63) / 64;
    return 16 + al8(nseg) + al8(nseg * 2) + nw * 8;
}

typedef struct { int64_t* total; uint32_t* nseg; uint8_t* flags; uint16_t* pop; uint64_t* bits; } selv_t;

static selv_t sel_view(uint8_t* p, int64_t nrows) {
    selv_t v;
    int64_t nseg = (nrows + SEG - 1) / SEG;
    v.total = (int64_t*)p;
    v.nseg = (uint32_t*)(p + 8);
    v.flags = p + 16;
    v.pop = (uint16_t*)(p + 16 + al8(nseg));
    v.bits = (uint64_t*)(p + 16 + al8(nseg) + al8(nseg * 2));
    return v;
}

/* td_sel_recompute (sel.c:77-122): per-segment popcount, NONE/ALL/MIX. */
static void sel_recompute(uint8_t* p, int64_t nrows) {
    selv_t v = sel_view(p, nrows);
    int64_t nseg = (nrows + SEG - 1) / SEG, total = 0;
    *v.nseg = (uint32_t)nseg;
    for (int64_t s = 0; s < nseg; s++) {
        int64_t lo = s * SEG, hi = lo + SEG < nrows ? lo + SEG : nrows, pc = 0;
        for (int64_t r = lo; r < hi; r++) pc += (v.bits[r >> 6] >> (r & 63)) & 1;
        v.pop[s] = (uint16_t)pc;
        v.flags[s] = pc == 0 ? 0 : (pc == hi - lo ? 1 : 2);
        total += pc;
    }
    *v.total = total;
}

void orc_sel_from_bool(const uint8_t* pred, int64_t nrows, uint8_t* payload) {
    memset(payload, 0, (size_t)orc_sel_payload_bytes(nrows));
    selv_t v = sel_view(payload, nrows);
    for (int64_t r = 0; r < nrows; r++)
        if (pred[r]) v.bits[r >> 6] |= 1ull << (r & 63);
    sel_recompute(payload, nrows);
}

void orc_sel_and(const uint8_t* a, const uint8_t* b, int64_t nrows, uint8_t* out) {
    memset(out, 0, (size_t)orc_sel_payload_bytes(nrows));
    selv_t va = sel_view((uint8_t*)a, nrows), vb = sel_view((uint8_t*)b, nrows);
    selv_t vo = sel_view(out, nrows);
    for (int64_t w = 0; w < (nrows + 63) / 64; w++) vo.bits[w] = va.bits[w] & vb.bits[w];
    sel_recompute(out, nrows);
}

int64_t orc_sel_to_idx(const uint8_t* payload, int64_t nrows, int64_t* idx) {
    selv_t v = sel_view((uint8_t*)payload, nrows);
    int64_t j = 0;
    for (int64_t r = 0; r < nrows; r++)
        if ((v.bits[r >> 6] >> (r & 63)) & 1) idx[j++] = r;
    return j;
}

/* ------------------------------------------------------------ predicates -- */

static int cmp_i(int op, int64_t a, int64_t b) {
    switch (op) {
    case O_EQ: return a == b; case O_NE: return a != b; case O_LT: return a < b;
    case O_LE: return a <= b; case O_GT: return a > b; default: return a >= b;
    }
}
static int cmp_f(int op, double a, double b) {
    switch (op) {
    case O_EQ: return a == b; case O_NE: return a != b; case O_LT: return a < b;
    case O_LE: return a <= b; case O_GT: return a > b; default: return a >= b;
    }
}

void orc_cmp_const(orc_col_t col, int64_t n, int op, int const_is_f64,
                   int64_t ci, double cf, uint8_t* out) {
    int as_f = col.type == T_F64 || const_is_f64;   /* mixed -> f64 (exec.c:865-873) */
    for (int64_t r = 0; r < n; r++)
        out[r] = as_f ? (uint8_t)cmp_f(op, rd_f64(col, r), const_is_f64 ? cf : (double)ci)
                      : (uint8_t)cmp_i(op, rd_i64(col, r), ci);
}

void orc_cmp_cols(orc_col_t a, orc_col_t b, int64_t n, int op, uint8_t* out) {
    int as_f = a.type == T_F64 || b.type == T_F64;
    for (int64_t r = 0; r < n; r++)
        out[r] = as_f ? (uint8_t)cmp_f(op, rd_f64(a, r), rd_f64(b, r))
                      : (uint8_t)cmp_i(op, rd_i64(a, r), rd_i64(b, r));
}

void orc_bool_op(int op, const uint8_t* a, const uint8_t* b, int64_t n, uint8_t* out) {
    for (int64_t r = 0; r < n; r++)
        out[r] = op == O_AND ? (a[r] && b[r]) : op == O_OR ? (a[r] || b[r]) : !a[r];
}

void orc_gather(const void* src, int esz, const int64_t* idx, int64_t m,
                void* dst, int nullable) {
    for (int64_t i = 0; i < m; i++) {
        if (idx[i] < 0 && nullable) memset((char*)dst + i * esz, 0, (size_t)esz);
        else memcpy((char*)dst + i * esz, (const char*)src + idx[i] * esz, (size_t)esz);
    }
}

/* ------------------------------------------------------- key hash tables -- */

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Key word of row r: F64 normalised so that -0.0 and 0.0 hash alike (they
 * compare equal); everything else as its int64 value. */
static uint64_t key_word(orc_col_t c, int64_t r) {
    if (c.type == T_F64) {
        double d = ((const double*)c.data)[r];
        if (d == 0.0) d = 0.0;
        uint64_t u;
        memcpy(&u, &d, 8);
        return u;
    }
    return (uint64_t)rd_i64(c, r);
}

static int keys_eq(const orc_col_t* a, int64_t ra, const orc_col_t* b, int64_t rb, int nk) {
    for (int k = 0; k < nk; k++) {
        if (!a[k].data || !b[k].data) return 0;
        if (a[k].type == T_F64) {
            if (((const double*)a[k].data)[ra] != ((const double*)b[k].data)[rb]) return 0;
        } else if (rd_i64(a[k], ra) != rd_i64(b[k], rb)) {
            return 0;
        }
    }
    return 1;
}

static uint64_t row_hash(const orc_col_t* c, int nk, int64_t r) {
    uint64_t h = 0x12345;
    for (int k = 0; k < nk; k++) h = mix64(h ^ (c[k].data ? key_word(c[k], r) : 0));
    return h;
}

/* ------------------------------------------------------------------ join -- */

int64_t orc_join(const orc_col_t* lkeys, const orc_col_t* rkeys, int nk,
                 int64_t nl, int64_t nr, int join_type,
                 int64_t* l_idx, int64_t* r_idx, int64_t cap) {
    for (int k = 0; k < nk; k++)
        if (lkeys[k].data && rkeys[k].data &&
            ((lkeys[k].type == T_F64) != (rkeys[k].type == T_F64)))
            return -1;                                  /* mixed f64/int keys */
    /* Chained table over right rows; chains kept in ascending row order. */
    int64_t cap_ht = 16;
    while (cap_ht < 2 * nr) cap_ht *= 2;
    int64_t* head = malloc((size_t)cap_ht * sizeof(int64_t));
    int64_t* tail = malloc((size_t)cap_ht * sizeof(int64_t));
    int64_t* next = malloc((size_t)(nr > 0 ? nr : 1) * sizeof(int64_t));
    uint8_t* matched = calloc((size_t)(nr > 0 ? nr : 1), 1);
    for (int64_t i = 0; i < cap_ht; i++) head[i] = tail[i] = -1;
    for (int64_t r = 0; r < nr; r++) {
        uint64_t b = row_hash(rkeys, nk, r) & (uint64_t)(cap_ht - 1);
        next[r] = -1;
        if (tail[b] < 0) head[b] = r; else next[tail[b]] = r;
        tail[b] = r;
    }
    int64_t cnt = 0;
#define EMIT(L, R) do { if (cnt < cap) { l_idx[cnt] = (L); r_idx[cnt] = (R); } cnt++; } while (0)
    for (int64_t l = 0; l < nl; l++) {
        uint64_t b = row_hash(lkeys, nk, l) & (uint64_t)(cap_ht - 1);
        int any = 0;
        for (int64_t r = head[b]; r >= 0; r = next[r])
            if (keys_eq(lkeys, l, rkeys, r, nk)) { EMIT(l, r); matched[r] = 1; any = 1; }
        if (!any && join_type >= 1) EMIT(l, -1);
    }
    if (join_type == 2)
        for (int64_t r = 0; r < nr; r++)
            if (!matched[r]) EMIT(-1, r);
#undef EMIT
    free(head); free(tail); free(next); free(matched);
    return cnt;
}

/* ---------------------------------------------------------------- group -- */

typedef struct {
    int64_t  count;
    int64_t  first_row, last_row;
    double   fsum[8], fcomp[8], fcc[8], fabs_[8], fsq[8], fmin[8], fmax[8];
    int64_t  isum[8], imin[8], imax[8];
} gacc_t;

static int is_int_key(int8_t t) {
    return t == T_I64 || t == T_SYM || t == T_I32 || t == T_TIMESTAMP || t == T_DATE ||
           t == T_TIME || t == T_BOOL || t == T_U8 || t == T_I16;
}

static int64_t* g_sort_keys;
static int g_sort_nk;
static int64_t g_sort_stride;
static int cmp_group(const void* pa, const void* pb) {
    int64_t a = *(const int64_t*)pa, b = *(const int64_t*)pb;
    for (int k = 0; k < g_sort_nk; k++) {
        int64_t x = g_sort_keys[k * g_sort_stride + a], y = g_sort_keys[k * g_sort_stride + b];
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

/* Reference direct-array eligibility (exec.c:5951-6036): integer keys, the
 * product of masked key ranges <= 262144 and one worker's accumulator arrays
 * <= 6 MB (MIN/MAX weighted twice). */
static int da_eligible(const orc_col_t* keys, int nk, const uint16_t* ops, int na,
                       int64_t nrows, const uint64_t* mask) {
    if (nrows <= 0 || nk < 1 || nk > 8) return 0;
    uint64_t slots = 1;
    for (int k = 0; k < nk; k++) {
        if (!keys[k].data || !is_int_key(keys[k].type)) return 0;
        int64_t mn = INT64_MAX, mx = INT64_MIN;
        for (int64_t r = 0; r < nrows; r++) {
            if (mask && !((mask[r >> 6] >> (r & 63)) & 1)) continue;
            int64_t v = rd_i64(keys[k], r);
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        int64_t range = (int64_t)((uint64_t)mx - (uint64_t)mn + 1u);
        if (range <= 0) return 0;
        slots *= (uint64_t)range;
        if (slots > 262144) return 0;
    }
    int sum = 0, mn = 0, mx = 0, sq = 0;
    for (int a = 0; a < na; a++) {
        uint16_t op = ops[a];
        if (op == O_SUM || op == O_AVG || op == O_FIRST || op == O_LAST) sum = 1;
        else if (op == O_STDDEV || op == O_STDDEV_POP || op == O_VAR || op == O_VAR_POP) sum = sq = 1;
        else if (op == O_MIN) mn = 1;
        else if (op == O_MAX) mx = 1;
    }
    uint64_t per = (uint64_t)(sum + 2 * mn + 2 * mx + sq);
    return slots * (per * (uint64_t)na + 1) * 8 <= (6ull << 20);
}

int64_t orc_group(const orc_col_t* keys, int nk, const uint16_t* ops,
                  const orc_col_t* aggs, int na, int64_t nrows,
                  const uint64_t* mask, int64_t cap, orc_group_out_t* out) {
    if (nk < 1 || nk > 8 || na > 8) return -1;
    for (int a = 0; a < na; a++)
        if (ops[a] == O_PROD) return -1;
    for (int a = 0; a < na; a++) {                   /* output types, exec.c:4645-4656 */
        int f = aggs[a].data && aggs[a].type == T_F64;
        switch (ops[a]) {
        case O_AVG: case O_STDDEV: case O_STDDEV_POP: case O_VAR: case O_VAR_POP:
            out->out_types[a] = T_F64; break;
        case O_COUNT: out->out_types[a] = T_I64; break;
        case O_SUM: out->out_types[a] = f ? T_F64 : T_I64; break;
        default: out->out_types[a] = aggs[a].data ? aggs[a].type : T_I64; break;
        }
    }
    int64_t hcap = 16;
    while (hcap < 2 * (nrows + 1)) hcap *= 2;
    int64_t* ht = malloc((size_t)hcap * sizeof(int64_t));
    for (int64_t i = 0; i < hcap; i++) ht[i] = -1;
    int64_t gcap = 1024, ng = 0;
    int64_t* grow = malloc((size_t)gcap * sizeof(int64_t));      /* first row */
    gacc_t* acc = malloc((size_t)gcap * sizeof(gacc_t));
    for (int64_t r = 0; r < nrows; r++) {
        if (mask && !((mask[r >> 6] >> (r & 63)) & 1)) continue;
        uint64_t b = row_hash(keys, nk, r) & (uint64_t)(hcap - 1);
        int64_t g = -1;
        for (;;) {
            int64_t e = ht[b];
            if (e < 0) break;
            /* group keys compare by raw 8-byte value (exec.c:4162) */
            int same = 1;
            for (int k = 0; k < nk && same; k++) {
                uint64_t x, y;
                if (keys[k].type == T_F64) {
                    memcpy(&x, (const double*)keys[k].data + grow[e], 8);
                    memcpy(&y, (const double*)keys[k].data + r, 8);
                } else {
                    x = (uint64_t)rd_i64(keys[k], grow[e]);
                    y = (uint64_t)rd_i64(keys[k], r);
                }
                same = x == y;
            }
            if (same) { g = e; break; }
            b = (b + 1) & (uint64_t)(hcap - 1);
        }
        if (g < 0) {
            if (ng == gcap) {
                gcap *= 2;
                grow = realloc(grow, (size_t)gcap * sizeof(int64_t));
                acc = realloc(acc, (size_t)gcap * sizeof(gacc_t));
            }
            g = ng++;
            ht[b] = g;
            grow[g] = r;
            gacc_t* z = &acc[g];
            memset(z, 0, sizeof(*z));
            z->first_row = r;
            for (int a = 0; a < na; a++) {
                z->fmin[a] = DBL_MAX; z->fmax[a] = -DBL_MAX;
                z->imin[a] = INT64_MAX; z->imax[a] = INT64_MIN;
            }
        }
        gacc_t* z = &acc[g];
        z->count++;
        z->last_row = r;
        for (int a = 0; a < na; a++) {
            if (!aggs[a].data) continue;
            if (aggs[a].type == T_F64) {
                double v = ((const double*)aggs[a].data)[r];
                z->fsum[a] += v;
                double t = z->fcomp[a] + v;           /* Neumaier compensation */
                if (fabs(z->fcomp[a]) >= fabs(v)) z->fcc[a] += (z->fcomp[a] - t) + v;
                else z->fcc[a] += (v - t) + z->fcomp[a];
                z->fcomp[a] = t;
                z->fabs_[a] += fabs(v);
                z->fsq[a] += v * v;
                if (v < z->fmin[a]) z->fmin[a] = v;
                if (v > z->fmax[a]) z->fmax[a] = v;
            } else {
                int64_t v = rd_i64(aggs[a], r);
                z->isum[a] = (int64_t)((uint64_t)z->isum[a] + (uint64_t)v);
                z->fabs_[a] += fabs((double)v);
                z->fsq[a] += (double)v * (double)v;
                if (v < z->imin[a]) z->imin[a] = v;
                if (v > z->imax[a]) z->imax[a] = v;
            }
        }
    }
    free(ht);
    int64_t* order = malloc((size_t)(ng > 0 ? ng : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < ng; i++) order[i] = i;
    out->order = da_eligible(keys, nk, ops, na, nrows, mask);
    int64_t* kv = malloc((size_t)(ng > 0 ? ng : 1) * (size_t)nk * sizeof(int64_t));
    for (int k = 0; k < nk; k++)
        for (int64_t g = 0; g < ng; g++)
            kv[k * ng + g] = (int64_t)key_word(keys[k], grow[g]);
    if (out->order) {
        g_sort_keys = kv; g_sort_nk = nk; g_sort_stride = ng;
        qsort(order, (size_t)ng, sizeof(int64_t), cmp_group);
    }
    for (int64_t i = 0; i < ng && i < cap; i++) {
        int64_t g = order[i];
        gacc_t* z = &acc[g];
        for (int k = 0; k < nk; k++) {
            if (keys[k].type == T_F64) {
                uint64_t u;
                memcpy(&u, (const double*)keys[k].data + grow[g], 8);
                out->keys[k * cap + i] = (int64_t)u;
            } else {
                out->keys[k * cap + i] = rd_i64(keys[k], grow[g]);
            }
        }
        for (int a = 0; a < na; a++) {
            int f = aggs[a].data && aggs[a].type == T_F64;
            double cnt = (double)z->count;
            double sumd = f ? z->fsum[a] : (double)z->isum[a];
            double dv = 0.0;
            int64_t iv = 0;
            int isf = out->out_types[a] == T_F64;
            switch (ops[a]) {
            case O_SUM: if (f) dv = z->fsum[a]; else iv = z->isum[a]; break;
            case O_COUNT: iv = z->count; break;
            case O_AVG: dv = sumd / cnt; break;
            case O_MIN: if (f) dv = z->fmin[a]; else iv = z->imin[a]; break;
            case O_MAX: if (f) dv = z->fmax[a]; else iv = z->imax[a]; break;
            case O_FIRST: case O_LAST: {
                int64_t rr = ops[a] == O_FIRST ? z->first_row : z->last_row;
                if (f) dv = ((const double*)aggs[a].data)[rr]; else iv = rd_i64(aggs[a], rr);
                break;
            }
            default: {                                /* VAR family, exec.c:4691-4701 */
                double mean = z->count > 0 ? sumd / cnt : 0.0;
                double vp = z->count > 0 ? z->fsq[a] / cnt - mean * mean : 0.0;
                if (vp < 0) vp = 0;
                if (ops[a] == O_VAR_POP) dv = z->count > 0 ? vp : NAN;
                else if (ops[a] == O_VAR) dv = z->count > 1 ? vp * cnt / (cnt - 1) : NAN;
                else if (ops[a] == O_STDDEV_POP) dv = z->count > 0 ? sqrt(vp) : NAN;
                else dv = z->count > 1 ? sqrt(vp * cnt / (cnt - 1)) : NAN;
            }
            }
            if (isf) memcpy(&out->vals[a * cap + i], &dv, 8);
            else out->vals[a * cap + i] = (uint64_t)iv;
            if (out->sum_abs) out->sum_abs[a * cap + i] = z->fabs_[a];
            if (out->comp) {
                double c = f ? z->fcomp[a] + z->fcc[a] : (double)z->isum[a];
                if (ops[a] == O_AVG) c /= cnt;
                out->comp[a * cap + i] = c;
            }
        }
    }
    free(kv); free(order); free(grow); free(acc);
    return ng;
}
