/*
 * arena.c -- the HBM arena: a buddy allocator over device memory, plus the
 * host proxies (mmod 3) that give device-resident vectors a td_t identity.
 *
 * Device memory is reserved in 1 GiB chunks and split buddy-style down to
 * 256-byte blocks; the bookkeeping lives on the host (per chunk: one state
 * byte per 256-byte unit and one free bitmap per order), since the device
 * memory itself is not host-addressable.  Requests above 512 MiB get their
 * own allocation, and freed large allocations are kept for reuse, because
 * allocating and releasing device memory synchronises the device.
 *
 * The reference's heap (src/mem/heap.c) has the same shape on the host side:
 * buddy blocks inside fixed pools, direct mappings for very large blocks.
 */
#include "internal.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>

#define A_MIN_ORDER   8                        /* 256 B */
#define A_CHUNK_ORDER 30                       /* 1 GiB */
#define A_CHUNK       ((size_t)1 << A_CHUNK_ORDER)
#define A_UNITS       (A_CHUNK >> A_MIN_ORDER)
#define A_NORDERS     (A_CHUNK_ORDER - A_MIN_ORDER + 1)
#define A_MAX_CHUNKS  512
#define A_MAX_DIRECT  1024

typedef struct {
    char*     base;                            /* device address */
    uint8_t*  state;                           /* per unit: 0, or order+1 of an allocated block */
    uint64_t* freebits[A_NORDERS];             /* bit i: block i of that order is free */
    int64_t   nfree[A_NORDERS];
} chunk_t;

typedef struct { char* p; size_t bytes; int live; } direct_t;

static pthread_mutex_t g_amu = PTHREAD_MUTEX_INITIALIZER;
static chunk_t  g_chunks[A_MAX_CHUNKS];
static int      g_nchunks;
static direct_t g_direct[A_MAX_DIRECT];
static size_t   g_in_use, g_peak, g_reserved, g_budget;
static int      g_dev_state;                   /* 0 unknown, 1 ready, -1 unavailable */
static size_t   g_fail_min;                    /* test hook: requests of at least this size */
static int64_t  g_fail_left;                   /* ... fail, this many more times */

void td_dev_fail_allocs(uint64_t min_bytes, int64_t count) {
    pthread_mutex_lock(&g_amu);
    g_fail_min = (size_t)min_bytes;
    g_fail_left = count > 0 ? count : 0;
    pthread_mutex_unlock(&g_amu);
}

static uint64_t* fb(chunk_t* c, int order) { return c->freebits[order - A_MIN_ORDER]; }

static void set_free(chunk_t* c, int order, size_t idx, int on) {
    uint64_t* b = fb(c, order);
    if (on) { b[idx >> 6] |= 1ull << (idx & 63); c->nfree[order - A_MIN_ORDER]++; }
    else { b[idx >> 6] &= ~(1ull << (idx & 63)); c->nfree[order - A_MIN_ORDER]--; }
}

static int is_free(chunk_t* c, int order, size_t idx) {
    return (fb(c, order)[idx >> 6] >> (idx & 63)) & 1;
}

static int64_t find_free(chunk_t* c, int order) {
    if (!c->nfree[order - A_MIN_ORDER]) return -1;
    size_t nblk = A_CHUNK >> order, nw = (nblk + 63) / 64;
    uint64_t* b = fb(c, order);
    for (size_t w = 0; w < nw; w++)
        if (b[w]) return (int64_t)(w * 64 + (size_t)__builtin_ctzll(b[w]));
    return -1;
}

int td__dev_ready(void) {
    if (g_dev_state) return g_dev_state > 0;
    g_dev_state = tdd_init() == TD_OK ? 1 : -1;
    if (g_dev_state > 0 && !g_budget) {
        uint64_t f = 0, t = 0;
        const char* env = getenv("TEIDE_HBM_BUDGET_GB");
        if (env && atof(env) > 0) g_budget = (size_t)(atof(env) * (double)(1ull << 30));
        else if (tdd_mem_info(&f, &t) == TD_OK) g_budget = (size_t)((double)t * 0.9);
    }
    return g_dev_state > 0;
}

static int add_chunk(void) {
    if (g_nchunks >= A_MAX_CHUNKS || g_reserved + A_CHUNK > g_budget) return 0;
    void* d = NULL;
    if (tdd_malloc(&d, A_CHUNK) != TD_OK) return 0;
    chunk_t* c = &g_chunks[g_nchunks];
    memset(c, 0, sizeof(*c));
    c->base = d;
    c->state = calloc(A_UNITS, 1);
    for (int o = A_MIN_ORDER; o <= A_CHUNK_ORDER; o++)
        c->freebits[o - A_MIN_ORDER] = calloc(((A_CHUNK >> o) + 63) / 64, 8);
    g_nchunks++;
    g_reserved += A_CHUNK;
    set_free(c, A_CHUNK_ORDER, 0, 1);
    return 1;
}

static void* chunk_alloc(int order) {
    for (int pass = 0; pass < 2; pass++) {
        for (int ci = 0; ci < g_nchunks; ci++) {
            chunk_t* c = &g_chunks[ci];
            for (int o = order; o <= A_CHUNK_ORDER; o++) {
                int64_t idx = find_free(c, o);
                if (idx < 0) continue;
                set_free(c, o, (size_t)idx, 0);
                size_t off = (size_t)idx << o;
                while (o > order) {                       /* split, keep the low half */
                    o--;
                    set_free(c, o, (off >> o) + 1, 1);
                }
                c->state[off >> A_MIN_ORDER] = (uint8_t)(order + 1);
                return c->base + off;
            }
        }
        if (pass == 0 && !add_chunk()) break;
    }
    return NULL;
}

static int chunk_free(void* p) {
    for (int ci = 0; ci < g_nchunks; ci++) {
        chunk_t* c = &g_chunks[ci];
        if ((char*)p < c->base || (char*)p >= c->base + A_CHUNK) continue;
        size_t off = (size_t)((char*)p - c->base);
        int order = c->state[off >> A_MIN_ORDER] - 1;
        if (order < A_MIN_ORDER) return -1;
        c->state[off >> A_MIN_ORDER] = 0;
        g_in_use -= (size_t)1 << order;
        size_t idx = off >> order;
        while (order < A_CHUNK_ORDER && is_free(c, order, idx ^ 1)) {   /* coalesce */
            set_free(c, order, idx ^ 1, 0);
            idx >>= 1;
            order++;
        }
        set_free(c, order, idx, 1);
        return 0;
    }
    return -1;
}

void* td__arena_alloc(size_t bytes) {
    if (!td__dev_ready()) return NULL;
    if (bytes == 0) bytes = 1;
    pthread_mutex_lock(&g_amu);
    void* r = NULL;
    if (g_fail_left > 0 && bytes >= g_fail_min) {   /* injected out-of-memory */
        g_fail_left--;
        pthread_mutex_unlock(&g_amu);
        return NULL;
    }
    if (bytes > A_CHUNK / 2) {
        size_t want = (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
        int best = -1;
        for (int i = 0; i < A_MAX_DIRECT; i++) {   /* reuse a cached block that fits */
            direct_t* d = &g_direct[i];
            if (d->p && !d->live && d->bytes >= want && d->bytes <= want + want / 4 &&
                (best < 0 || d->bytes < g_direct[best].bytes))
                best = i;
        }
        if (best >= 0) {
            g_direct[best].live = 1;
            r = g_direct[best].p;
            g_in_use += g_direct[best].bytes;
        } else {
            int slot = -1;
            for (int i = 0; i < A_MAX_DIRECT && slot < 0; i++) if (!g_direct[i].p) slot = i;
            if (slot >= 0 && g_reserved + want > g_budget) {   /* make room from the cache */
                for (int i = 0; i < A_MAX_DIRECT && g_reserved + want > g_budget; i++) {
                    if (g_direct[i].p && !g_direct[i].live) {
                        tdd_free(g_direct[i].p);
                        g_reserved -= g_direct[i].bytes;
                        g_direct[i].p = NULL;
                    }
                }
            }
            void* d = NULL;
            if (slot >= 0 && g_reserved + want <= g_budget && tdd_malloc(&d, want) == TD_OK) {
                g_direct[slot].p = d;
                g_direct[slot].bytes = want;
                g_direct[slot].live = 1;
                g_reserved += want;
                g_in_use += want;
                r = d;
            }
        }
    } else {
        int order = A_MIN_ORDER;
        while (((size_t)1 << order) < bytes) order++;
        r = chunk_alloc(order);
        if (r) g_in_use += (size_t)1 << order;
    }
    if (g_in_use > g_peak) g_peak = g_in_use;
    pthread_mutex_unlock(&g_amu);
    return r;
}

void td__arena_free(void* p) {
    if (!p) return;
    pthread_mutex_lock(&g_amu);
    if (chunk_free(p) != 0) {
        for (int i = 0; i < A_MAX_DIRECT; i++) {
            if (g_direct[i].p == p && g_direct[i].live) {
                g_direct[i].live = 0;
                g_in_use -= g_direct[i].bytes;
                break;
            }
        }
    }
    pthread_mutex_unlock(&g_amu);
}

/* Return cached large blocks to the device (td_heap_gc). */
void td__arena_release_cached(void) {
    pthread_mutex_lock(&g_amu);
    for (int i = 0; i < A_MAX_DIRECT; i++) {
        if (g_direct[i].p && !g_direct[i].live) {
            tdd_free(g_direct[i].p);
            g_reserved -= g_direct[i].bytes;
            g_direct[i].p = NULL;
        }
    }
    pthread_mutex_unlock(&g_amu);
}

size_t td__arena_in_use(void) { return g_in_use; }
size_t td__arena_peak(void) { return g_peak; }
size_t td__arena_reserved(void) { return g_reserved; }
size_t td__arena_budget(void) { return g_budget; }
void   td__arena_set_budget(size_t b) { g_budget = b; }

void td__dev_forget(td_t* v) { (void)v; }

/* ---------------------------------------------------------------- proxies -- */

td_t* td__hbm_vec_new(int8_t type, uint8_t attrs, int64_t len, size_t bytes) {
    void* d = td__arena_alloc(bytes ? bytes : 8);
    if (!d) return TD_ERR_PTR(TD_ERR_OOM);
    td_t* v = td_alloc(sizeof(td_hbm_ref_t));
    if (!v) { td__arena_free(d); return TD_ERR_PTR(TD_ERR_OOM); }
    v->mmod = TD_MMOD_HBM;
    v->type = type;
    v->attrs = attrs;
    v->len = len;
    memset(td_hbm_ref(v), 0, sizeof(td_hbm_ref_t));
    td_hbm_ref(v)->dptr = d;
    td_hbm_ref(v)->bytes = bytes;
    return v;
}

/* Proxy around an arena block the caller already owns (ownership moves). */
td_t* td__hbm_wrap(int8_t type, uint8_t attrs, int64_t len, void* dptr, size_t bytes) {
    td_t* v = td_alloc(sizeof(td_hbm_ref_t));
    if (!v) { td__arena_free(dptr); return TD_ERR_PTR(TD_ERR_OOM); }
    v->mmod = TD_MMOD_HBM;
    v->type = type;
    v->attrs = attrs;
    v->len = len;
    memset(td_hbm_ref(v), 0, sizeof(td_hbm_ref_t));
    td_hbm_ref(v)->dptr = dptr;
    td_hbm_ref(v)->bytes = bytes;
    return v;
}

void td__hbm_free(td_t* v) {
    td_hbm_ref_t* r = td_hbm_ref(v);
    if (r->owner) td_release(r->owner);
    else td__arena_free(r->dptr);
    r->dptr = NULL;
    r->owner = NULL;
}

td_t* td__hbm_view(td_t* owner, int8_t type, uint8_t attrs, int64_t len, size_t off,
                   size_t bytes) {
    td_t* v = td_alloc(sizeof(td_hbm_ref_t));
    if (!v) return TD_ERR_PTR(TD_ERR_OOM);
    v->mmod = TD_MMOD_HBM;
    v->type = type;
    v->attrs = attrs;
    v->len = len;
    memset(td_hbm_ref(v), 0, sizeof(td_hbm_ref_t));
    td_hbm_ref(v)->dptr = (char*)td_hbm_ref(owner)->dptr + off;
    td_hbm_ref(v)->bytes = bytes;
    td_retain(owner);
    td_hbm_ref(v)->owner = owner;
    return v;
}
