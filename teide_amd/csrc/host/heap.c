/*
 * heap.c -- host block allocator and reference counting.
 *
 * Blocks are powers of two (2^order bytes, order >= 6) with the 32-byte td_t
 * header first, the same contract as the reference's buddy heap
 * (src/mem/heap.c:468-633): td_alloc(n) returns a block of order
 * ceil(log2(n + 32)), rc = 1, mmod = 0, and a vector's capacity is
 * (2^order - 32) / elem_size.
 *
 * Small blocks (order <= 25) are carved from 32 MB pools by a classic buddy
 * system with one free list per order and coalescing on free; larger blocks
 * are mapped directly.  One process-wide lock serialises the heap: the td.h
 * API is single-caller, and all parallel work happens on the GPU.
 *
 * A block whose mmod is 3 is the host proxy of a device-resident object;
 * freeing it also returns its device memory to the HBM arena.
 */
#define _GNU_SOURCE
#include "internal.h"

#include <pthread.h>
#include <stdlib.h>
#include <sys/mman.h>

#define POOL_ORDER   25
#define POOL_SIZE    ((size_t)1 << POOL_ORDER)
#define MAX_ORDER    40
#define MAX_POOLS    4096

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static td_t* g_free[POOL_ORDER + 1];        /* circular lists, NULL = empty */
static char* g_pools[MAX_POOLS];
static int   g_npools;
static td_mem_stats_t g_stats;

void td__heap_lock(void)   { pthread_mutex_lock(&g_lock); }
void td__heap_unlock(void) { pthread_mutex_unlock(&g_lock); }

uint8_t td_order_for_size(size_t data_size) {
    size_t need = data_size + 32;
    uint8_t o = TD_ORDER_MIN;
    while (((size_t)1 << o) < need && o < 63) o++;
    return o;
}

/* ------------------------------------------------------------ free lists -- */

static void fl_push(td_t* b, uint8_t order) {
    b->mmod = TD_MMOD_FREE;
    b->order = order;
    atomic_store_explicit(&b->rc, 0, memory_order_relaxed);
    td_t* h = g_free[order];
    if (!h) {
        b->fl_prev = b->fl_next = b;
    } else {
        b->fl_next = h;
        b->fl_prev = h->fl_prev;
        h->fl_prev->fl_next = b;
        h->fl_prev = b;
    }
    g_free[order] = b;
}

static void fl_remove(td_t* b, uint8_t order) {
    if (b->fl_next == b) {
        g_free[order] = NULL;
    } else {
        b->fl_prev->fl_next = b->fl_next;
        b->fl_next->fl_prev = b->fl_prev;
        if (g_free[order] == b) g_free[order] = b->fl_next;
    }
    b->mmod = TD_MMOD_HEAP;
}

static int add_pool(void) {
    if (g_npools >= MAX_POOLS) return 0;
    /* Map twice the size and trim so the pool is self-aligned: a block's
     * buddy is then addr ^ size within the pool. */
    char* raw = mmap(NULL, 2 * POOL_SIZE, PROT_READ | PROT_WRITE,
                     MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (raw == MAP_FAILED) return 0;
    uintptr_t a = ((uintptr_t)raw + POOL_SIZE - 1) & ~(uintptr_t)(POOL_SIZE - 1);
    char* pool = (char*)a;
    if (pool > raw) munmap(raw, (size_t)(pool - raw));
    char* end = pool + POOL_SIZE;
    if (end < raw + 2 * POOL_SIZE) munmap(end, (size_t)(raw + 2 * POOL_SIZE - end));
    g_pools[g_npools++] = pool;
    g_stats.sys_current += POOL_SIZE;
    if (g_stats.sys_current > g_stats.sys_peak) g_stats.sys_peak = g_stats.sys_current;
    fl_push((td_t*)pool, POOL_ORDER);
    return 1;
}

static char* pool_of(const void* p) {
    uintptr_t base = (uintptr_t)p & ~(uintptr_t)(POOL_SIZE - 1);
    for (int i = 0; i < g_npools; i++)
        if ((uintptr_t)g_pools[i] == base) return g_pools[i];
    return NULL;
}

/* ------------------------------------------------------------ alloc/free -- */

static td_t* heap_alloc_locked(uint8_t order) {
    int o = order;
    while (o <= POOL_ORDER && !g_free[o]) o++;
    if (o > POOL_ORDER) {
        if (!add_pool()) return NULL;
        o = POOL_ORDER;
    }
    td_t* b = g_free[o];
    fl_remove(b, (uint8_t)o);
    while (o > order) {                          /* split, keep the low half */
        o--;
        td_t* hi = (td_t*)((char*)b + ((size_t)1 << o));
        fl_push(hi, (uint8_t)o);
    }
    return b;
}

td_t* td_alloc(size_t data_size) {
    uint8_t order = td_order_for_size(data_size);
    if (order > MAX_ORDER) return NULL;
    td_t* b;
    if (order > POOL_ORDER) {
        void* m = mmap(NULL, (size_t)1 << order, PROT_READ | PROT_WRITE,
                       MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (m == MAP_FAILED) return NULL;
        b = (td_t*)m;
        td__heap_lock();
        g_stats.direct_count++;
        g_stats.direct_bytes += (size_t)1 << order;
    } else {
        td__heap_lock();
        b = heap_alloc_locked(order);
        if (!b) { td__heap_unlock(); return NULL; }
    }
    g_stats.alloc_count++;
    g_stats.bytes_allocated += (size_t)1 << order;
    if (g_stats.bytes_allocated > g_stats.peak_bytes) g_stats.peak_bytes = g_stats.bytes_allocated;
    td__heap_unlock();
    memset(b, 0, 32);
    b->mmod = TD_MMOD_HEAP;
    b->order = order;
    atomic_store_explicit(&b->rc, 1, memory_order_relaxed);
    return b;
}

static int atom_owns_obj(const td_t* v) {
    if (v->type == TD_ATOM_GUID) return v->obj != NULL;
    if (v->type != TD_ATOM_STR) return 0;
    return !((v->slen >= 1 && v->slen <= 7) || (v->slen == 0 && v->obj == NULL));
}

/* Drop the references a block holds on other blocks (children). */
static void release_children(td_t* v) {
    if (td_is_atom(v)) {
        if (atom_owns_obj(v) && v->obj && !TD_IS_ERR(v->obj)) td_release(v->obj);
        return;
    }
    if (v->mmod == TD_MMOD_HBM) return;
    if (v->attrs & TD_ATTR_SLICE) {
        if (v->slice_parent && !TD_IS_ERR(v->slice_parent)) td_release(v->slice_parent);
        return;
    }
    if ((v->attrs & TD_ATTR_NULLMAP_EXT) && v->ext_nullmap && !TD_IS_ERR(v->ext_nullmap))
        td_release(v->ext_nullmap);
    if (v->type == TD_TABLE) {
        td_t** slots = (td_t**)td_data(v);
        if (slots[0] && !TD_IS_ERR(slots[0])) td_release(slots[0]);
        for (int64_t i = 0; i < v->len; i++)
            if (slots[1 + i] && !TD_IS_ERR(slots[1 + i])) td_release(slots[1 + i]);
    } else if (v->type == TD_LIST || TD_IS_PARTED(v->type) || v->type == TD_MAPCOMMON) {
        /* lists, parted columns (one vector per partition) and MAPCOMMON
         * (key values, row counts) hold references to their items */
        td_t** items = (td_t**)td_data(v);
        for (int64_t i = 0; i < v->len; i++)
            if (items[i] && !TD_IS_ERR(items[i])) td_release(items[i]);
    }
}

void td_free(td_t* v) {
    if (!v || TD_IS_ERR(v)) return;
    atomic_store_explicit(&v->rc, 1, memory_order_relaxed);
    release_children(v);
    if (v->mmod == TD_MMOD_HBM) td__hbm_free(v);
    else td__dev_forget(v);                      /* drop any upload cache entry */
    if (v->mmod == TD_MMOD_FILE) return;         /* file mappings are not produced here */
    uint8_t order = v->order;
    if (order < TD_ORDER_MIN || order > MAX_ORDER) return;
    td__heap_lock();
    g_stats.free_count++;
    g_stats.bytes_allocated -= (size_t)1 << order;
    if (order > POOL_ORDER) {
        g_stats.direct_count--;
        g_stats.direct_bytes -= (size_t)1 << order;
        td__heap_unlock();
        munmap(v, (size_t)1 << order);
        return;
    }
    char* pool = pool_of(v);
    if (!pool) { td__heap_unlock(); return; }
    td_t* b = v;
    while (order < POOL_ORDER) {                 /* coalesce with free buddies */
        uintptr_t off = (uintptr_t)((char*)b - pool);
        td_t* buddy = (td_t*)(pool + (off ^ ((uintptr_t)1 << order)));
        if (buddy->mmod != TD_MMOD_FREE || buddy->order != order) break;
        fl_remove(buddy, order);
        if (buddy < b) b = buddy;
        order++;
    }
    fl_push(b, order);
    td__heap_unlock();
}

td_t* td_alloc_copy(td_t* v) {
    if (!v || TD_IS_ERR(v)) return NULL;
    if (v->mmod == TD_MMOD_HBM) return td_dev_put(v);   /* device-side copy */
    size_t data = ((size_t)1 << v->order) - 32;
    td_t* c = td_alloc(data);
    if (!c) return NULL;
    uint8_t order = c->order;
    memcpy(c, v, (size_t)1 << (v->order < order ? v->order : order));
    c->mmod = TD_MMOD_HEAP;
    c->order = order;
    atomic_store_explicit(&c->rc, 1, memory_order_relaxed);
    c->attrs &= (uint8_t)~TD_ATTR_SLICE;
    /* The copy holds its own references to children. */
    if (c->type == TD_TABLE) {
        td_t** s = (td_t**)td_data(c);
        if (s[0]) { td_t* sc = td_alloc_copy(s[0]); s[0] = sc; }
        for (int64_t i = 0; i < c->len; i++) td_retain(s[1 + i]);
    } else if (c->type == TD_LIST || TD_IS_PARTED(c->type) || c->type == TD_MAPCOMMON) {
        td_t** it = (td_t**)td_data(c);
        for (int64_t i = 0; i < c->len; i++) if (it[i]) td_retain(it[i]);
    } else if (td_is_atom(c) && atom_owns_obj(c)) {
        td_retain(c->obj);
    } else if ((c->attrs & TD_ATTR_NULLMAP_EXT) && c->ext_nullmap) {
        td_retain(c->ext_nullmap);
    }
    return c;
}

td_t* td_scratch_alloc(size_t data_size) { return td_alloc(data_size); }

td_t* td_scratch_realloc(td_t* v, size_t new_data_size) {
    if (!v || TD_IS_ERR(v)) return td_alloc(new_data_size);
    td_t* n = td_alloc(new_data_size);
    if (!n) return TD_ERR_PTR(TD_ERR_OOM);
    size_t old_data = ((size_t)1 << v->order) - 32;
    size_t keep = old_data < new_data_size ? old_data : new_data_size;
    uint8_t order = n->order;
    memcpy(n, v, 32);
    memcpy(td_data(n), td_data(v), keep);
    n->order = order;
    n->mmod = TD_MMOD_HEAP;
    atomic_store_explicit(&n->rc, 1, memory_order_relaxed);
    /* Children move with the payload: free the old block without
     * releasing them. */
    v->type = TD_U8;                              /* no children to release */
    v->attrs = 0;
    td_free(v);
    return n;
}

void td_heap_init(void) {}
void td_heap_destroy(void) {}
void td_heap_merge(td_heap_t* src) { (void)src; }
void td_heap_gc(void) { td__arena_release_cached(); }
void td_heap_release_pages(void) {}

void td_mem_stats(td_mem_stats_t* out) {
    if (!out) return;
    td__heap_lock();
    *out = g_stats;
    td__heap_unlock();
}

/* ---------------------------------------------------- reference counting -- */

void td_retain(td_t* v) {
    if (!v || TD_IS_ERR(v)) return;
    atomic_fetch_add_explicit(&v->rc, 1, memory_order_relaxed);
}

void td_release(td_t* v) {
    if (!v || TD_IS_ERR(v)) return;
    if (atomic_fetch_sub_explicit(&v->rc, 1, memory_order_acq_rel) == 1) td_free(v);
}

td_t* td_cow(td_t* v) {
    if (!v || TD_IS_ERR(v)) return v;
    if (atomic_load_explicit(&v->rc, memory_order_acquire) == 1) return v;
    td_t* c = td_alloc_copy(v);
    if (!c || TD_IS_ERR(c)) return c ? c : TD_ERR_PTR(TD_ERR_OOM);
    td_release(v);
    return c;
}
