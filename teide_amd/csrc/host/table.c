/*
 * table.c -- tables and the global symbol table.
 *
 * A table's payload is [schema | col_0 | col_1 ...]: slot 0 holds an I64
 * vector of column-name symbol ids, slot 1+i the i-th column, len is the
 * column count and the row count is the first column's length (the
 * reference's src/table/table.c:28-228).  Adding a column copies a shared
 * table first and may move it, so callers use the returned pointer.
 *
 * Symbols are interned strings numbered 0, 1, 2 ... in interning order
 * (src/table/sym.c:170-303); they name columns and are the dictionary of
 * TD_SYM columns.
 */
#include "internal.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>

static td_t** schema_slot(td_t* t) { return (td_t**)td_data(t); }
static td_t** col_slots(td_t* t) { return (td_t**)td_data(t) + 1; }

td_t* td_table_new(int64_t ncols) {
    if (ncols < 0) return TD_ERR_PTR(TD_ERR_RANGE);
    td_t* t = td_alloc((size_t)(ncols + 1) * sizeof(td_t*));
    if (!t) return TD_ERR_PTR(TD_ERR_OOM);
    t->type = TD_TABLE;
    t->len = 0;
    memset(td_data(t), 0, (size_t)(ncols + 1) * sizeof(td_t*));
    td_t* schema = td_vec_new(TD_I64, ncols);
    if (!schema || TD_IS_ERR(schema)) { td_free(t); return schema; }
    *schema_slot(t) = schema;
    return t;
}

td_t* td_table_add_col(td_t* tbl, int64_t name_id, td_t* col) {
    if (!tbl || TD_IS_ERR(tbl)) return tbl;
    if (!col || TD_IS_ERR(col)) return TD_ERR_PTR(TD_ERR_TYPE);
    tbl = td_cow(tbl);
    if (!tbl || TD_IS_ERR(tbl)) return tbl;
    int64_t idx = tbl->len;
    int64_t max_cols = (int64_t)((((size_t)1 << tbl->order) - 32) / sizeof(td_t*)) - 1;
    if (idx >= max_cols) {
        td_t* n = td_scratch_realloc(tbl, (size_t)(1 + (idx + 1) * 2) * sizeof(td_t*));
        if (!n || TD_IS_ERR(n)) return n;
        tbl = n;
    }
    td_t* schema = td_vec_append(*schema_slot(tbl), &name_id);
    if (!schema || TD_IS_ERR(schema)) return TD_ERR_PTR(TD_ERR_OOM);
    *schema_slot(tbl) = schema;
    col_slots(tbl)[idx] = col;
    td_retain(col);
    tbl->len = idx + 1;
    return tbl;
}

td_t* td_table_get_col(td_t* tbl, int64_t name_id) {
    if (!tbl || TD_IS_ERR(tbl) || tbl->type != TD_TABLE) return NULL;
    td_t* schema = *schema_slot(tbl);
    if (!schema) return NULL;
    const int64_t* ids = (const int64_t*)td_data(schema);
    for (int64_t i = 0; i < tbl->len; i++)
        if (ids[i] == name_id) return col_slots(tbl)[i];
    return NULL;
}

td_t* td_table_get_col_idx(td_t* tbl, int64_t idx) {
    if (!tbl || TD_IS_ERR(tbl) || idx < 0 || idx >= tbl->len) return NULL;
    return col_slots(tbl)[idx];
}

int64_t td_table_col_name(td_t* tbl, int64_t idx) {
    if (!tbl || TD_IS_ERR(tbl) || idx < 0 || idx >= tbl->len) return -1;
    td_t* schema = *schema_slot(tbl);
    return schema ? ((const int64_t*)td_data(schema))[idx] : -1;
}

void td_table_set_col_name(td_t* tbl, int64_t idx, int64_t name_id) {
    if (!tbl || TD_IS_ERR(tbl) || idx < 0 || idx >= tbl->len) return;
    td_t* schema = td_cow(*schema_slot(tbl));
    if (!schema || TD_IS_ERR(schema)) return;
    *schema_slot(tbl) = schema;
    ((int64_t*)td_data(schema))[idx] = name_id;
}

int64_t td_table_ncols(td_t* tbl) {
    if (!tbl || TD_IS_ERR(tbl)) return 0;
    return tbl->len;
}

int64_t td_parted_nrows(td_t* v) {
    if (!v || TD_IS_ERR(v)) return 0;
    if (!TD_IS_PARTED(v->type) && v->type != TD_MAPCOMMON) return v->len;
    if (v->type == TD_MAPCOMMON) {
        td_t* counts = ((td_t**)td_data(v))[1];
        int64_t total = 0;
        for (int64_t i = 0; counts && i < counts->len; i++) total += ((int64_t*)td_data(counts))[i];
        return total;
    }
    int64_t total = 0;
    td_t** segs = (td_t**)td_data(v);
    for (int64_t i = 0; i < v->len; i++) if (segs[i]) total += segs[i]->len;
    return total;
}

int64_t td_table_nrows(td_t* tbl) {
    if (!tbl || TD_IS_ERR(tbl) || tbl->len <= 0) return 0;
    td_t* c0 = col_slots(tbl)[0];
    if (!c0 || TD_IS_ERR(c0)) return 0;
    return td_parted_nrows(c0);
}

td_t* td_table_schema(td_t* tbl) {
    if (!tbl || TD_IS_ERR(tbl)) return NULL;
    return *schema_slot(tbl);
}

/* --------------------------------------------------------------- symbols -- */

typedef struct {
    uint64_t* slots;   /* (fnv32 << 32) | (id + 1); 0 = empty */
    uint32_t  cap;
    td_t**    strs;
    uint32_t  count, str_cap;
    int       live;
} symtab_t;

static symtab_t g_sym;
static pthread_mutex_t g_sym_lock = PTHREAD_MUTEX_INITIALIZER;

static uint32_t fnv1a32(const char* s, size_t n) {
    uint32_t h = 2166136261u;
    for (size_t i = 0; i < n; i++) { h ^= (uint8_t)s[i]; h *= 16777619u; }
    return h;
}

void td_sym_init(void) {
    pthread_mutex_lock(&g_sym_lock);
    if (!g_sym.live) {
        g_sym.cap = 1024;
        g_sym.slots = calloc(g_sym.cap, sizeof(uint64_t));
        g_sym.str_cap = 1024;
        g_sym.strs = calloc(g_sym.str_cap, sizeof(td_t*));
        g_sym.count = 0;
        g_sym.live = g_sym.slots && g_sym.strs;
    }
    pthread_mutex_unlock(&g_sym_lock);
}

void td_sym_destroy(void) {
    pthread_mutex_lock(&g_sym_lock);
    if (g_sym.live) {
        for (uint32_t i = 0; i < g_sym.count; i++) td_release(g_sym.strs[i]);
        free(g_sym.slots);
        free(g_sym.strs);
        memset(&g_sym, 0, sizeof(g_sym));
    }
    pthread_mutex_unlock(&g_sym_lock);
}

static int64_t find_locked(const char* s, size_t n, uint32_t h, uint32_t* slot_out) {
    uint32_t mask = g_sym.cap - 1, i = h & mask;
    for (;;) {
        uint64_t e = g_sym.slots[i];
        if (!e) { *slot_out = i; return -1; }
        if ((uint32_t)(e >> 32) == h) {
            uint32_t id = (uint32_t)e - 1;
            td_t* str = g_sym.strs[id];
            if (td_str_len(str) == n && memcmp(td_str_ptr(str), s, n) == 0) return id;
        }
        i = (i + 1) & mask;
    }
}

static void grow_locked(void) {
    uint32_t ncap = g_sym.cap * 2;
    uint64_t* ns = calloc(ncap, sizeof(uint64_t));
    if (!ns) return;
    for (uint32_t i = 0; i < g_sym.cap; i++) {
        uint64_t e = g_sym.slots[i];
        if (!e) continue;
        uint32_t j = (uint32_t)(e >> 32) & (ncap - 1);
        while (ns[j]) j = (j + 1) & (ncap - 1);
        ns[j] = e;
    }
    free(g_sym.slots);
    g_sym.slots = ns;
    g_sym.cap = ncap;
}

int64_t td_sym_intern(const char* s, size_t n) {
    if (!g_sym.live) td_sym_init();
    pthread_mutex_lock(&g_sym_lock);
    uint32_t h = fnv1a32(s, n), slot;
    int64_t id = find_locked(s, n, h, &slot);
    if (id < 0) {
        if (g_sym.count == g_sym.str_cap) {
            td_t** ns = realloc(g_sym.strs, (size_t)g_sym.str_cap * 2 * sizeof(td_t*));
            if (!ns) { pthread_mutex_unlock(&g_sym_lock); return -1; }
            g_sym.strs = ns;
            g_sym.str_cap *= 2;
        }
        td_t* str = td_str(s, n);
        if (!str || TD_IS_ERR(str)) { pthread_mutex_unlock(&g_sym_lock); return -1; }
        id = g_sym.count++;
        g_sym.strs[id] = str;
        g_sym.slots[slot] = ((uint64_t)h << 32) | (uint64_t)(id + 1);
        if ((uint64_t)g_sym.count * 10 > (uint64_t)g_sym.cap * 7) grow_locked();
    }
    pthread_mutex_unlock(&g_sym_lock);
    return id;
}

int64_t td_sym_find(const char* s, size_t n) {
    if (!g_sym.live) return -1;
    pthread_mutex_lock(&g_sym_lock);
    uint32_t slot;
    int64_t id = find_locked(s, n, fnv1a32(s, n), &slot);
    pthread_mutex_unlock(&g_sym_lock);
    return id;
}

td_t* td_sym_str(int64_t id) {
    if (!g_sym.live || id < 0) return NULL;
    pthread_mutex_lock(&g_sym_lock);
    td_t* r = (uint64_t)id < g_sym.count ? g_sym.strs[id] : NULL;
    pthread_mutex_unlock(&g_sym_lock);
    return r;
}

uint32_t td_sym_count(void) {
    return g_sym.live ? g_sym.count : 0;
}

/* File format: "TSYM", u32 count, then per symbol u32 length + bytes. */
td_err_t td_sym_save(const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return TD_ERR_IO;
    uint32_t n = td_sym_count();
    fwrite("TSYM", 1, 4, f);
    fwrite(&n, 4, 1, f);
    for (uint32_t i = 0; i < n; i++) {
        td_t* s = td_sym_str(i);
        uint32_t l = (uint32_t)td_str_len(s);
        fwrite(&l, 4, 1, f);
        fwrite(td_str_ptr(s), 1, l, f);
    }
    return fclose(f) == 0 ? TD_OK : TD_ERR_IO;
}

td_err_t td_sym_load(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) return TD_ERR_IO;
    char magic[4];
    uint32_t n;
    if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "TSYM", 4) || fread(&n, 4, 1, f) != 1) {
        fclose(f);
        return TD_ERR_CORRUPT;
    }
    char buf[4096];
    for (uint32_t i = 0; i < n; i++) {
        uint32_t l;
        if (fread(&l, 4, 1, f) != 1 || l > sizeof(buf) || fread(buf, 1, l, f) != l) {
            fclose(f);
            return TD_ERR_CORRUPT;
        }
        td_sym_intern(buf, l);
    }
    fclose(f);
    return TD_OK;
}

int64_t td__intern_cstr(const char* s) { return td_sym_intern(s, strlen(s)); }

const char* td__sym_cstr(int64_t id, size_t* len) {
    td_t* s = td_sym_str(id);
    if (len) *len = s ? td_str_len(s) : 0;
    return s ? td_str_ptr(s) : NULL;
}
