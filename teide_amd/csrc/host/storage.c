/*
 * storage.c -- column files and splayed tables, host and HBM.
 *
 * File format (the reference's, store/col.c:35-44, so files written by
 * either library load in the other): a 32-byte header laid out like the
 * td_t header -- bytes 0-15 the inline null bitmap (zero without nulls),
 * byte 16 mmod = 0, 17 order = 0, 18 type, 19 attrs (slice flag cleared),
 * 20-23 rc = 0, 24-31 len -- followed by the raw elements.  A splayed
 * table is a directory holding one such file per column, named by the
 * column, and `.d`, an I64 column file of the column name ids, with the
 * symbol table saved beside it (store/splay.c:46-247).
 *
 * td_col_* / td_splay_* / td_read_splayed follow the reference (host
 * vectors).  td_dev_col_load / td_dev_splay_load stream the elements from
 * the file straight into the HBM arena through pinned staging buffers and
 * give the columns their statistics, as td_dev_put does.
 *
 * A partitioned database (store/part.c) is a root directory holding the
 * symbol file `sym` and one directory per partition, named by digits and
 * dots (YYYY.MM.DD for dates), each holding the table splayed.
 * td_part_load concatenates the partitions; td_read_parted keeps them as
 * parted columns (one vector per partition) beside a MAPCOMMON column of
 * partition keys and row counts.  td_dev_parted_load reads the same into
 * HBM, each parted column's partitions back to back in one arena block.
 */
#define _GNU_SOURCE
#include "internal.h"

#include <dirent.h>
#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>
#include <unistd.h>

#define HDR 32

static int serializable(int8_t t) {
    switch (t) {
    case TD_BOOL: case TD_U8: case TD_CHAR: case TD_I16: case TD_I32: case TD_I64:
    case TD_F64: case TD_DATE: case TD_TIME: case TD_TIMESTAMP: case TD_GUID: case TD_SYM:
        return 1;
    default:
        return 0;
    }
}

/* ------------------------------------------------------------------ save -- */

td_err_t td_col_save(td_t* vec, const char* path) {
    if (!vec || TD_IS_ERR(vec)) return TD_ERR_TYPE;
    if (!path) return TD_ERR_IO;
    if (vec->type == TD_LIST || vec->type == TD_TABLE) return TD_ERR_NYI;   /* nested forms */
    if (!serializable(vec->type)) return TD_ERR_NYI;
    if (vec->len < 0) return TD_ERR_CORRUPT;
    if ((vec->attrs & TD_ATTR_HAS_NULLS) && (vec->attrs & TD_ATTR_NULLMAP_EXT)) return TD_ERR_NYI;
    uint8_t esz = td_sym_elem_size(vec->type, vec->attrs);
    if (esz == 0 && vec->len > 0) return TD_ERR_TYPE;
    size_t bytes = (size_t)vec->len * esz;

    uint8_t h[HDR];
    memset(h, 0, sizeof(h));
    if ((vec->attrs & TD_ATTR_HAS_NULLS) && !(vec->attrs & TD_ATTR_SLICE))
        memcpy(h, vec->nullmap, 16);
    h[18] = (uint8_t)vec->type;
    h[19] = (uint8_t)(vec->attrs & ~TD_ATTR_SLICE);
    if (!(vec->attrs & TD_ATTR_HAS_NULLS)) h[19] &= (uint8_t)~TD_ATTR_NULLMAP_EXT;
    memcpy(h + 24, &vec->len, 8);

    /* the elements: a slice reads its parent; resident data comes back first */
    td_t* src = vec;
    size_t off = 0;
    if (vec->attrs & TD_ATTR_SLICE) {
        src = vec->slice_parent;
        if (!src || vec->slice_offset < 0 || vec->slice_offset + vec->len > src->len)
            return TD_ERR_IO;
        off = (size_t)vec->slice_offset * esz;
    }
    void* hostbuf = NULL;
    const void* data = (const char*)td_data(src) + off;
    if (src->mmod == TD_MMOD_HBM && bytes) {
        hostbuf = malloc(bytes);
        if (!hostbuf) return TD_ERR_OOM;
        if (tdd_d2h(hostbuf, (const char*)td_hbm_ref(src)->dptr + off, bytes) != TD_OK) {
            free(hostbuf);
            return TD_ERR_IO;
        }
        data = hostbuf;
    }
    FILE* f = fopen(path, "wb");
    td_err_t err = TD_OK;
    if (!f) err = TD_ERR_IO;
    else {
        if (fwrite(h, 1, HDR, f) != HDR) err = TD_ERR_IO;
        if (!err && bytes && fwrite(data, 1, bytes, f) != bytes) err = TD_ERR_IO;
        if (fclose(f) != 0 && !err) err = TD_ERR_IO;
    }
    free(hostbuf);
    return err;
}

/* ------------------------------------------------------------------ load -- */

/* Header of a column file: validated type, attrs, length and element size. */
typedef struct { int8_t type; uint8_t attrs; int64_t len; uint8_t esz; uint8_t nullmap[16]; } colhdr_t;

static td_err_t read_header(int fd, size_t file_size, colhdr_t* c) {
    uint8_t h[HDR];
    if (file_size < HDR || pread(fd, h, HDR, 0) != HDR) return TD_ERR_CORRUPT;
    uint32_t magic;
    memcpy(&magic, h, 4);
    /* string lists, nested lists and tables (STRL / LSTG / TTBL) */
    if (magic == 0x4C525453U || magic == 0x4754534CU || magic == 0x4C425454U) return TD_ERR_NYI;
    c->type = (int8_t)h[18];
    c->attrs = h[19];
    memcpy(&c->len, h + 24, 8);
    memcpy(c->nullmap, h, 16);
    if (!serializable(c->type)) return TD_ERR_NYI;
    if (c->len < 0) return TD_ERR_CORRUPT;
    if ((c->attrs & TD_ATTR_HAS_NULLS) && (c->attrs & TD_ATTR_NULLMAP_EXT)) return TD_ERR_NYI;
    c->esz = td_sym_elem_size(c->type, c->attrs);
    if (c->esz == 0 && c->len > 0) return TD_ERR_TYPE;
    if ((uint64_t)c->len > (SIZE_MAX - HDR) / (c->esz ? c->esz : 1)) return TD_ERR_IO;
    if (HDR + (size_t)c->len * c->esz > file_size) return TD_ERR_CORRUPT;
    return TD_OK;
}

static int open_col(const char* path, size_t* size) {
    int fd = open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return -1;
    struct stat st;
    if (fstat(fd, &st) != 0) { close(fd); return -1; }
    *size = (size_t)st.st_size;
    return fd;
}

static td_t* new_host_vec(const colhdr_t* c) {
    td_t* v = c->type == TD_SYM ? td_sym_vec_new(c->attrs & TD_SYM_W_MASK, c->len)
                                : td_vec_new(c->type, c->len);
    if (TD_IS_ERR(v)) return v;
    v->attrs = (uint8_t)(c->attrs & ~TD_ATTR_SLICE);
    if (v->attrs & TD_ATTR_HAS_NULLS) memcpy(v->nullmap, c->nullmap, 16);
    v->len = c->len;
    return v;
}

td_t* td_col_load(const char* path) {
    if (!path) return TD_ERR_PTR(TD_ERR_IO);
    size_t size;
    int fd = open_col(path, &size);
    if (fd < 0) return TD_ERR_PTR(TD_ERR_IO);
    colhdr_t c;
    td_err_t e = read_header(fd, size, &c);
    if (e) { close(fd); return TD_ERR_PTR(e); }
    td_t* v = new_host_vec(&c);
    if (TD_IS_ERR(v)) { close(fd); return v; }
    size_t bytes = (size_t)c.len * c.esz, got = 0;
    while (got < bytes) {
        ssize_t r = pread(fd, (char*)td_data(v) + got, bytes - got, HDR + (off_t)got);
        if (r <= 0) { close(fd); td_release(v); return TD_ERR_PTR(TD_ERR_IO); }
        got += (size_t)r;
    }
    close(fd);
    return v;
}

/* The reference maps the file (mmod 1); here the column is read into an
 * ordinary vector, which behaves the same through the API. */
td_t* td_col_mmap(const char* path) { return td_col_load(path); }

td_t* td_dev_col_load(const char* path) {
    if (!path) return TD_ERR_PTR(TD_ERR_IO);
    if (!td__dev_ready()) return TD_ERR_PTR(TD_ERR_IO);
    size_t size;
    int fd = open_col(path, &size);
    if (fd < 0) return TD_ERR_PTR(TD_ERR_IO);
    colhdr_t c;
    td_err_t e = read_header(fd, size, &c);
    if (!e && (c.attrs & TD_ATTR_HAS_NULLS)) e = TD_ERR_NYI;      /* resident: no null bitmap */
    if (e) { close(fd); return TD_ERR_PTR(e); }
    size_t bytes = (size_t)c.len * c.esz;
    td_t* d = td__hbm_vec_new(c.type, c.attrs & TD_SYM_W_MASK, c.len, bytes);
    if (TD_IS_ERR(d)) { close(fd); return d; }
    int rc = bytes ? tdd_file_h2d(fd, HDR, td_hbm_ref(d)->dptr, bytes) : TD_OK;
    close(fd);
    if (!rc) rc = td__dev_stats(d);
    if (rc) { td_release(d); return TD_ERR_PTR(rc); }
    return d;
}

/* --------------------------------------------------------------- splayed -- */

static int name_ok(const char* name, size_t len) {
    return len > 0 && name[0] != '.' && !memchr(name, '/', len) && !memchr(name, '\\', len) &&
           !memchr(name, '\0', len);
}

td_err_t td_splay_save(td_t* tbl, const char* dir, const char* sym_path) {
    if (!tbl || TD_IS_ERR(tbl) || tbl->type != TD_TABLE) return TD_ERR_TYPE;
    if (!dir) return TD_ERR_IO;
    if (sym_path) {
        td_err_t e = td_sym_save(sym_path);
        if (e) return e;
    }
    if (mkdir(dir, 0755) != 0 && errno != EEXIST) return TD_ERR_IO;
    char path[1024];
    td_t* schema = td_table_schema(tbl);
    if (schema && !TD_IS_ERR(schema)) {
        int n = snprintf(path, sizeof(path), "%s/.d", dir);
        if (n < 0 || (size_t)n >= sizeof(path)) return TD_ERR_RANGE;
        td_err_t e = td_col_save(schema, path);
        if (e) return e;
    }
    for (int64_t c = 0; c < td_table_ncols(tbl); c++) {
        td_t* col = td_table_get_col_idx(tbl, c);
        td_t* name = td_sym_str(td_table_col_name(tbl, c));
        if (!col || !name || !name_ok(td_str_ptr(name), td_str_len(name))) continue;
        int n = snprintf(path, sizeof(path), "%s/%.*s", dir, (int)td_str_len(name), td_str_ptr(name));
        if (n < 0 || (size_t)n >= sizeof(path)) return TD_ERR_RANGE;
        td_err_t e = td_col_save(col, path);
        if (e) return e;
    }
    return TD_OK;
}

static td_t* splay_load(const char* dir, int dev) {
    if (!dir) return TD_ERR_PTR(TD_ERR_IO);
    char path[1024];
    int n = snprintf(path, sizeof(path), "%s/.d", dir);
    if (n < 0 || (size_t)n >= sizeof(path)) return TD_ERR_PTR(TD_ERR_RANGE);
    td_t* schema = td_col_load(path);
    if (TD_IS_ERR(schema)) return schema;
    if (schema->type != TD_I64 && schema->type != TD_SYM) {
        td_release(schema);
        return TD_ERR_PTR(TD_ERR_CORRUPT);
    }
    int64_t ncols = schema->len;
    td_t* tbl = td_table_new(ncols);
    for (int64_t c = 0; c < ncols && !TD_IS_ERR(tbl); c++) {
        int64_t id = ((const int64_t*)td_data(schema))[c];
        td_t* name = td_sym_str(id);
        if (!name || !name_ok(td_str_ptr(name), td_str_len(name))) continue;
        n = snprintf(path, sizeof(path), "%s/%.*s", dir, (int)td_str_len(name), td_str_ptr(name));
        if (n < 0 || (size_t)n >= sizeof(path)) { td_release(tbl); tbl = TD_ERR_PTR(TD_ERR_RANGE); break; }
        td_t* col = dev ? td_dev_col_load(path) : td_col_load(path);
        if (TD_IS_ERR(col)) { td_release(tbl); tbl = col; break; }
        tbl = td_table_add_col(tbl, id, col);
        td_release(col);
    }
    td_release(schema);
    return tbl;
}

td_t* td_splay_load(const char* dir) { return splay_load(dir, 0); }

td_t* td_read_splayed(const char* dir, const char* sym_path) {
    if (sym_path) (void)td_sym_load(sym_path);     /* as the reference: names may stay unresolved */
    return splay_load(dir, 0);
}

td_t* td_dev_splay_load(const char* dir, const char* sym_path) {
    if (sym_path) (void)td_sym_load(sym_path);
    return splay_load(dir, 1);
}

/* ----------------------------------------------------------- partitioned -- */

/* A partition directory name: digits and dots with at least one dot
 * (part.c:128-140). */
static int part_name(const char* s) {
    int dot = 0;
    for (; *s; s++) {
        if (*s == '.') dot = 1;
        else if (*s < '0' || *s > '9') return 0;
    }
    return dot;
}

static int cmp_str(const void* a, const void* b) {
    return strcmp(*(char* const*)a, *(char* const*)b);
}

typedef struct {
    char**  names;
    int64_t n;
} parts_t;

static void parts_free(parts_t* p) {
    for (int64_t i = 0; i < p->n; i++) free(p->names[i]);
    free(p->names);
    p->names = NULL;
    p->n = 0;
}

/* Partition directories of `root`, in byte order. */
static td_err_t parts_list(const char* root, parts_t* out) {
    out->names = NULL;
    out->n = 0;
    DIR* d = opendir(root);
    if (!d) return TD_ERR_IO;
    int64_t cap = 0;
    struct dirent* e;
    while ((e = readdir(d)) != NULL) {
        if (e->d_name[0] == '.' || !part_name(e->d_name)) continue;
        if (out->n == cap) {
            cap = cap ? cap * 2 : 16;
            char** nn = realloc(out->names, (size_t)cap * sizeof(char*));
            if (!nn) { closedir(d); parts_free(out); return TD_ERR_OOM; }
            out->names = nn;
        }
        out->names[out->n] = strdup(e->d_name);
        if (!out->names[out->n]) { closedir(d); parts_free(out); return TD_ERR_OOM; }
        out->n++;
    }
    closedir(d);
    if (out->n == 0) { parts_free(out); return TD_ERR_IO; }
    qsort(out->names, (size_t)out->n, sizeof(char*), cmp_str);
    return TD_OK;
}

static int table_name_ok(const char* t) {
    return t && t[0] && t[0] != '.' && !strchr(t, '/') && !strchr(t, '\\') && !strstr(t, "..");
}

static int is_date_name(const char* s) {
    if (strlen(s) != 10 || s[4] != '.' || s[7] != '.') return 0;
    for (int i = 0; i < 10; i++)
        if (i != 4 && i != 7 && (s[i] < '0' || s[i] > '9')) return 0;
    int m = (s[5] - '0') * 10 + (s[6] - '0'), dd = (s[8] - '0') * 10 + (s[9] - '0');
    return m >= 1 && m <= 12 && dd >= 1 && dd <= 31;
}

static int is_int_name(const char* s) {
    if (*s == '-') s++;
    if (!*s) return 0;
    for (; *s; s++)
        if (*s < '0' || *s > '9') return 0;
    return 1;
}

/* Days from 2000-01-01 of a proleptic Gregorian date (the civil-to-days
 * count of H. Hinnant's published date algorithms, shifted from the 1970
 * epoch by 10957 days), as the reference dates its partitions
 * (part.c:71-84). */
static int32_t days_2000(int64_t y, int64_t m, int64_t d) {
    y -= m <= 2;
    const int64_t era = (y >= 0 ? y : y - 399) / 400;
    const int64_t yoe = y - era * 400;
    const int64_t doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
    const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
    return (int32_t)(era * 146097 + doe - 719468 - 10957);
}

/* The MAPCOMMON column of the partitions: key values typed by what every
 * name is (DATE when all are YYYY.MM.DD, I64 when all are integers, else
 * SYM of the names; part.c:56-68, 399-452) and row counts. */
static td_t* mapcommon_new(const parts_t* P, const int64_t* counts, int64_t* name_id) {
    int all_date = 1, all_int = 1;
    for (int64_t i = 0; i < P->n; i++) {
        all_date = all_date && is_date_name(P->names[i]);
        all_int = all_int && is_int_name(P->names[i]);
    }
    uint8_t mc = all_date ? TD_MC_DATE : all_int ? TD_MC_I64 : TD_MC_SYM;
    td_t* kv = mc == TD_MC_DATE ? td_vec_new(TD_DATE, P->n)
             : mc == TD_MC_I64 ? td_vec_new(TD_I64, P->n)
             : td_sym_vec_new(TD_SYM_W64, P->n);
    td_t* rc = td_vec_new(TD_I64, P->n);
    td_t* m = td_alloc(2 * sizeof(td_t*));
    if (TD_IS_ERR(kv) || TD_IS_ERR(rc) || !m) {
        if (!TD_IS_ERR(kv)) td_release(kv);
        if (!TD_IS_ERR(rc)) td_release(rc);
        if (m) td_release(m);
        return TD_ERR_PTR(TD_ERR_OOM);
    }
    for (int64_t i = 0; i < P->n; i++) {
        const char* s = P->names[i];
        if (mc == TD_MC_DATE)
            ((int32_t*)td_data(kv))[i] = days_2000((s[0] - '0') * 1000 + (s[1] - '0') * 100 +
                                                   (s[2] - '0') * 10 + (s[3] - '0'),
                                                   (s[5] - '0') * 10 + (s[6] - '0'),
                                                   (s[8] - '0') * 10 + (s[9] - '0'));
        else if (mc == TD_MC_I64)
            ((int64_t*)td_data(kv))[i] = strtoll(s, NULL, 10);
        else
            ((int64_t*)td_data(kv))[i] = td_sym_intern(s, strlen(s));
        ((int64_t*)td_data(rc))[i] = counts[i];
    }
    kv->len = P->n;
    rc->len = P->n;
    m->type = TD_MAPCOMMON;
    m->attrs = mc;
    m->len = 2;
    ((td_t**)td_data(m))[0] = kv;
    ((td_t**)td_data(m))[1] = rc;
    const char* nm = mc == TD_MC_DATE ? "date" : "part";
    *name_id = td_sym_intern(nm, strlen(nm));
    return m;
}

static void release_all(td_t** v, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        if (v[i] && !TD_IS_ERR(v[i])) td_release(v[i]);
}

td_t* td_part_load(const char* db_root, const char* table_name) {
    if (!db_root || !table_name_ok(table_name)) return TD_ERR_PTR(TD_ERR_IO);
    parts_t P;
    td_err_t e = parts_list(db_root, &P);
    if (e) return TD_ERR_PTR(e);
    td_t** tabs = calloc((size_t)P.n, sizeof(td_t*));
    td_t* out = tabs ? NULL : TD_ERR_PTR(TD_ERR_OOM);
    char path[1024];
    for (int64_t p = 0; p < P.n && !out; p++) {
        int k = snprintf(path, sizeof(path), "%s/%s/%s", db_root, P.names[p], table_name);
        tabs[p] = k < 0 || (size_t)k >= sizeof(path) ? TD_ERR_PTR(TD_ERR_IO) : td_splay_load(path);
        if (TD_IS_ERR(tabs[p])) {
            /* the first partition's own error; any later failure is an IO error */
            out = p ? TD_ERR_PTR(TD_ERR_IO) : tabs[p];
            tabs[p] = NULL;
        }
    }
    if (!out && P.n == 1) {
        out = tabs[0];
        tabs[0] = NULL;
    } else if (!out) {
        /* columns concatenated across partitions, by position (part.c:214-249) */
        int64_t nc = td_table_ncols(tabs[0]);
        out = td_table_new(nc);
        for (int64_t c = 0; c < nc && !TD_IS_ERR(out); c++) {
            td_t* acc = td_table_get_col_idx(tabs[0], c);
            td_retain(acc);
            for (int64_t p = 1; p < P.n && !TD_IS_ERR(acc); p++) {
                td_t* seg = td_table_get_col_idx(tabs[p], c);
                if (!seg) continue;
                td_t* nx = td_vec_concat(acc, seg);
                td_release(acc);
                acc = nx;
            }
            if (TD_IS_ERR(acc)) { td_release(out); out = acc; break; }
            out = td_table_add_col(out, td_table_col_name(tabs[0], c), acc);
            td_release(acc);
        }
    }
    if (tabs) {
        release_all(tabs, P.n);
        free(tabs);
    }
    parts_free(&P);
    return out;
}

/* A parted column over `n` segments (held by the column). */
static td_t* parted_new(int8_t base, td_t** segs, int64_t n) {
    td_t* v = td_alloc((size_t)n * sizeof(td_t*));
    if (!v) return TD_ERR_PTR(TD_ERR_OOM);
    v->type = (int8_t)(TD_PARTED_BASE + base);
    v->attrs = 0;
    v->len = n;
    for (int64_t i = 0; i < n; i++) {
        ((td_t**)td_data(v))[i] = segs[i];
        if (segs[i]) td_retain(segs[i]);
    }
    return v;
}

/* The parted table of `P`'s partitions of `table_name` under `root`, each
 * partition's table given by load(path) (host vectors, or resident ones
 * for the device loader).  Columns by position of the first partition. */
static td_t* parted_table(const char* root, const char* table_name, const parts_t* P,
                          td_t* (*load)(const char* path, void* arg), void* arg) {
    td_t** tabs = calloc((size_t)P->n, sizeof(td_t*));
    int64_t* counts = calloc((size_t)P->n, sizeof(int64_t));
    td_t** segs = calloc((size_t)P->n, sizeof(td_t*));
    td_t* out = tabs && counts && segs ? NULL : TD_ERR_PTR(TD_ERR_OOM);
    char path[1024];
    for (int64_t p = 0; p < P->n && !out; p++) {
        int k = snprintf(path, sizeof(path), "%s/%s/%s", root, P->names[p], table_name);
        tabs[p] = k < 0 || (size_t)k >= sizeof(path) ? TD_ERR_PTR(TD_ERR_IO) : load(path, arg);
        if (TD_IS_ERR(tabs[p])) { out = TD_ERR_PTR(TD_ERR_IO); tabs[p] = NULL; }
        else counts[p] = td_table_nrows(tabs[p]);
    }
    int64_t nc = out ? 0 : td_table_ncols(tabs[0]);
    if (!out && nc <= 0) out = TD_ERR_PTR(TD_ERR_IO);
    if (!out) {
        int64_t mc_name;
        td_t* mc = mapcommon_new(P, counts, &mc_name);
        out = TD_IS_ERR(mc) ? mc : td_table_new(nc + 1);
        if (!TD_IS_ERR(mc)) {
            out = td_table_add_col(out, mc_name, mc);
            td_release(mc);
        }
        for (int64_t c = 0; c < nc && !TD_IS_ERR(out); c++) {
            td_t* first = td_table_get_col_idx(tabs[0], c);
            for (int64_t p = 0; p < P->n; p++) segs[p] = td_table_get_col_idx(tabs[p], c);
            td_t* pc = parted_new(first->type, segs, P->n);
            if (TD_IS_ERR(pc)) { td_release(out); out = pc; break; }
            out = td_table_add_col(out, td_table_col_name(tabs[0], c), pc);
            td_release(pc);
        }
    }
    if (tabs) release_all(tabs, P->n);
    free(tabs);
    free(counts);
    free(segs);
    return out;
}

static td_t* load_host(const char* path, void* sym_path) {
    return td_read_splayed(path, (const char*)sym_path);
}

td_t* td_read_parted(const char* db_root, const char* table_name) {
    if (!db_root || !table_name_ok(table_name)) return TD_ERR_PTR(TD_ERR_IO);
    char sym[1024];
    int k = snprintf(sym, sizeof(sym), "%s/sym", db_root);
    if (k < 0 || (size_t)k >= sizeof(sym)) return TD_ERR_PTR(TD_ERR_IO);
    td_err_t e = td_sym_load(sym);
    if (e) return TD_ERR_PTR(e);
    parts_t P;
    if ((e = parts_list(db_root, &P))) return TD_ERR_PTR(e);
    td_t* t = parted_table(db_root, table_name, &P, load_host, sym);
    parts_free(&P);
    return t;
}

/* Device loader: one partition's splayed table with each column streamed
 * into its place in a per-column arena block shared by all partitions
 * (offsets advance as partitions are read), as views into that block. */
typedef struct {
    td_t*   owner[256];       /* per column: the block (a resident vector of all rows) */
    int64_t at[256];          /* rows already placed */
    int64_t rows[256];        /* total rows per column */
    int8_t  type[256];
    uint8_t attrs[256];
    int64_t nc;
} devpart_t;

static td_t* load_dev(const char* dir, void* arg) {
    devpart_t* D = arg;
    char path[1024];
    int n = snprintf(path, sizeof(path), "%s/.d", dir);
    if (n < 0 || (size_t)n >= sizeof(path)) return TD_ERR_PTR(TD_ERR_RANGE);
    td_t* schema = td_col_load(path);
    if (TD_IS_ERR(schema)) return schema;
    int64_t nc = schema->len;
    td_t* tbl = nc == D->nc ? td_table_new(nc) : TD_ERR_PTR(TD_ERR_CORRUPT);
    for (int64_t c = 0; c < nc && !TD_IS_ERR(tbl); c++) {
        int64_t id = ((const int64_t*)td_data(schema))[c];
        td_t* name = td_sym_str(id);
        if (!name || !name_ok(td_str_ptr(name), td_str_len(name))) { td_release(tbl); tbl = TD_ERR_PTR(TD_ERR_CORRUPT); break; }
        n = snprintf(path, sizeof(path), "%s/%.*s", dir, (int)td_str_len(name), td_str_ptr(name));
        size_t size;
        int fd = n < 0 || (size_t)n >= sizeof(path) ? -1 : open_col(path, &size);
        colhdr_t h;
        td_err_t e = fd < 0 ? TD_ERR_IO : read_header(fd, size, &h);
        td_t* own = D->owner[c];
        if (!e && (h.attrs & TD_ATTR_HAS_NULLS)) e = TD_ERR_NYI;
        if (!e && (h.type != own->type || td__esz(own) != h.esz ||
                   D->at[c] + h.len > D->rows[c]))
            e = TD_ERR_CORRUPT;
        size_t off = (size_t)D->at[c] * h.esz, bytes = (size_t)h.len * h.esz;
        int rc = e ? (int)e : bytes ? tdd_file_h2d(fd, HDR, (char*)td_hbm_ref(own)->dptr + off, bytes)
                                    : TD_OK;
        if (fd >= 0) close(fd);
        if (rc) { td_release(tbl); tbl = TD_ERR_PTR(rc); break; }
        td_t* seg = td__hbm_view(own, h.type, h.attrs & TD_SYM_W_MASK, h.len, off, bytes);
        if (TD_IS_ERR(seg)) { td_release(tbl); tbl = seg; break; }
        D->at[c] += h.len;
        tbl = td_table_add_col(tbl, id, seg);
        td_release(seg);
    }
    td_release(schema);
    return tbl;
}

td_t* td_dev_parted_load(const char* db_root, const char* table_name) {
    if (!db_root || !table_name_ok(table_name)) return TD_ERR_PTR(TD_ERR_IO);
    if (!td__dev_ready()) return TD_ERR_PTR(TD_ERR_IO);
    char sym[1024], path[1024];
    int k = snprintf(sym, sizeof(sym), "%s/sym", db_root);
    if (k < 0 || (size_t)k >= sizeof(sym)) return TD_ERR_PTR(TD_ERR_IO);
    td_err_t e = td_sym_load(sym);
    if (e) return TD_ERR_PTR(e);
    parts_t P;
    if ((e = parts_list(db_root, &P))) return TD_ERR_PTR(e);
    /* sizes first: every partition's column headers, so each column's
     * partitions can land back to back in one block */
    devpart_t* D = calloc(1, sizeof(devpart_t));
    td_t* out = D ? NULL : TD_ERR_PTR(TD_ERR_OOM);
    for (int64_t p = 0; p < P.n && !out; p++) {
        k = snprintf(path, sizeof(path), "%s/%s/%s/.d", db_root, P.names[p], table_name);
        td_t* schema = k < 0 || (size_t)k >= sizeof(path) ? TD_ERR_PTR(TD_ERR_IO) : td_col_load(path);
        if (TD_IS_ERR(schema)) { out = TD_ERR_PTR(TD_ERR_IO); break; }
        if (p == 0) D->nc = schema->len;
        if (schema->len != D->nc || D->nc > 256) { td_release(schema); out = TD_ERR_PTR(TD_ERR_CORRUPT); break; }
        for (int64_t c = 0; c < D->nc && !out; c++) {
            td_t* name = td_sym_str(((const int64_t*)td_data(schema))[c]);
            if (!name || !name_ok(td_str_ptr(name), td_str_len(name))) { out = TD_ERR_PTR(TD_ERR_CORRUPT); break; }
            k = snprintf(path, sizeof(path), "%s/%s/%s/%.*s", db_root, P.names[p], table_name,
                         (int)td_str_len(name), td_str_ptr(name));
            size_t size;
            int fd = k < 0 || (size_t)k >= sizeof(path) ? -1 : open_col(path, &size);
            colhdr_t h;
            e = fd < 0 ? TD_ERR_IO : read_header(fd, size, &h);
            if (fd >= 0) close(fd);
            if (e) { out = TD_ERR_PTR(e); break; }
            if (p == 0) {
                D->type[c] = h.type;
                D->attrs[c] = h.attrs & TD_SYM_W_MASK;
            } else if (h.type != D->type[c] || (h.attrs & TD_SYM_W_MASK) != D->attrs[c]) {
                out = TD_ERR_PTR(TD_ERR_CORRUPT);           /* partitions disagree */
                break;
            }
            D->rows[c] += h.len;
        }
        td_release(schema);
    }
    /* the blocks, then the data */
    for (int64_t c = 0; c < (D ? D->nc : 0) && !out; c++) {
        size_t bytes = (size_t)D->rows[c] * td_sym_elem_size(D->type[c], D->attrs[c]);
        td_t* b = td__hbm_vec_new(D->type[c], D->attrs[c], D->rows[c], bytes);
        D->owner[c] = TD_IS_ERR(b) ? NULL : b;
        if (TD_IS_ERR(b)) out = b;
    }
    if (!out) out = parted_table(db_root, table_name, &P, load_dev, D);
    for (int64_t c = 0; D && c < D->nc && !TD_IS_ERR(out); c++) {
        int rc = D->at[c] == D->rows[c] ? td__dev_stats(D->owner[c]) : TD_ERR_CORRUPT;
        if (rc) { td_release(out); out = TD_ERR_PTR(rc); }
    }
    for (int64_t c = 0; D && c < D->nc; c++)
        if (D->owner[c]) td_release(D->owner[c]);
    free(D);
    parts_free(&P);
    return out;
}
