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
 */
#define _GNU_SOURCE
#include "internal.h"

#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
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
