/*
 * dev_api.c -- the td_dev.h extensions: residency (device-resident copies
 * of vectors, selections and tables), arena accounting, kernel timing and
 * physical-plan hints.
 */
#include "internal.h"

#include <stdarg.h>
#include <stdio.h>

static int g_force_join, g_force_group, g_force_bits;
static char g_plan[256] = "-";

int td__force_join_path(void) { return g_force_join; }
int td__force_group_path(void) { return g_force_group; }
int td__force_radix_bits(void) { return g_force_bits; }

void td__set_plan(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_plan, sizeof(g_plan), fmt, ap);
    va_end(ap);
}

const char* td_dev_last_plan(void) { return g_plan; }

void td_dev_force_paths(int join_path, int group_path, int radix_bits) {
    g_force_join = join_path;
    g_force_group = group_path;
    g_force_bits = radix_bits;
}

int td_dev_available(void) { return td__dev_ready(); }

td_err_t td_dev_sync(void) {
    if (!td__dev_ready()) return TD_ERR_IO;
    return (td_err_t)tdd_sync();
}

static size_t payload_bytes(td_t* v) {
    if (v->type == TD_SEL) return td__sel_payload_bytes(v->len);
    return (size_t)v->len * td__esz(v);
}

/* Column statistics for an immutable resident vector: min / max of an
 * integer column, computed once so queries size direct maps and group
 * tables without a scan. */
int td__dev_stats(td_t* d) {
    int8_t t = d->type;
    int int_col = t == TD_BOOL || t == TD_U8 || t == TD_I16 || t == TD_I32 || t == TD_I64 ||
                  t == TD_DATE || t == TD_TIME || t == TD_TIMESTAMP || t == TD_SYM;
    if (!int_col || d->len <= 0) return TD_OK;
    tdd_col_t c;
    memset(&c, 0, sizeof(c));
    c.data = td_hbm_ref(d)->dptr;
    c.type = t;
    c.attrs = d->attrs & TD_SYM_W_MASK;
    c.esz = (uint8_t)td__esz(d);
    int64_t lo, hi;
    int rc = tdd_minmax(&c, d->len, NULL, &lo, &hi);
    if (!rc) {
        td_hbm_ref(d)->lo = lo;
        td_hbm_ref(d)->hi = hi;
        td_hbm_ref(d)->ranged = 1;
    }
    return rc;
}

/* Device-resident copy of one vector / selection. */
static td_t* put_vec(td_t* v) {
    if (v->type <= 0 || v->type == TD_TABLE || v->type == TD_LIST || TD_IS_PARTED(v->type) ||
        v->type == TD_MAPCOMMON)
        return TD_ERR_PTR(TD_ERR_TYPE);
    size_t bytes = payload_bytes(v);
    td_t* d = td__hbm_vec_new(v->type, v->attrs & TD_SYM_W_MASK, v->len, bytes);
    if (TD_IS_ERR(d)) return d;
    int rc;
    if (v->mmod == TD_MMOD_HBM) {
        rc = tdd_d2d(td_hbm_ref(d)->dptr, td_hbm_ref(v)->dptr, bytes);
    } else {
        const void* src = td_data(v);
        if (v->attrs & TD_ATTR_SLICE) {
            td_t* p = v->slice_parent;
            if (p->mmod == TD_MMOD_HBM) {
                rc = tdd_d2d(td_hbm_ref(d)->dptr,
                             (char*)td_hbm_ref(p)->dptr + v->slice_offset * td__esz(p), bytes);
                goto done;
            }
            src = (const char*)td_data(p) + v->slice_offset * td__esz(p);
        }
        rc = tdd_h2d(td_hbm_ref(d)->dptr, src, bytes);
    }
done:
    if (!rc) rc = td__dev_stats(d);
    if (!rc) rc = tdd_sync();
    if (rc) { td_release(d); return TD_ERR_PTR(rc); }
    return d;
}

/* A parted column with all its segments back to back in one arena block,
 * each a view into it; the block carries the column statistics. */
static td_t* put_parted(td_t* v) {
    td_t** segs = (td_t**)td_data(v);
    int8_t base = (int8_t)TD_PARTED_BASETYPE(v->type);
    uint8_t attrs = 0;
    for (int64_t i = 0; i < v->len; i++)
        if (segs[i]) { attrs = segs[i]->attrs & TD_SYM_W_MASK; break; }
    uint8_t esz = td_sym_elem_size(base, attrs);
    int64_t total = td_parted_nrows(v);
    td_t* block = td__hbm_vec_new(base, attrs, total, (size_t)total * esz);
    if (TD_IS_ERR(block)) return block;
    td_t* out = td_alloc((size_t)v->len * sizeof(td_t*));
    if (!out) { td_release(block); return TD_ERR_PTR(TD_ERR_OOM); }
    out->type = v->type;
    out->attrs = 0;
    out->len = v->len;
    memset(td_data(out), 0, (size_t)v->len * sizeof(td_t*));
    int64_t at = 0;
    int rc = TD_OK;
    for (int64_t i = 0; i < v->len && !rc; i++) {
        td_t* s = segs[i];
        if (!s) continue;
        if (s->type != base || td__esz(s) != esz) { rc = TD_ERR_TYPE; break; }
        size_t off = (size_t)at * esz, bytes = (size_t)s->len * esz;
        char* dst = (char*)td_hbm_ref(block)->dptr + off;
        if (bytes) {
            if (s->mmod == TD_MMOD_HBM) {
                rc = tdd_d2d(dst, td_hbm_ref(s)->dptr, bytes);
            } else if (s->attrs & TD_ATTR_SLICE) {
                td_t* p = s->slice_parent;
                rc = p->mmod == TD_MMOD_HBM
                   ? tdd_d2d(dst, (char*)td_hbm_ref(p)->dptr + s->slice_offset * esz, bytes)
                   : tdd_h2d(dst, (const char*)td_data(p) + s->slice_offset * esz, bytes);
            } else {
                rc = tdd_h2d(dst, td_data(s), bytes);
            }
        }
        td_t* view = rc ? NULL : td__hbm_view(block, base, attrs, s->len, off, bytes);
        if (view && TD_IS_ERR(view)) { rc = TD_ERR_CODE(view); break; }
        ((td_t**)td_data(out))[i] = view;
        at += s->len;
    }
    if (!rc) rc = td__dev_stats(block);
    if (!rc) rc = tdd_sync();
    td_release(block);                       /* the views hold it */
    if (rc) { td_release(out); return TD_ERR_PTR(rc); }
    return out;
}

td_t* td_dev_put(td_t* v) {
    if (!v || TD_IS_ERR(v)) return v;
    if (!td__dev_ready()) return TD_ERR_PTR(TD_ERR_IO);
    if (TD_IS_PARTED(v->type)) return put_parted(v);
    if (v->type != TD_TABLE) return put_vec(v);
    int64_t nc = td_table_ncols(v);
    td_t* t = td_table_new(nc);
    if (TD_IS_ERR(t)) return t;
    for (int64_t c = 0; c < nc; c++) {
        td_t* col = td_table_get_col_idx(v, c);
        td_t* d;
        if (col->type == TD_MAPCOMMON) {
            td_retain(col);                  /* partition keys: host metadata */
            d = col;
        } else if (TD_IS_PARTED(col->type)) {
            d = put_parted(col);
        } else {
            d = put_vec(col);
        }
        if (TD_IS_ERR(d)) { td_release(t); return d; }
        t = td_table_add_col(t, td_table_col_name(v, c), d);
        td_release(d);
        if (TD_IS_ERR(t)) return t;
    }
    return t;
}

/* Host copy of one resident vector / selection. */
static td_t* get_vec(td_t* v) {
    size_t bytes = payload_bytes(v);
    td_t* h;
    if (v->type == TD_SEL) {
        h = td_sel_new(v->len);
    } else if (v->type == TD_SYM) {
        h = td_sym_vec_new(v->attrs & TD_SYM_W_MASK, v->len);
    } else {
        h = td_vec_new(v->type, v->len);
    }
    if (TD_IS_ERR(h)) return h;
    if (v->type != TD_SEL) h->len = v->len;
    int rc = tdd_d2h(td_data(h), td_hbm_ref(v)->dptr, bytes);
    if (rc) { td_release(h); return TD_ERR_PTR(rc); }
    return h;
}

/* Host copy of a parted column: resident segments copied, host ones held. */
static td_t* get_parted(td_t* v) {
    td_t* out = td_alloc((size_t)v->len * sizeof(td_t*));
    if (!out) return TD_ERR_PTR(TD_ERR_OOM);
    out->type = v->type;
    out->attrs = 0;
    out->len = v->len;
    memset(td_data(out), 0, (size_t)v->len * sizeof(td_t*));
    for (int64_t i = 0; i < v->len; i++) {
        td_t* s = ((td_t**)td_data(v))[i];
        if (!s) continue;
        td_t* h = s->mmod == TD_MMOD_HBM ? get_vec(s) : (td_retain(s), s);
        if (TD_IS_ERR(h)) { td_release(out); return h; }
        ((td_t**)td_data(out))[i] = h;
    }
    return out;
}

static int col_resident(td_t* c) {
    if (c->type == TD_MAPCOMMON) return 1;
    if (TD_IS_PARTED(c->type)) {
        for (int64_t i = 0; i < c->len; i++) {
            td_t* s = ((td_t**)td_data(c))[i];
            if (s && s->mmod != TD_MMOD_HBM) return 0;
        }
        return 1;
    }
    return c->mmod == TD_MMOD_HBM;
}

td_t* td_dev_get(td_t* v) {
    if (!v || TD_IS_ERR(v)) return v;
    if (v->type == TD_TABLE) {
        int64_t nc = td_table_ncols(v);
        td_t* t = td_table_new(nc);
        if (TD_IS_ERR(t)) return t;
        for (int64_t c = 0; c < nc; c++) {
            td_t* col = td_table_get_col_idx(v, c);
            td_t* h = col->mmod == TD_MMOD_HBM ? get_vec(col)
                    : TD_IS_PARTED(col->type) ? get_parted(col)
                    : (td_retain(col), col);
            if (TD_IS_ERR(h)) { td_release(t); return h; }
            t = td_table_add_col(t, td_table_col_name(v, c), h);
            td_release(h);
            if (TD_IS_ERR(t)) return t;
        }
        return t;
    }
    if (TD_IS_PARTED(v->type)) return get_parted(v);
    if (v->mmod != TD_MMOD_HBM) { td_retain(v); return v; }
    return get_vec(v);
}

td_t* td__to_host(td_t* v) {
    if (!v || TD_IS_ERR(v)) return v;
    int need = v->mmod == TD_MMOD_HBM;
    if (v->type == TD_TABLE)
        for (int64_t c = 0; c < td_table_ncols(v) && !need; c++)
            need = td_table_get_col_idx(v, c)->mmod == TD_MMOD_HBM;
    if (!need) return v;
    td_t* h = td_dev_get(v);
    td_release(v);
    return h;
}

int td_dev_is_resident(td_t* v) {
    if (!v || TD_IS_ERR(v)) return 0;
    if (v->type == TD_TABLE) {
        for (int64_t c = 0; c < td_table_ncols(v); c++)
            if (!col_resident(td_table_get_col_idx(v, c))) return 0;
        return 1;
    }
    return col_resident(v);
}

void td_dev_mem_stats(td_dev_mem_t* out) {
    memset(out, 0, sizeof(*out));
    if (!td__dev_ready()) return;
    out->in_use = td__arena_in_use();
    out->peak = td__arena_peak();
    out->reserved = td__arena_reserved();
    out->budget = td__arena_budget();
    uint64_t f = 0, t = 0;
    if (tdd_mem_info(&f, &t) == TD_OK) out->device_total = t;
}

void td_dev_set_budget(uint64_t bytes) {
    if (!td__dev_ready()) return;
    uint64_t f = 0, t = 0;
    if (!bytes && tdd_mem_info(&f, &t) == TD_OK) bytes = (uint64_t)((double)t * 0.9);
    td__arena_set_budget((size_t)bytes);
}

void td_dev_prof_enable(int on) { if (td__dev_ready()) tdd_prof_enable(on); }
void td_dev_prof_reset(void) { if (td__dev_ready()) tdd_prof_reset(); }

int td_dev_prof_read(td_dev_prof_t* rows, int cap) {
    if (!td__dev_ready()) return 0;
    int n = 0;
    tdd_prof_row_t tmp[128];
    if (cap > 128) cap = 128;
    if (tdd_prof_read(tmp, cap, &n) != TD_OK) return 0;
    for (int i = 0; i < n; i++) {
        memcpy(rows[i].name, tmp[i].name, sizeof(rows[i].name));
        rows[i].launches = tmp[i].launches;
        rows[i].total_ms = tmp[i].total_ms;
    }
    return n;
}
