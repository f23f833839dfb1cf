/*
 * sel.c -- the td.h selection API (TD_SEL bitmaps).
 *
 * Layout (td.h:546-589, the reference's src/vec/sel.c:37-47): meta
 * {total_pass, n_segs} | seg_flags[n_segs] padded to 8 | seg_popcnt[n_segs]
 * padded to 8 | bits[ceil(n/64)].  td_sel_new only allocates; building,
 * AND-ing and recomputing run on the GPU (tdd_sel_*), with host selections
 * copied in and out and device-resident selections (mmod 3) used in place.
 */
#include "internal.h"

size_t td__sel_payload_bytes(int64_t nrows) {
    size_t off_flags, off_pop, off_bits;
    td__sel_layout(nrows, &off_flags, &off_pop, &off_bits);
    return off_bits + (size_t)((nrows + 63) / 64) * 8;
}

void td__sel_layout(int64_t nrows, size_t* off_flags, size_t* off_pop, size_t* off_bits) {
    size_t n_segs = (size_t)((nrows + TD_MORSEL_ELEMS - 1) / TD_MORSEL_ELEMS);
    *off_flags = sizeof(td_sel_meta_t);
    *off_pop = *off_flags + ((n_segs + 7) & ~(size_t)7);
    *off_bits = *off_pop + ((n_segs * 2 + 7) & ~(size_t)7);
}

td_t* td_sel_new(int64_t nrows) {
    if (nrows < 0) return TD_ERR_PTR(TD_ERR_RANGE);
    size_t bytes = td__sel_payload_bytes(nrows);
    td_t* s = td_alloc(bytes);
    if (!s) return TD_ERR_PTR(TD_ERR_OOM);
    s->type = TD_SEL;
    s->len = nrows;
    memset(td_data(s), 0, bytes);
    td_sel_meta(s)->n_segs = (uint32_t)((nrows + TD_MORSEL_ELEMS - 1) / TD_MORSEL_ELEMS);
    return s;
}

/* Device view of a TD_SEL payload at device address base. */
tdd_sel_t td__sel_view(void* base, int64_t nrows) {
    size_t of, op, ob;
    td__sel_layout(nrows, &of, &op, &ob);
    tdd_sel_t v;
    v.base = base;
    v.flags = (uint8_t*)base + of;
    v.popcnt = (uint16_t*)((uint8_t*)base + op);
    v.bits = (uint64_t*)((uint8_t*)base + ob);
    v.nrows = nrows;
    v.n_segs = (uint32_t)((nrows + TD_MORSEL_ELEMS - 1) / TD_MORSEL_ELEMS);
    v.pad = 0;
    return v;
}

/* A new device-resident TD_SEL (mmod 3) for nrows. */
td_t* td__sel_new_dev(int64_t nrows) {
    return td__hbm_vec_new(TD_SEL, 0, nrows, td__sel_payload_bytes(nrows));
}

/* Device address of a selection's payload; host selections are copied into
 * a temporary arena block (*tmp, to be freed by the caller). */
void* td__sel_dev_payload(td_t* s, void** tmp) {
    *tmp = NULL;
    if (s->mmod == TD_MMOD_HBM) return td_hbm_ref(s)->dptr;
    size_t bytes = td__sel_payload_bytes(s->len);
    void* d = td__arena_alloc(bytes);
    if (!d) return NULL;
    if (tdd_h2d(d, td_data(s), bytes) != TD_OK) { td__arena_free(d); return NULL; }
    *tmp = d;
    return d;
}

/* Result selection shaped like the inputs: device-resident when `resident`,
 * else a host copy of the device payload `dev`. */
static td_t* finish_sel(void* dev, int64_t nrows, int resident, td_t* dev_owner) {
    if (resident) return dev_owner;
    td_t* h = td_sel_new(nrows);
    if (TD_IS_ERR(h)) { td_release(dev_owner); return h; }
    int rc = tdd_d2h(td_data(h), dev, td__sel_payload_bytes(nrows));
    td_release(dev_owner);
    if (rc) { td_release(h); return TD_ERR_PTR(rc); }
    return h;
}

td_t* td_sel_from_pred(td_t* pred) {
    if (!pred || TD_IS_ERR(pred)) return pred;
    if (pred->type != TD_BOOL) return TD_ERR_PTR(TD_ERR_TYPE);
    if (!td__dev_ready()) return TD_ERR_PTR(TD_ERR_IO);
    int64_t n = pred->len;
    td_t* out = td__sel_new_dev(n);
    if (TD_IS_ERR(out)) return out;
    void* src = NULL;
    void* tmp = NULL;
    if (pred->mmod == TD_MMOD_HBM) {
        src = td_hbm_ref(pred)->dptr;
    } else {
        tmp = td__arena_alloc((size_t)(n > 0 ? n : 1));
        if (!tmp) { td_release(out); return TD_ERR_PTR(TD_ERR_OOM); }
        const void* hsrc = (pred->attrs & TD_ATTR_SLICE) ? td_vec_get(pred, 0) : td_data(pred);
        if (n && tdd_h2d(tmp, hsrc, (size_t)n) != TD_OK) {
            td__arena_free(tmp);
            td_release(out);
            return TD_ERR_PTR(TD_ERR_IO);
        }
        src = tmp;
    }
    tdd_sel_t v = td__sel_view(td_hbm_ref(out)->dptr, n);
    int rc = tdd_sel_from_bool((const uint8_t*)src, &v);
    if (!rc) rc = tdd_sync();
    td__arena_free(tmp);
    if (rc) { td_release(out); return TD_ERR_PTR(rc); }
    return finish_sel(td_hbm_ref(out)->dptr, n, pred->mmod == TD_MMOD_HBM, out);
}

td_t* td_sel_and(td_t* a, td_t* b) {
    if (!a || TD_IS_ERR(a)) return a;
    if (!b || TD_IS_ERR(b)) return b;
    if (a->type != TD_SEL || b->type != TD_SEL) return TD_ERR_PTR(TD_ERR_TYPE);
    if (a->len != b->len) return TD_ERR_PTR(TD_ERR_RANGE);
    if (!td__dev_ready()) return TD_ERR_PTR(TD_ERR_IO);
    int64_t n = a->len;
    td_t* out = td__sel_new_dev(n);
    if (TD_IS_ERR(out)) return out;
    void *ta = NULL, *tb = NULL;
    void* da = td__sel_dev_payload(a, &ta);
    void* db = td__sel_dev_payload(b, &tb);
    int rc = (!da || !db) ? TD_ERR_OOM : TD_OK;
    if (!rc) {
        tdd_sel_t va = td__sel_view(da, n), vb = td__sel_view(db, n);
        tdd_sel_t vo = td__sel_view(td_hbm_ref(out)->dptr, n);
        rc = tdd_sel_and(&va, &vb, &vo);
        if (!rc) rc = tdd_sync();
    }
    td__arena_free(ta);
    td__arena_free(tb);
    if (rc) { td_release(out); return TD_ERR_PTR(rc); }
    int resident = a->mmod == TD_MMOD_HBM && b->mmod == TD_MMOD_HBM;
    return finish_sel(td_hbm_ref(out)->dptr, n, resident, out);
}

void td_sel_recompute(td_t* sel) {
    if (!sel || TD_IS_ERR(sel) || sel->type != TD_SEL || !td__dev_ready()) return;
    void* tmp = NULL;
    void* d = td__sel_dev_payload(sel, &tmp);
    if (!d) return;
    tdd_sel_t v = td__sel_view(d, sel->len);
    int rc = tdd_sel_recompute(&v);
    if (!rc && tmp) rc = tdd_d2h(td_data(sel), tmp, td__sel_payload_bytes(sel->len));
    if (!rc) tdd_sync();
    td__arena_free(tmp);
}
