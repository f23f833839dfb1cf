/*
 * vec.c -- vectors, atoms, strings, lists and the morsel iterator.
 *
 * Semantics follow the reference's src/vec/{vec,atom,str,list}.c: vectors are
 * td_t blocks whose capacity comes from the block order, appends grow by
 * doubling, slices are header-only views onto a parent, TD_SYM columns carry
 * their index width in attrs.  Device-resident vectors (mmod 3) are not
 * host-addressable: element access on them reports TD_ERR_TYPE / NULL.
 */
#include "internal.h"

static int64_t capacity_of(const td_t* v) {
    uint8_t esz = td_sym_elem_size(v->type, v->attrs);
    if (!esz) return 0;
    return (int64_t)((((size_t)1 << v->order) - 32) / esz);
}

td_t* td_vec_new(int8_t type, int64_t capacity) {
    if (type <= 0 || type >= TD_TYPE_COUNT) return TD_ERR_PTR(TD_ERR_TYPE);
    if (type == TD_SYM) return td_sym_vec_new(TD_SYM_W64, capacity);
    if (capacity < 0) return TD_ERR_PTR(TD_ERR_RANGE);
    uint8_t esz = td_elem_size(type);
    if (esz > 1 && (uint64_t)capacity > SIZE_MAX / esz) return TD_ERR_PTR(TD_ERR_OOM);
    td_t* v = td_alloc((size_t)capacity * esz);
    if (!v) return TD_ERR_PTR(TD_ERR_OOM);
    v->type = type;
    v->len = 0;
    return v;
}

td_t* td_sym_vec_new(uint8_t sym_width, int64_t capacity) {
    if (sym_width & ~TD_SYM_W_MASK) return TD_ERR_PTR(TD_ERR_TYPE);
    if (capacity < 0) return TD_ERR_PTR(TD_ERR_RANGE);
    uint8_t esz = (uint8_t)TD_SYM_ELEM(sym_width);
    if ((uint64_t)capacity > SIZE_MAX / esz) return TD_ERR_PTR(TD_ERR_OOM);
    td_t* v = td_alloc((size_t)capacity * esz);
    if (!v) return TD_ERR_PTR(TD_ERR_OOM);
    v->type = TD_SYM;
    v->attrs = sym_width;
    v->len = 0;
    return v;
}

td_t* td_vec_from_raw(int8_t type, const void* data, int64_t count) {
    if (type <= 0 || type >= TD_TYPE_COUNT) return TD_ERR_PTR(TD_ERR_TYPE);
    if (count < 0) return TD_ERR_PTR(TD_ERR_RANGE);
    uint8_t w = type == TD_SYM ? TD_SYM_W64 : 0;
    uint8_t esz = td_sym_elem_size(type, w);
    td_t* v = td_alloc((size_t)count * esz);
    if (!v) return TD_ERR_PTR(TD_ERR_OOM);
    v->type = type;
    v->attrs = w;
    v->len = count;
    if (count) memcpy(td_data(v), data, (size_t)count * esz);
    if (type == TD_LIST || type == TD_TABLE) {
        td_t** p = (td_t**)td_data(v);
        for (int64_t i = 0; i < count; i++) if (p[i]) td_retain(p[i]);
    }
    return v;
}

td_t* td_vec_append(td_t* vec, const void* elem) {
    if (!vec || TD_IS_ERR(vec)) return vec;
    if (vec->type <= 0 || vec->type >= TD_TYPE_COUNT || vec->mmod == TD_MMOD_HBM)
        return TD_ERR_PTR(TD_ERR_TYPE);
    vec = td_cow(vec);
    if (!vec || TD_IS_ERR(vec)) return vec;
    uint8_t esz = td__esz(vec);
    if (vec->len >= capacity_of(vec)) {
        size_t want = (size_t)(vec->len + 1) * esz, s = 32;
        while (s < want) s *= 2;
        td_t* n = td_scratch_realloc(vec, s);
        if (!n || TD_IS_ERR(n)) return n;
        vec = n;
    }
    memcpy((char*)td_data(vec) + vec->len * esz, elem, esz);
    vec->len++;
    return vec;
}

td_t* td_vec_set(td_t* vec, int64_t idx, const void* elem) {
    if (!vec || TD_IS_ERR(vec)) return vec;
    if (vec->mmod == TD_MMOD_HBM) return TD_ERR_PTR(TD_ERR_TYPE);
    if (idx < 0 || idx >= vec->len) return TD_ERR_PTR(TD_ERR_RANGE);
    vec = td_cow(vec);
    if (!vec || TD_IS_ERR(vec)) return vec;
    uint8_t esz = td__esz(vec);
    memcpy((char*)td_data(vec) + idx * esz, elem, esz);
    return vec;
}

void* td_vec_get(td_t* vec, int64_t idx) {
    if (!vec || TD_IS_ERR(vec) || vec->mmod == TD_MMOD_HBM) return NULL;
    if (idx < 0 || idx >= vec->len) return NULL;
    if (vec->attrs & TD_ATTR_SLICE) {
        td_t* p = vec->slice_parent;
        return (char*)td_data(p) + (vec->slice_offset + idx) * td__esz(p);
    }
    return (char*)td_data(vec) + idx * td__esz(vec);
}

/* A header-only view of elements [offset, offset + len): it records the
 * vector that owns the bytes (for a view of a view, that view's owner) and
 * holds a reference on it. */
td_t* td_vec_slice(td_t* vec, int64_t offset, int64_t len) {
    if (!vec || TD_IS_ERR(vec)) return vec;
    const int in_range = offset >= 0 && len >= 0 && offset <= vec->len &&
                         len <= vec->len - offset;
    if (!in_range) return TD_ERR_PTR(TD_ERR_RANGE);
    const int of_view = (vec->attrs & TD_ATTR_SLICE) != 0;
    td_t* owner = of_view ? vec->slice_parent : vec;
    td_t* view = td_alloc(0);
    if (!view) return TD_ERR_PTR(TD_ERR_OOM);
    view->type = owner->type;
    view->attrs = (uint8_t)((owner->attrs & TD_SYM_W_MASK) | TD_ATTR_SLICE);
    view->len = len;
    view->slice_offset = (of_view ? vec->slice_offset : 0) + offset;
    view->slice_parent = owner;
    td_retain(owner);
    return view;
}

td_t* td_vec_concat(td_t* a, td_t* b) {
    if (!a || TD_IS_ERR(a)) return a;
    if (!b || TD_IS_ERR(b)) return b;
    if (a->type != b->type || a->mmod == TD_MMOD_HBM || b->mmod == TD_MMOD_HBM)
        return TD_ERR_PTR(TD_ERR_TYPE);
    uint8_t ea = td__esz(a), eb = td__esz(b);
    uint8_t attrs = ea >= eb ? a->attrs : b->attrs;
    uint8_t esz = ea >= eb ? ea : eb;
    int64_t n = a->len + b->len;
    td_t* r = td_alloc((size_t)n * esz);
    if (!r) return TD_ERR_PTR(TD_ERR_OOM);
    r->type = a->type;
    r->attrs = attrs & (uint8_t)~TD_ATTR_SLICE;
    r->len = n;
    for (int side = 0; side < 2; side++) {
        td_t* s = side ? b : a;
        int64_t base = side ? a->len : 0;
        for (int64_t i = 0; i < s->len; i++) {
            const void* src = td_vec_get(s, i);
            if (s->type == TD_SYM && td__esz(s) != esz)
                td_write_sym(td_data(r), base + i,
                             (uint64_t)td_read_sym(src, 0, TD_SYM, s->attrs), TD_SYM, r->attrs);
            else
                memcpy((char*)td_data(r) + (base + i) * esz, src, esz);
        }
    }
    if (r->type == TD_LIST || r->type == TD_TABLE) {
        td_t** p = (td_t**)td_data(r);
        for (int64_t i = 0; i < n; i++) if (p[i]) td_retain(p[i]);
    }
    return r;
}

/* Null bitmaps: inline in nullmap[] for <= 128 rows, else an external U8
 * vector (td.h attrs TD_ATTR_NULLMAP_EXT).  The filter/join/group path has no
 * null semantics (the reference never reads these bitmaps there). */
void td_vec_set_null(td_t* vec, int64_t idx, bool is_null) {
    if (!vec || TD_IS_ERR(vec) || (vec->attrs & TD_ATTR_SLICE) || vec->mmod == TD_MMOD_HBM) return;
    if (idx < 0 || idx >= vec->len) return;
    if (is_null) vec->attrs |= TD_ATTR_HAS_NULLS;
    uint8_t* bits;
    if (vec->len <= 128 && !(vec->attrs & TD_ATTR_NULLMAP_EXT)) {
        bits = vec->nullmap;
    } else {
        if (!(vec->attrs & TD_ATTR_NULLMAP_EXT)) {
            td_t* ext = td_vec_new(TD_U8, (vec->len + 7) / 8);
            if (!ext || TD_IS_ERR(ext)) return;
            ext->len = (vec->len + 7) / 8;
            memset(td_data(ext), 0, (size_t)ext->len);
            if (vec->len <= 128) memcpy(td_data(ext), vec->nullmap, 16);
            vec->ext_nullmap = ext;
            vec->attrs |= TD_ATTR_NULLMAP_EXT;
        }
        bits = (uint8_t*)td_data(vec->ext_nullmap);
    }
    if (is_null) bits[idx >> 3] |= (uint8_t)(1u << (idx & 7));
    else bits[idx >> 3] &= (uint8_t)~(1u << (idx & 7));
}

bool td_vec_is_null(td_t* vec, int64_t idx) {
    if (!vec || TD_IS_ERR(vec) || !(vec->attrs & TD_ATTR_HAS_NULLS)) return false;
    if (idx < 0 || idx >= vec->len) return false;
    const uint8_t* bits = (vec->attrs & TD_ATTR_NULLMAP_EXT)
                              ? (const uint8_t*)td_data(vec->ext_nullmap) : vec->nullmap;
    return (bits[idx >> 3] >> (idx & 7)) & 1;
}

/* ---------------------------------------------------------------- atoms -- */

static td_t* atom(int8_t t) {
    td_t* v = td_alloc(0);
    if (!v) return TD_ERR_PTR(TD_ERR_OOM);
    v->type = t;
    return v;
}
td_t* td_bool(bool x)      { td_t* v = atom(TD_ATOM_BOOL); if (!TD_IS_ERR(v)) v->b8 = x ? 1 : 0; return v; }
td_t* td_u8(uint8_t x)     { td_t* v = atom(TD_ATOM_U8);   if (!TD_IS_ERR(v)) v->u8 = x; return v; }
td_t* td_char(char x)      { td_t* v = atom(TD_ATOM_CHAR); if (!TD_IS_ERR(v)) v->c8 = x; return v; }
td_t* td_i16(int16_t x)    { td_t* v = atom(TD_ATOM_I16);  if (!TD_IS_ERR(v)) v->i16 = x; return v; }
td_t* td_i32(int32_t x)    { td_t* v = atom(TD_ATOM_I32);  if (!TD_IS_ERR(v)) v->i32 = x; return v; }
td_t* td_i64(int64_t x)    { td_t* v = atom(TD_ATOM_I64);  if (!TD_IS_ERR(v)) v->i64 = x; return v; }
td_t* td_f64(double x)     { td_t* v = atom(TD_ATOM_F64);  if (!TD_IS_ERR(v)) v->f64 = x; return v; }
td_t* td_sym(int64_t id)   { td_t* v = atom(TD_ATOM_SYM);  if (!TD_IS_ERR(v)) v->i64 = id; return v; }
td_t* td_date(int64_t x)   { td_t* v = atom(TD_ATOM_DATE); if (!TD_IS_ERR(v)) v->i64 = x; return v; }
td_t* td_time(int64_t x)   { td_t* v = atom(TD_ATOM_TIME); if (!TD_IS_ERR(v)) v->i64 = x; return v; }
td_t* td_timestamp(int64_t x) { td_t* v = atom(TD_ATOM_TIMESTAMP); if (!TD_IS_ERR(v)) v->i64 = x; return v; }

td_t* td_guid(const uint8_t* bytes) {
    td_t* v = atom(TD_ATOM_GUID);
    if (TD_IS_ERR(v)) return v;
    td_t* g = td_vec_new(TD_U8, 16);
    if (!g || TD_IS_ERR(g)) { td_free(v); return g; }
    g->len = 16;
    memcpy(td_data(g), bytes, 16);
    v->obj = g;
    return v;
}

/* Strings shorter than 7 bytes live inline (slen + sdata, NUL-terminated);
 * longer ones in a CHAR vector held by obj.  slen and the low byte of obj
 * share storage: a 32-byte-aligned pointer never has a low byte in 1..7, so
 * slen in 1..7, or all-zero bytes, means inline (the reference's
 * src/vec/str.c:29-45 uses the same encoding). */
static int str_inline(const td_t* s) {
    return (s->slen >= 1 && s->slen <= 7) || (s->slen == 0 && s->obj == NULL);
}

td_t* td_str(const char* s, size_t len) {
    td_t* v = atom(TD_ATOM_STR);
    if (TD_IS_ERR(v)) return v;
    if (len < 7) {
        v->slen = (uint8_t)len;
        if (len) memcpy(v->sdata, s, len);
        v->sdata[len] = 0;
        return v;
    }
    td_t* c = td_alloc(len + 1);
    if (!c) { td_free(v); return TD_ERR_PTR(TD_ERR_OOM); }
    c->type = TD_CHAR;
    c->len = (int64_t)len;
    memcpy(td_data(c), s, len);
    ((char*)td_data(c))[len] = 0;
    v->obj = c;
    return v;
}

const char* td_str_ptr(td_t* s) {
    if (!s || TD_IS_ERR(s) || s->type != TD_ATOM_STR) return NULL;
    return str_inline(s) ? s->sdata : (const char*)td_data(s->obj);
}

size_t td_str_len(td_t* s) {
    if (!s || TD_IS_ERR(s) || s->type != TD_ATOM_STR) return 0;
    return str_inline(s) ? s->slen : (size_t)s->obj->len;
}

int td_str_cmp(td_t* a, td_t* b) {
    size_t la = td_str_len(a), lb = td_str_len(b);
    int c = memcmp(td_str_ptr(a), td_str_ptr(b), la < lb ? la : lb);
    if (c) return c;
    return la < lb ? -1 : la > lb;
}

/* ---------------------------------------------------------------- lists -- */

td_t* td_list_new(int64_t capacity) {
    if (capacity < 0) return TD_ERR_PTR(TD_ERR_RANGE);
    td_t* l = td_alloc((size_t)capacity * sizeof(td_t*));
    if (!l) return TD_ERR_PTR(TD_ERR_OOM);
    l->type = TD_LIST;
    l->len = 0;
    return l;
}

td_t* td_list_append(td_t* list, td_t* item) {
    td_t* r = td_vec_append(list, &item);
    if (r && !TD_IS_ERR(r) && item) td_retain(item);
    return r;
}

td_t* td_list_get(td_t* list, int64_t idx) {
    if (!list || TD_IS_ERR(list) || idx < 0 || idx >= list->len) return NULL;
    return ((td_t**)td_data(list))[idx];
}

td_t* td_list_set(td_t* list, int64_t idx, td_t* item) {
    if (!list || TD_IS_ERR(list)) return list;
    if (idx < 0 || idx >= list->len) return TD_ERR_PTR(TD_ERR_RANGE);
    list = td_cow(list);
    if (!list || TD_IS_ERR(list)) return list;
    td_t** p = (td_t**)td_data(list);
    if (item) td_retain(item);
    if (p[idx]) td_release(p[idx]);
    p[idx] = item;
    return list;
}

/* --------------------------------------------------------------- morsels -- */

void td_morsel_init_range(td_morsel_t* m, td_t* vec, int64_t start, int64_t end) {
    m->vec = vec;
    m->offset = start;
    m->len = end;
    m->elem_size = vec ? td__esz(vec) : 0;
    m->morsel_len = 0;
    m->morsel_ptr = NULL;
    m->null_bits = NULL;
}

void td_morsel_init(td_morsel_t* m, td_t* vec) {
    td_morsel_init_range(m, vec, 0, vec ? vec->len : 0);
}

bool td_morsel_next(td_morsel_t* m) {
    if (!m->vec || m->vec->mmod == TD_MMOD_HBM || m->offset >= m->len) return false;
    if (m->morsel_ptr) m->offset += m->morsel_len;
    if (m->offset >= m->len) return false;
    int64_t n = m->len - m->offset;
    m->morsel_len = n < TD_MORSEL_ELEMS ? n : TD_MORSEL_ELEMS;
    m->morsel_ptr = td_vec_get(m->vec, m->offset);
    return true;
}
