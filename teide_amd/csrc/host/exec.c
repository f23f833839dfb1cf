/*
 * exec.c -- td_execute: maps a td.h query DAG onto the device layer.
 *
 * The node semantics are the reference's exec_node (src/ops/exec.c:11561-
 * 12321); every operator on the filter -> join -> group-by path runs on the
 * GPU through tdd.h and anything else reports TD_ERR_NYI, so a query either
 * runs natively or fails loudly.  Supported shapes:
 *
 *   FILTER(table, P)        P (comparisons of columns with constants or
 *                           columns under AND/OR/NOT) is evaluated over the
 *                           bound table into a device TD_SEL, ANDed into
 *                           g->selection (exec.c:11706-11731); the table is
 *                           compacted by td_execute's final step
 *                           (exec.c:12312-12319).
 *   FILTER(vector, P)       eager compaction of the vector (exec.c:1978-2002).
 *   GROUP(keys; aggs)       over g->table, masked by g->selection, which the
 *                           group consumes.  (The reference compacts the group
 *                           output against the leftover selection as well,
 *                           reading past the output's end; here a selection
 *                           always means "rows of the bound table".)
 *                           Keys and aggregate inputs of the form FILTER(SCAN
 *                           c, P) with one shared P (the README quick-start
 *                           shape) mean "filter by P, then group".
 *   JOIN(L, R)              L masked by g->selection (the reference compacts
 *                           it first, exec.c:11846-11852); output columns are
 *                           all left columns then the right non-key columns
 *                           (exec.c:8607-8728).
 *   GROUP over a JOIN       when the group's columns are missing from
 *                           g->table but present in the output of the graph's
 *                           JOIN(FILTER*(CONST table), CONST table), the group
 *                           runs over the join output, fused on the device
 *                           (the reference connects EXPAND -> GROUP the same
 *                           way, exec.c:11766-11800).
 *   SUM/COUNT/MIN/MAX/AVG/FIRST/LAST/VAR.. of a vector; ALIAS, MATERIALIZE.
 *
 * Residency: when every column a query reads is device-resident its result
 * stays device-resident; otherwise inputs are copied in for the call and
 * the result comes back as ordinary host vectors.
 */
#include "internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#define MAX_TMP 256

#define MAX_UPLOADS 64

typedef struct {
    td_graph_t* g;
    int   all_resident;
    void* tmp[MAX_TMP];
    int   ntmp;
    /* host columns uploaded during this execution, so a column used by a
     * predicate and an aggregate is one device copy (each held, so its
     * address cannot be reused while the entry lives) */
    td_t* up_v[MAX_UPLOADS];
    void* up_d[MAX_UPLOADS];
    int   nup;
    /* parted tables of this execution and their flat forms (held) */
    td_t* fl_src[8];
    td_t* fl_dst[8];
    int   nfl;
} xctx_t;

static td_t* xnode(xctx_t* x, td_op_t* op);

static int keep_tmp(xctx_t* x, void* d) {
    if (x->ntmp >= MAX_TMP) { td__arena_free(d); return TD_ERR_OOM; }
    x->tmp[x->ntmp++] = d;
    return TD_OK;
}

static void drop_tmps(xctx_t* x) {
    for (int i = 0; i < x->ntmp; i++) td__arena_free(x->tmp[i]);
    x->ntmp = 0;
    for (int i = 0; i < x->nup; i++) td_release(x->up_v[i]);
    x->nup = 0;
    for (int i = 0; i < x->nfl; i++) td_release(x->fl_dst[i]);
    x->nfl = 0;
}

/* ------------------------------------------------------ parted tables -- */

/* A table of parted columns (one vector per partition) and a MAPCOMMON
 * partition-key column queries as the concatenation of its partitions
 * with the key repeated over each partition's rows, as the reference's
 * concatenating paths read it (exec.c:201-241, 5443-5537, 11581-11608).
 * Here the concatenation is a flat device column: in place when the
 * segments are consecutive views of one arena block (td_dev_put,
 * td_dev_parted_load), else copied into a new block; the key column is
 * filled on the device. */
static int has_parted(td_t* t) {
    for (int64_t c = 0; t && t->type == TD_TABLE && c < td_table_ncols(t); c++) {
        td_t* col = td_table_get_col_idx(t, c);
        if (col && (TD_IS_PARTED(col->type) || col->type == TD_MAPCOMMON)) return 1;
    }
    return 0;
}

static td_t* flat_parted(xctx_t* x, td_t* v) {
    td_t** segs = (td_t**)td_data(v);
    int8_t base = (int8_t)TD_PARTED_BASETYPE(v->type);
    uint8_t attrs = 0;
    for (int64_t i = 0; i < v->len; i++)
        if (segs[i]) { attrs = segs[i]->attrs & TD_SYM_W_MASK; break; }
    uint8_t esz = td_sym_elem_size(base, attrs);
    int64_t total = td_parted_nrows(v);
    td_t* own = NULL;
    const char* next = NULL;
    int inplace = 1;
    for (int64_t i = 0; i < v->len && inplace; i++) {
        td_t* s = segs[i];
        if (!s) continue;
        if (s->mmod != TD_MMOD_HBM || !td_hbm_ref(s)->owner) { inplace = 0; break; }
        if (!own) {
            own = td_hbm_ref(s)->owner;
            next = (const char*)td_hbm_ref(own)->dptr;
        }
        inplace = td_hbm_ref(s)->owner == own && (const char*)td_hbm_ref(s)->dptr == next &&
                  td__esz(s) == esz;
        next += (size_t)s->len * esz;
    }
    if (inplace && own && own->len == total && own->type == base && td__esz(own) == esz) {
        td_retain(own);
        return own;
    }
    td_t* d = td__hbm_vec_new(base, attrs, total, (size_t)total * esz);
    if (TD_IS_ERR(d)) return d;
    int64_t at = 0;
    int rc = TD_OK;
    for (int64_t i = 0; i < v->len && !rc; i++) {
        td_t* s = segs[i];
        if (!s || s->len == 0) continue;
        if (s->type != base || td__esz(s) != esz) { rc = TD_ERR_TYPE; break; }
        char* dst = (char*)td_hbm_ref(d)->dptr + (size_t)at * esz;
        size_t bytes = (size_t)s->len * esz;
        td_t* src = s;
        size_t off = 0;
        if (s->mmod != TD_MMOD_HBM && (s->attrs & TD_ATTR_SLICE)) {
            src = s->slice_parent;
            off = (size_t)s->slice_offset * esz;
        }
        if (src->mmod == TD_MMOD_HBM) {
            rc = tdd_d2d(dst, (const char*)td_hbm_ref(src)->dptr + off, bytes);
        } else {
            x->all_resident = 0;                  /* a host partition: results go back */
            rc = tdd_h2d(dst, (const char*)td_data(src) + off, bytes);
        }
        at += s->len;
    }
    if (!rc) rc = td__dev_stats(d);
    if (rc) { td_release(d); return TD_ERR_PTR(rc); }
    return d;
}

static td_t* flat_mapcommon(xctx_t* x, td_t* v) {
    td_t* kv = ((td_t**)td_data(v))[0];
    td_t* rc_v = ((td_t**)td_data(v))[1];
    if (!kv || !rc_v || kv->len != rc_v->len || kv->len < 1 || rc_v->type != TD_I64)
        return TD_ERR_PTR(TD_ERR_CORRUPT);
    int64_t n = kv->len, total = 0;
    uint8_t esz = td__esz(kv);
    int64_t* offs = malloc((size_t)(n + 1) * sizeof(int64_t));
    if (!offs) return TD_ERR_PTR(TD_ERR_OOM);
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t p = 0; p < n; p++) {
        offs[p] = total;
        total += ((const int64_t*)td_data(rc_v))[p];
        int64_t k = kv->type == TD_DATE || kv->type == TD_I32 ? ((const int32_t*)td_data(kv))[p]
                                                              : ((const int64_t*)td_data(kv))[p];
        if (k < lo) lo = k;
        if (k > hi) hi = k;
    }
    offs[n] = total;
    td_t* d = td__hbm_vec_new(kv->type, kv->attrs & TD_SYM_W_MASK, total, (size_t)total * esz);
    void* meta = TD_IS_ERR(d) ? NULL : td__arena_alloc((size_t)n * 8 + (size_t)(n + 1) * 8);
    int rc = TD_IS_ERR(d) ? TD_ERR_CODE(d) : meta ? keep_tmp(x, meta) : TD_ERR_OOM;
    if (!rc) rc = tdd_h2d(meta, td_data(kv), (size_t)n * esz);
    if (!rc) rc = tdd_h2d((char*)meta + (size_t)n * 8, offs, (size_t)(n + 1) * sizeof(int64_t));
    if (!rc) rc = tdd_mapcommon_fill(td_hbm_ref(d)->dptr, esz, meta,
                                     (const int64_t*)((char*)meta + (size_t)n * 8), (int)n, total);
    /* the staged keys and offsets are read before this call returns */
    if (!rc) rc = tdd_sync();
    free(offs);
    if (rc) {
        if (!TD_IS_ERR(d)) td_release(d);
        return TD_ERR_PTR(rc);
    }
    td_hbm_ref(d)->lo = lo;
    td_hbm_ref(d)->hi = hi;
    td_hbm_ref(d)->ranged = 1;
    return d;
}

/* The flat form of a table with parted columns (the table itself, held,
 * when it has none), made once per execution. */
static td_t* flat_table(xctx_t* x, td_t* t) {
    if (!has_parted(t)) { td_retain(t); return t; }
    for (int i = 0; i < x->nfl; i++)
        if (x->fl_src[i] == t) { td_retain(x->fl_dst[i]); return x->fl_dst[i]; }
    if (x->nfl >= 8) return TD_ERR_PTR(TD_ERR_NYI);
    int64_t nc = td_table_ncols(t);
    td_t* out = td_table_new(nc);
    for (int64_t c = 0; c < nc && !TD_IS_ERR(out); c++) {
        td_t* col = td_table_get_col_idx(t, c);
        td_t* f = col->type == TD_MAPCOMMON ? flat_mapcommon(x, col)
                : TD_IS_PARTED(col->type) ? flat_parted(x, col)
                : (td_retain(col), col);
        if (TD_IS_ERR(f)) { td_release(out); return f; }
        out = td_table_add_col(out, td_table_col_name(t, c), f);
        td_release(f);
    }
    if (TD_IS_ERR(out)) return out;
    x->fl_src[x->nfl] = t;
    x->fl_dst[x->nfl++] = out;
    td_retain(out);
    return out;
}

/* Device view of a column: resident columns in place, slices as an offset
 * into their parent, host columns copied into a temporary arena block. */
static int dev_col(xctx_t* x, td_t* v, tdd_col_t* out) {
    memset(out, 0, sizeof(*out));
    if (!v || TD_IS_ERR(v)) return TD_ERR_SCHEMA;
    if (v->type <= 0 || v->type == TD_TABLE || v->type == TD_LIST || v->type == TD_SEL ||
        TD_IS_PARTED(v->type) || v->type == TD_MAPCOMMON)
        return TD_ERR_NYI;
    out->type = v->type;
    out->attrs = v->attrs & TD_SYM_W_MASK;
    out->esz = td__esz(v);
    if (v->mmod == TD_MMOD_HBM) {
        td_hbm_ref_t* r = td_hbm_ref(v);
        out->data = r->dptr;
        out->ranged = (uint8_t)(r->ranged != 0);
        out->lo = r->lo;
        out->hi = r->hi;
        return TD_OK;
    }
    td_t* src = v;
    int64_t off = 0;
    if (v->attrs & TD_ATTR_SLICE) {
        src = v->slice_parent;
        off = v->slice_offset;
        if (src->mmod == TD_MMOD_HBM) {
            td_hbm_ref_t* r = td_hbm_ref(src);
            out->data = (char*)r->dptr + off * out->esz;
            out->ranged = (uint8_t)(r->ranged != 0);   /* parent bounds hold for a slice */
            out->lo = r->lo;
            out->hi = r->hi;
            return TD_OK;
        }
    }
    x->all_resident = 0;
    for (int i = 0; i < x->nup; i++)
        if (x->up_v[i] == v) { out->data = x->up_d[i]; return TD_OK; }
    size_t bytes = (size_t)v->len * out->esz;
    void* d = td__arena_alloc(bytes ? bytes : 8);
    if (!d) return TD_ERR_OOM;
    int rc = keep_tmp(x, d);
    if (rc) return rc;
    if (bytes && tdd_h2d(d, (const char*)td_data(src) + off * out->esz, bytes) != TD_OK)
        return TD_ERR_IO;
    out->data = d;
    if (x->nup < MAX_UPLOADS) {
        td_retain(v);
        x->up_v[x->nup] = v;
        x->up_d[x->nup++] = d;
    }
    return TD_OK;
}

/* Device view of a selection (host selections copied in). */
static int dev_sel(xctx_t* x, td_t* s, tdd_sel_t* out) {
    void* tmp = NULL;
    void* d = td__sel_dev_payload(s, &tmp);
    if (!d) return TD_ERR_OOM;
    if (tmp) {
        x->all_resident = 0;
        int rc = keep_tmp(x, tmp);
        if (rc) return rc;
    }
    *out = td__sel_view(d, s->len);
    return TD_OK;
}

static td_t* scan_table(td_graph_t* g, td_op_ext_t* e) {
    uint16_t stored = 0;
    memcpy(&stored, e->base.pad, sizeof(stored));
    if (stored && g->tables && stored - 1 < g->n_tables) return g->tables[stored - 1];
    return g->table;
}

/* Column a SCAN node names, in `tbl` (or its own table when tbl is NULL). */
static td_t* scan_col(td_graph_t* g, td_op_t* op, td_t* tbl) {
    td_op_ext_t* e = td__find_ext(g, op->id);
    if (!e || e->base.opcode != OP_SCAN) return NULL;
    if (!tbl) tbl = scan_table(g, e);
    return tbl ? td_table_get_col(tbl, e->sym) : NULL;
}

/* ---------------------------------------------------------- predicates -- */

static int mirror_cmp(int op) {
    switch (op) {
    case OP_LT: return OP_GT;
    case OP_LE: return OP_GE;
    case OP_GT: return OP_LT;
    case OP_GE: return OP_LE;
    default: return op;
    }
}

/* Numeric value of an atom (exec.c atom_to_numeric): 1 when f64. */
static int atom_num(td_t* a, int64_t* i, double* f, int* is_f) {
    switch (a->type) {
    case TD_ATOM_F64: *f = a->f64; *i = (int64_t)a->f64; *is_f = 1; return 1;
    case TD_ATOM_I64: case TD_ATOM_SYM: case TD_ATOM_TIMESTAMP:
    case TD_ATOM_DATE: case TD_ATOM_TIME:
        *i = a->i64; *f = (double)a->i64; *is_f = 0; return 1;
    case TD_ATOM_I32: *i = a->i32; *f = a->i32; *is_f = 0; return 1;
    case TD_ATOM_I16: *i = a->i16; *f = a->i16; *is_f = 0; return 1;
    case TD_ATOM_U8: case TD_ATOM_BOOL: *i = a->u8; *f = a->u8; *is_f = 0; return 1;
    default: return 0;
    }
}

typedef struct {
    xctx_t*     x;
    td_t*       tbl;
    tdd_pred_t* P;
    td_t*       cols[TDD_PRED_MAX_COLS];
} pcomp_t;

static int pred_col(pcomp_t* c, td_t* col) {
    for (int i = 0; i < c->P->n_cols; i++) if (c->cols[i] == col) return i;
    if (c->P->n_cols >= TDD_PRED_MAX_COLS) return -TD_ERR_NYI;
    int rc = dev_col(c->x, col, &c->P->cols[c->P->n_cols]);
    if (rc) return -rc;
    c->cols[c->P->n_cols] = col;
    return c->P->n_cols++;
}

static tdd_pred_ins_t* emit(pcomp_t* c) {
    if (c->P->n_ins >= TDD_PRED_MAX_INS) return NULL;
    tdd_pred_ins_t* in = &c->P->ins[c->P->n_ins++];
    memset(in, 0, sizeof(*in));
    return in;
}

static int pred_compile(pcomp_t* c, td_op_t* op) {
    td_graph_t* g = c->x->g;
    tdd_pred_ins_t* in;
    switch (op->opcode) {
    case OP_AND: case OP_OR: {
        int rc = pred_compile(c, op->inputs[0]);
        if (!rc) rc = pred_compile(c, op->inputs[1]);
        if (rc) return rc;
        if (!(in = emit(c))) return TD_ERR_NYI;
        in->kind = op->opcode == OP_AND ? TDD_P_AND : TDD_P_OR;
        return TD_OK;
    }
    case OP_NOT: {
        int rc = pred_compile(c, op->inputs[0]);
        if (rc) return rc;
        if (!(in = emit(c))) return TD_ERR_NYI;
        in->kind = TDD_P_NOT;
        return TD_OK;
    }
    case OP_SCAN: {                                  /* a BOOL column as predicate */
        td_t* col = scan_col(g, op, c->tbl);
        if (!col) return TD_ERR_SCHEMA;
        if (col->type != TD_BOOL && col->type != TD_U8) return TD_ERR_TYPE;
        int ci = pred_col(c, col);
        if (ci < 0) return -ci;
        if (!(in = emit(c))) return TD_ERR_NYI;
        in->kind = TDD_P_BOOLCOL;
        in->a = (uint8_t)ci;
        return TD_OK;
    }
    case OP_CONST: {
        td_op_ext_t* e = td__find_ext(g, op->id);
        int64_t i; double f; int isf;
        if (!e || !e->literal || !atom_num(e->literal, &i, &f, &isf)) return TD_ERR_NYI;
        if (!(in = emit(c))) return TD_ERR_NYI;
        in->kind = TDD_P_CONST;
        in->ci = isf ? (f != 0.0) : (i != 0);
        return TD_OK;
    }
    case OP_EQ: case OP_NE: case OP_LT: case OP_LE: case OP_GT: case OP_GE: {
        td_op_t* a = op->inputs[0];
        td_op_t* b = op->inputs[1];
        int cmp = op->opcode;
        if (a->opcode == OP_CONST && b->opcode == OP_SCAN) {
            td_op_t* t = a; a = b; b = t;
            cmp = mirror_cmp(cmp);
        }
        if (a->opcode != OP_SCAN) return TD_ERR_NYI;   /* general expressions: not yet */
        td_t* ca = scan_col(g, a, c->tbl);
        if (!ca) return TD_ERR_SCHEMA;
        int ia = pred_col(c, ca);
        if (ia < 0) return -ia;
        if (b->opcode == OP_SCAN) {
            td_t* cb = scan_col(g, b, c->tbl);
            if (!cb) return TD_ERR_SCHEMA;
            int ib = pred_col(c, cb);
            if (ib < 0) return -ib;
            if (!(in = emit(c))) return TD_ERR_NYI;
            in->kind = TDD_P_CMP_COL;
            in->cmp = (uint8_t)cmp;
            in->a = (uint8_t)ia;
            in->b = (uint8_t)ib;
            in->as_f64 = ca->type == TD_F64 || cb->type == TD_F64;
            return TD_OK;
        }
        if (b->opcode != OP_CONST) return TD_ERR_NYI;
        td_op_ext_t* e = td__find_ext(g, b->id);
        if (!e || !e->literal) return TD_ERR_NYI;
        if (!(in = emit(c))) return TD_ERR_NYI;
        int64_t i; double f; int isf;
        if (!atom_num(e->literal, &i, &f, &isf)) {
            if (e->literal->type != TD_ATOM_STR) return TD_ERR_NYI;
            /* A string constant compares as its symbol id (exec.c:826-839);
             * an unknown string matches nothing. */
            int64_t sid = td_sym_find(td_str_ptr(e->literal), td_str_len(e->literal));
            if (sid < 0) {
                in->kind = TDD_P_CONST;
                in->ci = cmp == OP_NE;
                return TD_OK;
            }
            i = sid; f = (double)sid; isf = 0;
        }
        in->kind = TDD_P_CMP_CONST;
        in->cmp = (uint8_t)cmp;
        in->a = (uint8_t)ia;
        in->as_f64 = ca->type == TD_F64 || isf;
        in->ci = i;
        in->cf = isf ? f : (double)i;
        return TD_OK;
    }
    default:
        return TD_ERR_NYI;
    }
}

/* ---------------------------------------------------------- expressions -- */

static int expr_elementwise(uint16_t op) {
    return (op >= OP_NEG && op <= OP_CAST && op != OP_ISNULL) || (op >= OP_ADD && op <= OP_MAX2);
}

#define EXPR_FULL 1000             /* internal: registers or instructions exhausted */

typedef struct {
    xctx_t*     x;
    td_t*       tbl;
    tdd_expr_t* E;
    td_op_t*    pre_op[2];             /* inputs already evaluated into vectors */
    int         int_mode;              /* constant subtrees as i64 (integer-linear sums) */
    td_t*       pre[2];
    uint8_t     reg_of[4096];          /* node id -> register, 0xFF: none yet */
} ecomp_t;

static int ereg_new(ecomp_t* c, uint8_t kind, int8_t type) {
    tdd_expr_t* E = c->E;
    if (E->n_regs >= TDD_EXPR_MAX_REGS) return -1;
    int r = E->n_regs++;
    memset(&E->regs[r], 0, sizeof(E->regs[r]));
    E->regs[r].kind = kind;
    E->regs[r].type = type;
    return r;
}

static int eins(ecomp_t* c, uint8_t op, int dst, int s1, int s2) {
    if (c->E->n_ins >= TDD_EXPR_MAX_INS) return -1;
    c->E->ins[c->E->n_ins++] = (tdd_expr_ins_t){op, (uint8_t)dst, (uint8_t)s1,
                                                (uint8_t)(s2 < 0 ? 0xFF : s2)};
    return 0;
}

/* Register `r` as type t: itself, or a new register holding CAST(r). */
static int ecast(ecomp_t* c, int r, int8_t t) {
    if (c->E->regs[r].type == t) return r;
    int d = ereg_new(c, TDD_R_TMP, t);
    if (d < 0 || eins(c, OP_CAST, d, r, -1)) return -1;
    return d;
}

/* Register of a node (compiled once per node).  Typing as the reference's
 * expr_compile (exec.c:760-918): a column is f64 if its type is, else i64;
 * comparisons and logic give bool; arithmetic gives f64 when an operand is
 * f64 or the op is DIV / SQRT / LOG / EXP, else i64; f64 results promote
 * their operands, and mixed comparisons compare in the wider type. */
static int const_numeric(td_graph_t* g, td_op_t* op, double* f, int64_t* i, int* isf);

/* An integral value of a constant subtree (const_expr_to_i64,
 * exec.c:461-476). */
static int const_i64(td_graph_t* g, td_op_t* op, int64_t* out) {
    double f, ip;
    int64_t i;
    int isf;
    if (!const_numeric(g, op, &f, &i, &isf)) return 0;
    if (!isf) { *out = i; return 1; }
    if (!isfinite(f) || modf(f, &ip) != 0.0 || ip > (double)INT64_MAX || ip < (double)INT64_MIN)
        return 0;
    *out = (int64_t)ip;
    return 1;
}

static int ecompile(ecomp_t* c, td_op_t* op) {
    td_graph_t* g = c->x->g;
    if (op->id < 4096 && c->reg_of[op->id] != 0xFF) return c->reg_of[op->id];
    int r;
    int64_t ci;
    if (c->int_mode && op->opcode != OP_SCAN && const_i64(g, op, &ci)) {
        r = ereg_new(c, TDD_R_CONST, TD_I64);
        if (r < 0) return -EXPR_FULL;
        c->E->regs[r].ci = ci;
        c->E->regs[r].cf = (double)ci;
        if (op->id < 4096) c->reg_of[op->id] = (uint8_t)r;
        return r;
    }
    int pi = op == c->pre_op[0] ? 0 : op == c->pre_op[1] ? 1 : -1;
    if (pi >= 0) {
        td_t* v = c->pre[pi];
        r = ereg_new(c, TDD_R_COL, v->type);
        if (r < 0) return -EXPR_FULL;
        int rc = dev_col(c->x, v, &c->E->regs[r].col);
        if (rc) return -rc;
    } else if (op->opcode == OP_SCAN) {
        td_t* col = scan_col(g, op, c->tbl);
        if (!col) return -TD_ERR_SCHEMA;
        if (col->type == TD_MAPCOMMON || TD_IS_PARTED(col->type) || col->type <= 0 ||
            col->type == TD_TABLE || col->type == TD_LIST)
            return -TD_ERR_NYI;
        if (col->len != td_table_nrows(c->tbl)) return -TD_ERR_LENGTH;
        r = ereg_new(c, TDD_R_COL, col->type == TD_F64 ? TD_F64 : TD_I64);
        if (r < 0) return -EXPR_FULL;
        int rc = dev_col(c->x, col, &c->E->regs[r].col);
        if (rc) return -rc;
    } else if (op->opcode == OP_CONST) {
        td_op_ext_t* e = td__find_ext(g, op->id);
        int64_t i; double f; int isf;
        if (!e || !e->literal) return -TD_ERR_NYI;
        if (!atom_num(e->literal, &i, &f, &isf)) {
            if (e->literal->type != TD_ATOM_STR) return -TD_ERR_NYI;
            i = td_sym_find(td_str_ptr(e->literal), td_str_len(e->literal));
            if (i < 0) return -TD_ERR_NYI;
            f = (double)i;
            isf = 0;
        }
        r = ereg_new(c, TDD_R_CONST, isf ? TD_F64 : TD_I64);
        if (r < 0) return -EXPR_FULL;
        c->E->regs[r].ci = i;
        c->E->regs[r].cf = isf ? f : (double)i;
    } else if (expr_elementwise(op->opcode)) {
        if (!op->inputs[0]) return -TD_ERR_NYI;
        int s1 = ecompile(c, op->inputs[0]);
        if (s1 < 0) return s1;
        int s2 = -1;
        if (op->arity >= 2 && op->inputs[1]) {
            s2 = ecompile(c, op->inputs[1]);
            if (s2 < 0) return s2;
        }
        int8_t t1 = c->E->regs[s1].type, t2 = s2 >= 0 ? c->E->regs[s2].type : t1;
        uint16_t o = op->opcode;
        int8_t ot;
        if (o == OP_CAST) {
            ot = op->out_type;
            if (ot != TD_F64 && ot != TD_I64) return -TD_ERR_NYI;
        } else if ((o >= OP_EQ && o <= OP_GE) || o == OP_AND || o == OP_OR || o == OP_NOT) {
            ot = TD_BOOL;
        } else if (t1 == TD_F64 || t2 == TD_F64 || o == OP_DIV || o == OP_SQRT || o == OP_LOG ||
                   o == OP_EXP) {
            ot = TD_F64;
        } else {
            ot = TD_I64;
        }
        if (o != OP_CAST && ot == TD_F64) {
            s1 = ecast(c, s1, TD_F64);
            if (s2 >= 0) s2 = ecast(c, s2, TD_F64);
        } else if (ot == TD_BOOL && s2 >= 0 && t1 != t2) {
            int8_t pt = t1 == TD_F64 || t2 == TD_F64 ? TD_F64 : TD_I64;
            s1 = ecast(c, s1, pt);
            s2 = ecast(c, s2, pt);
        }
        if (s1 < 0 || (op->arity >= 2 && op->inputs[1] && s2 < 0)) return -EXPR_FULL;
        r = ereg_new(c, TDD_R_TMP, ot);
        if (r < 0 || eins(c, (uint8_t)o, r, s1, s2)) return -EXPR_FULL;
    } else {
        return -TD_ERR_NYI;
    }
    if (op->id < 4096) c->reg_of[op->id] = (uint8_t)r;
    return r;
}

/* An element-wise expression over the rows of `tbl`, as a device vector
 * (TD_I64, TD_F64 or TD_BOOL). */
static td_t* eval_expr_mode(xctx_t* x, td_t* tbl, td_op_t* op, int int_mode);

static td_t* eval_expr(xctx_t* x, td_t* tbl, td_op_t* op) {
    return eval_expr_mode(x, tbl, op, 0);
}

static td_t* eval_expr_mode(xctx_t* x, td_t* tbl, td_op_t* op, int int_mode) {
    if (!tbl || tbl->type != TD_TABLE) return TD_ERR_PTR(TD_ERR_SCHEMA);
    tdd_expr_t E;
    ecomp_t* c = calloc(1, sizeof(ecomp_t));
    if (!c) return TD_ERR_PTR(TD_ERR_OOM);
    c->x = x;
    c->tbl = tbl;
    c->E = &E;
    c->int_mode = int_mode;
    int r = -EXPR_FULL;
    for (int split = 0; split < 2 && r == -EXPR_FULL; split++) {
        if (split) {
            /* Too large for one kernel: the element-wise inputs become
             * vectors of their own (each split again as needed), and the
             * top operation reads them as columns. */
            for (int i = 0; i < 2 && i < op->arity; i++) {
                td_op_t* in = op->inputs[i];
                if (!in || !expr_elementwise(in->opcode)) continue;
                td_t* v = eval_expr_mode(x, tbl, in, int_mode);
                if (TD_IS_ERR(v)) { r = -(int)TD_ERR_CODE(v); break; }
                c->pre_op[i] = in;
                c->pre[i] = v;
            }
            if (r != -EXPR_FULL) break;
        }
        memset(&E, 0, sizeof(E));
        memset(c->reg_of, 0xFF, sizeof(c->reg_of));
        r = ecompile(c, op);
    }
    td_t* pre[2] = {c->pre[0], c->pre[1]};
    free(c);
    if (r < 0) {
        for (int i = 0; i < 2; i++) if (pre[i]) td_release(pre[i]);
        return TD_ERR_PTR(r == -EXPR_FULL ? TD_ERR_NYI : -r);
    }
    if (E.n_ins == 0) {                                        /* a bare SCAN / CONST */
        for (int i = 0; i < 2; i++) if (pre[i]) td_release(pre[i]);
        return TD_ERR_PTR(TD_ERR_NYI);
    }
    E.out_reg = r;
    int8_t ot = E.regs[r].type;
    int64_t n = td_table_nrows(tbl);
    size_t esz = ot == TD_BOOL ? 1 : 8;
    td_t* v = td__hbm_vec_new(ot, 0, n, (size_t)(n > 0 ? n : 1) * esz);
    if (TD_IS_ERR(v)) {
        for (int i = 0; i < 2; i++) if (pre[i]) td_release(pre[i]);
        return v;
    }
    int rc = tdd_expr_eval(&E, n, td_hbm_ref(v)->dptr);
    /* the kernel is queued on the library's stream: releasing the inputs
     * now returns their blocks only to work ordered after it */
    for (int i = 0; i < 2; i++) if (pre[i]) td_release(pre[i]);
    if (rc) { td_release(v); return TD_ERR_PTR(rc); }
    return v;
}

static int compile_pred(xctx_t* x, td_t* tbl, td_op_t* op, tdd_pred_t* P) {
    memset(P, 0, sizeof(*P));
    pcomp_t c;
    memset(&c, 0, sizeof(c));
    c.x = x;
    c.tbl = tbl;
    c.P = P;
    return pred_compile(&c, op);
}

/* Evaluate P over `nrows` into a new device selection, ANDed with `and_with`. */
static td_t* eval_sel(xctx_t* x, const tdd_pred_t* P, int64_t nrows, td_t* and_with) {
    td_t* s = td__sel_new_dev(nrows);
    if (TD_IS_ERR(s)) return s;
    tdd_sel_t v = td__sel_view(td_hbm_ref(s)->dptr, nrows), w;
    int rc = TD_OK;
    if (and_with) rc = dev_sel(x, and_with, &w);
    if (!rc) rc = tdd_sel_eval(P, and_with ? &w : NULL, &v);
    if (rc) { td_release(s); return TD_ERR_PTR(rc); }
    return s;
}

/* A predicate with arithmetic: its BOOL vector as a device selection,
 * ANDed with `and_with`. */
static td_t* pred_expr_sel(xctx_t* x, td_t* tbl, td_op_t* pred, td_t* and_with) {
    td_t* b = eval_expr(x, tbl, pred);
    if (TD_IS_ERR(b)) return b;
    if (b->type != TD_BOOL) { td_release(b); return TD_ERR_PTR(TD_ERR_TYPE); }
    int64_t n = b->len;
    td_t* s = td__sel_new_dev(n);
    if (TD_IS_ERR(s)) { td_release(b); return s; }
    tdd_sel_t v = td__sel_view(td_hbm_ref(s)->dptr, n);
    int rc = tdd_sel_from_bool((const uint8_t*)td_hbm_ref(b)->dptr, &v);
    td_release(b);
    if (!rc && and_with) {
        tdd_sel_t w;
        rc = dev_sel(x, and_with, &w);
        if (!rc) rc = tdd_sel_and(&v, &w, &v);
    }
    if (rc) { td_release(s); return TD_ERR_PTR(rc); }
    return s;
}

/* ----------------------------------------------------------- compaction -- */

/* Table rows selected by s, as a new table of device-resident columns. */
static td_t* compact_table(xctx_t* x, td_t* tbl, td_t* s) {
    int64_t nrows = td_table_nrows(tbl), ncols = td_table_ncols(tbl);
    if (s->len != nrows) return TD_ERR_PTR(TD_ERR_LENGTH);
    tdd_sel_t v;
    int rc = dev_sel(x, s, &v);
    if (rc) return TD_ERR_PTR(rc);
    int64_t pass = 0;
    rc = tdd_sel_total(&v, &pass);
    if (rc) return TD_ERR_PTR(rc);
    td_t* out = td_table_new(ncols);
    if (TD_IS_ERR(out)) return out;
    for (int64_t c0 = 0; c0 < ncols; c0 += 16) {
        int n = (int)(ncols - c0 < 16 ? ncols - c0 : 16);
        tdd_col_t src[16];
        void* dst[16];
        td_t* nc[16];
        for (int i = 0; i < n; i++) {
            td_t* col = td_table_get_col_idx(tbl, c0 + i);
            rc = dev_col(x, col, &src[i]);
            if (rc) goto fail;
            nc[i] = td__hbm_vec_new(col->type, col->attrs & TD_SYM_W_MASK, pass,
                                    (size_t)pass * src[i].esz);
            if (TD_IS_ERR(nc[i])) { rc = TD_ERR_OOM; goto fail; }
            dst[i] = td_hbm_ref(nc[i])->dptr;
            out = td_table_add_col(out, td_table_col_name(tbl, c0 + i), nc[i]);
            td_release(nc[i]);
            if (TD_IS_ERR(out)) return out;
        }
        rc = tdd_sel_compact(&v, n, src, dst);
        if (rc) goto fail;
    }
    return out;
fail:
    td_release(out);
    return TD_ERR_PTR(rc);
}

static td_t* compact_vec(xctx_t* x, td_t* vec, td_t* s) {
    if (s->len != vec->len) return TD_ERR_PTR(TD_ERR_LENGTH);
    tdd_sel_t v;
    tdd_col_t src;
    int rc = dev_sel(x, s, &v);
    if (!rc) rc = dev_col(x, vec, &src);
    int64_t pass = 0;
    if (!rc) rc = tdd_sel_total(&v, &pass);
    if (rc) return TD_ERR_PTR(rc);
    td_t* out = td__hbm_vec_new(vec->type, vec->attrs & TD_SYM_W_MASK, pass, (size_t)pass * src.esz);
    if (TD_IS_ERR(out)) return out;
    void* dst = td_hbm_ref(out)->dptr;
    rc = tdd_sel_compact(&v, 1, &src, &dst);
    if (rc) { td_release(out); return TD_ERR_PTR(rc); }
    return out;
}

/* --------------------------------------------------------------- group -- */

static const char* agg_suffix(uint16_t op) {
    switch (op) {
    case OP_SUM: return "_sum";
    case OP_COUNT: return "_count";
    case OP_AVG: return "_mean";
    case OP_MIN: return "_min";
    case OP_MAX: return "_max";
    case OP_FIRST: return "_first";
    case OP_LAST: return "_last";
    case OP_STDDEV: return "_stddev";
    case OP_STDDEV_POP: return "_stddev_pop";
    case OP_VAR: return "_var";
    case OP_VAR_POP: return "_var_pop";
    default: return "";
    }
}

/* Output column name of aggregate a (emit_agg_columns naming,
 * exec.c:4724-4771): "<input column><suffix>", or "_e<a><suffix>" for a
 * non-SCAN input. */
static int64_t agg_name(td_graph_t* g, td_op_t* in, int a, uint16_t op) {
    char buf[320];
    td_op_t* src = in;
    if (src->opcode == OP_FILTER) src = src->inputs[0];     /* README shape */
    td_op_ext_t* e = td__find_ext(g, src->id);
    if (e && e->base.opcode == OP_SCAN) {
        size_t n;
        const char* s = td__sym_cstr(e->sym, &n);
        const char* sfx = agg_suffix(op);
        if (s && n + strlen(sfx) < sizeof(buf)) {
            memcpy(buf, s, n);
            memcpy(buf + n, sfx, strlen(sfx));
            return td_sym_intern(buf, n + strlen(sfx));
        }
        return e->sym;
    }
    int len = snprintf(buf, sizeof(buf), "_e%d%s", a, agg_suffix(op));
    return td_sym_intern(buf, (size_t)len);
}

static int64_t key_name(td_graph_t* g, td_op_t* k, int idx) {
    td_op_t* src = k->opcode == OP_FILTER ? k->inputs[0] : k;
    td_op_ext_t* e = td__find_ext(g, src->id);
    return e && e->base.opcode == OP_SCAN ? e->sym : idx;
}

/* Wrap a device group result as a table of resident columns. */
static td_t* group_table(td_graph_t* g, td_op_ext_t* ext, const tdd_group_result_t* r,
                         td_t* const* key_src, td_t* const* agg_src) {
    int nk = ext->n_keys, na = ext->n_aggs;
    td_t* out = td_table_new(nk + na);
    if (TD_IS_ERR(out)) goto oom;
    for (int k = 0; k < nk; k++) {
        td_t* s = key_src[k];
        td_t* col = td__hbm_wrap(s->type, s->attrs & TD_SYM_W_MASK, r->n_groups,
                                 r->keys[k], (size_t)r->n_groups * td__esz(s));
        if (TD_IS_ERR(col)) { td_release(out); goto oom_from_key; }
        out = td_table_add_col(out, key_name(g, ext->keys[k], k), col);
        td_release(col);
        if (TD_IS_ERR(out)) return out;
        ((tdd_group_result_t*)r)->keys[k] = NULL;
    }
    for (int a = 0; a < na; a++) {
        int8_t t = r->out_types[a];
        uint8_t attrs = 0, esz = 8;
        if (t != TD_F64 && t != TD_I64 && agg_src[a]) {
            attrs = agg_src[a]->attrs & TD_SYM_W_MASK;
            esz = td__esz(agg_src[a]);
        }
        td_t* col = td__hbm_wrap(t, attrs, r->n_groups, r->vals[a], (size_t)r->n_groups * esz);
        if (TD_IS_ERR(col)) { td_release(out); goto oom_from_agg; }
        out = td_table_add_col(out, agg_name(g, ext->agg_ins[a], a, ext->agg_ops[a]), col);
        td_release(col);
        if (TD_IS_ERR(out)) return out;
        ((tdd_group_result_t*)r)->vals[a] = NULL;
    }
    return out;
oom:
oom_from_key:
oom_from_agg:
    for (int k = 0; k < nk; k++) td__arena_free(r->keys[k]);
    for (int a = 0; a < na; a++) td__arena_free(r->vals[a]);
    return TD_ERR_PTR(TD_ERR_OOM);
}

/* All FILTER(SCAN, P) nodes of a group sharing one P: that P, else NULL. */
static td_op_t* readme_pred(td_op_ext_t* ext) {
    td_op_t* p = NULL;
    int n = ext->n_keys + ext->n_aggs;
    for (int i = 0; i < n; i++) {
        td_op_t* in = i < ext->n_keys ? ext->keys[i] : ext->agg_ins[i - ext->n_keys];
        if (in->opcode != OP_FILTER) return NULL;
        if (in->inputs[0]->opcode != OP_SCAN) return NULL;
        if (p && in->inputs[1] != p) return NULL;
        p = in->inputs[1];
    }
    return p;
}

static td_op_t* strip_filter(td_op_t* op) {
    return op->opcode == OP_FILTER ? op->inputs[0] : op;
}

static td_t* exec_join_group(xctx_t* x, td_op_t* op, td_op_ext_t* ext, td_op_t* join);
static td_t* group_over(xctx_t* x, td_op_ext_t* ext, td_t* tbl, int use_sel);

/* A SCAN at or under `op` (through element-wise nodes) names a column
 * `tbl` lacks. */
static int scan_missing(td_graph_t* g, td_op_t* op, td_t* tbl) {
    if (!op) return 0;
    if (op->opcode == OP_SCAN) return !scan_col(g, op, tbl);
    if (!expr_elementwise(op->opcode)) return 0;
    for (int i = 0; i < op->arity && i < 2; i++)
        if (scan_missing(g, op->inputs[i], tbl)) return 1;
    return 0;
}

static td_t* exec_group(xctx_t* x, td_op_t* op) {
    td_graph_t* g = x->g;
    td_op_ext_t* ext = td__find_ext(g, op->id);
    if (!ext) return TD_ERR_PTR(TD_ERR_NYI);
    if (ext->n_keys < 0 || ext->n_keys > TDD_MAX_KEYS || ext->n_aggs > TDD_MAX_AGGS)
        return TD_ERR_PTR(TD_ERR_NYI);
    td_t* tbl = g->table;
    if (!tbl) return TD_ERR_PTR(TD_ERR_SCHEMA);

    /* GROUP over the graph's JOIN output? */
    int missing = 0;
    for (int i = 0; i < ext->n_keys + ext->n_aggs; i++) {
        td_op_t* in = strip_filter(i < ext->n_keys ? ext->keys[i] : ext->agg_ins[i - ext->n_keys]);
        missing |= scan_missing(g, in, tbl);
    }
    if (missing) {
        if (ext->n_keys == 0) return TD_ERR_PTR(TD_ERR_NYI);
        for (uint32_t i = 0; i < g->node_count; i++)
            if (g->nodes[i].opcode == OP_JOIN) return exec_join_group(x, op, ext, &g->nodes[i]);
        return TD_ERR_PTR(TD_ERR_SCHEMA);
    }
    return group_over(x, ext, tbl, 1);
}

/* ---------------------------------------------- affine aggregate inputs -- */

/* The value of a constant subtree as the reference's aggregate rewrites
 * evaluate it (eval_const_numeric_expr, exec.c:380-458): constants, NEG /
 * ABS, and binary arithmetic over constants, in f64 when an operand is f64,
 * the node is typed f64 or the op is DIV (division or modulo by zero give
 * 0), otherwise in wrapping i64. */
static int const_numeric(td_graph_t* g, td_op_t* op, double* f, int64_t* i, int* isf) {
    if (!op) return 0;
    if (op->opcode == OP_CONST) {
        td_op_ext_t* e = td__find_ext(g, op->id);
        if (!e || !e->literal || !td_is_atom(e->literal)) return 0;
        return atom_num(e->literal, i, f, isf);
    }
    if ((op->opcode == OP_NEG || op->opcode == OP_ABS) && op->arity == 1) {
        double af;
        int64_t ai;
        int af64;
        if (!const_numeric(g, op->inputs[0], &af, &ai, &af64)) return 0;
        if (af64 || op->out_type == TD_F64) {
            double v = af64 ? af : (double)ai;
            *f = op->opcode == OP_NEG ? -v : fabs(v);
            *i = (int64_t)*f;
            *isf = 1;
            return 1;
        }
        uint64_t u = (uint64_t)ai;
        *i = (int64_t)(op->opcode == OP_NEG || ai < 0 ? 0 - u : u);
        if (op->opcode == OP_ABS && ai >= 0) *i = ai;
        *f = (double)*i;
        *isf = 0;
        return 1;
    }
    if (op->arity != 2 || op->opcode < OP_ADD || op->opcode > OP_MAX2) return 0;
    double lf, rf;
    int64_t li, ri;
    int lf64, rf64;
    if (!const_numeric(g, op->inputs[0], &lf, &li, &lf64) ||
        !const_numeric(g, op->inputs[1], &rf, &ri, &rf64))
        return 0;
    if (op->out_type == TD_F64 || lf64 || rf64 || op->opcode == OP_DIV) {
        double a = lf64 ? lf : (double)li, b = rf64 ? rf : (double)ri, r;
        switch (op->opcode) {
        case OP_ADD: r = a + b; break;
        case OP_SUB: r = a - b; break;
        case OP_MUL: r = a * b; break;
        case OP_DIV: r = b != 0.0 ? a / b : 0.0; break;
        case OP_MOD: r = b != 0.0 ? fmod(a, b) : 0.0; break;
        case OP_MIN2: r = a < b ? a : b; break;
        case OP_MAX2: r = a > b ? a : b; break;
        default: return 0;
        }
        *f = r;
        *i = (int64_t)r;
        *isf = 1;
        return 1;
    }
    int ok = ri != 0 && !(li == INT64_MIN && ri == -1);
    int64_t r;
    switch (op->opcode) {
    case OP_ADD: r = (int64_t)((uint64_t)li + (uint64_t)ri); break;
    case OP_SUB: r = (int64_t)((uint64_t)li - (uint64_t)ri); break;
    case OP_MUL: r = (int64_t)((uint64_t)li * (uint64_t)ri); break;
    case OP_DIV: r = ok ? li / ri : 0; break;
    case OP_MOD: r = ok ? li % ri : 0; break;
    case OP_MIN2: r = li < ri ? li : ri; break;
    case OP_MAX2: r = li > ri ? li : ri; break;
    default: return 0;
    }
    *i = r;
    *f = (double)r;
    *isf = 0;
    return 1;
}

typedef struct {
    int     on;
    int64_t bias_i;
    double  bias_f;
} affine_t;

/* SUM / AVG(col + c), (c + col) or (col - c) over a column of `tbl`: the
 * column and the bias (try_affine_sumavg_input, exec.c:614-694).  An
 * integer column takes only an integral bias. */
static td_t* affine_input(td_graph_t* g, td_t* tbl, td_op_t* op, affine_t* A) {
    if ((op->opcode != OP_ADD && op->opcode != OP_SUB) || op->arity != 2) return NULL;
    double lf, rf, cf;
    int64_t li, ri, ci;
    int lf64, rf64, cf64;
    int lc = const_numeric(g, op->inputs[0], &lf, &li, &lf64);
    int rc = const_numeric(g, op->inputs[1], &rf, &ri, &rf64);
    td_op_t* base;
    int sign = 1;
    if (op->opcode == OP_ADD && lc) {
        base = op->inputs[1]; cf = lf; ci = li; cf64 = lf64;
    } else if (op->opcode == OP_ADD && rc) {
        base = op->inputs[0]; cf = rf; ci = ri; cf64 = rf64;
    } else if (op->opcode == OP_SUB && rc) {
        base = op->inputs[0]; cf = rf; ci = ri; cf64 = rf64; sign = -1;
    } else {
        return NULL;
    }
    if (base->opcode != OP_SCAN) return NULL;
    td_t* col = scan_col(g, base, tbl);
    if (!col) return NULL;
    int8_t t = col->type;
    if (t == TD_F64) {
        A->bias_f = (double)sign * (cf64 ? cf : (double)ci);
        A->bias_i = (int64_t)A->bias_f;
    } else if (t == TD_I64 || t == TD_TIMESTAMP || t == TD_I32 || t == TD_I16 || t == TD_U8 ||
               t == TD_BOOL || t == TD_SYM) {
        int64_t c = ci;
        if (cf64) {
            double ip;
            if (!isfinite(cf) || modf(cf, &ip) != 0.0 || ip > (double)INT64_MAX ||
                ip < (double)INT64_MIN)
                return NULL;
            c = (int64_t)ip;
        }
        A->bias_i = sign > 0 ? c : (int64_t)(0 - (uint64_t)c);
        A->bias_f = (double)A->bias_i;
    } else {
        return NULL;
    }
    A->on = 1;
    return col;
}

/* Integer-linear form of an aggregate input, sum of coefficient x column
 * plus a bias over integer-class columns (parse_linear_i64_expr,
 * exec.c:528-596): terms merged by column, zero terms dropped, at most 8. */
typedef struct {
    int     n;
    int64_t sym[8], coeff[8];
} lin_t;

static int lin_add(lin_t* L, int64_t sym, int64_t k) {
    if (k == 0) return 1;
    for (int i = 0; i < L->n; i++) {
        if (L->sym[i] != sym) continue;
        int64_t nx = (int64_t)((uint64_t)L->coeff[i] + (uint64_t)k);
        if (nx) { L->coeff[i] = nx; return 1; }
        for (int j = i + 1; j < L->n; j++) { L->sym[j - 1] = L->sym[j]; L->coeff[j - 1] = L->coeff[j]; }
        L->n--;
        return 1;
    }
    if (L->n >= 8) return 0;
    L->sym[L->n] = sym;
    L->coeff[L->n++] = k;
    return 1;
}

static int lin_parse(td_graph_t* g, td_op_t* op, int64_t scale, lin_t* L) {
    int64_t c;
    if (!op) return 0;
    if (const_i64(g, op, &c)) return 1;                   /* the bias: no terms */
    if (op->opcode == OP_SCAN) {
        td_op_ext_t* e = td__find_ext(g, op->id);
        return e && e->base.opcode == OP_SCAN && lin_add(L, e->sym, scale);
    }
    if (op->opcode == OP_NEG && op->arity == 1)
        return lin_parse(g, op->inputs[0], (int64_t)(0 - (uint64_t)scale), L);
    if ((op->opcode == OP_ADD || op->opcode == OP_SUB) && op->arity == 2)
        return lin_parse(g, op->inputs[0], scale, L) &&
               lin_parse(g, op->inputs[1], op->opcode == OP_ADD ? scale
                                                                 : (int64_t)(0 - (uint64_t)scale), L);
    if (op->opcode == OP_MUL && op->arity == 2) {
        if (const_i64(g, op->inputs[0], &c))
            return lin_parse(g, op->inputs[1], (int64_t)((uint64_t)scale * (uint64_t)c), L);
        if (const_i64(g, op->inputs[1], &c))
            return lin_parse(g, op->inputs[0], (int64_t)((uint64_t)scale * (uint64_t)c), L);
    }
    return 0;
}

/* SUM / AVG of an integer-linear input with no group keys: the reference
 * sums it in wrapping i64 (try_linear_sumavg_input_i64, exec.c:598-612),
 * integral f64 constants included. */
static int linear_i64(td_graph_t* g, td_t* tbl, td_op_t* op) {
    lin_t L;
    L.n = 0;
    if (op->opcode == OP_SCAN || !lin_parse(g, op, 1, &L)) return 0;
    for (int i = 0; i < L.n; i++) {
        td_t* col = td_table_get_col(tbl, L.sym[i]);
        if (!col) return 0;
        int8_t t = col->type;
        if (!(t == TD_I64 || t == TD_TIMESTAMP || t == TD_I32 || t == TD_DATE || t == TD_TIME ||
              t == TD_I16 || t == TD_U8 || t == TD_BOOL || t == TD_SYM))
            return 0;
    }
    return 1;
}

static td_t* group_over_ev(xctx_t* x, td_op_ext_t* ext, td_t* tbl, int use_sel, td_t** ev);

/* GROUP of `tbl`'s columns; with use_sel, the graph's lazy selection (or the
 * README shape's shared predicate) masks the rows and is consumed.  Key and
 * aggregate inputs that are expressions are evaluated over the rows first
 * (into `ev`, released once the group result holds its own columns). */
static td_t* group_over(xctx_t* x, td_op_ext_t* ext, td_t* tbl, int use_sel) {
    td_t* ev[TDD_MAX_KEYS + TDD_MAX_AGGS] = {0};
    td_t* r = group_over_ev(x, ext, tbl, use_sel, ev);
    for (int i = 0; i < TDD_MAX_KEYS + TDD_MAX_AGGS; i++)
        if (ev[i]) td_release(ev[i]);
    return r;
}

static td_t* group_over_ev(xctx_t* x, td_op_ext_t* ext, td_t* tbl, int use_sel, td_t** ev) {
    td_graph_t* g = x->g;
    int64_t nrows = td_table_nrows(tbl);
    tdd_group_spec_t S;
    memset(&S, 0, sizeof(S));
    S.nrows = nrows;
    S.n_keys = ext->n_keys;
    S.n_aggs = ext->n_aggs;
    S.force_path = td__force_group_path();
    td_t* key_src[TDD_MAX_KEYS];
    td_t* agg_src[TDD_MAX_AGGS];
    affine_t aff[TDD_MAX_AGGS];
    int need_count = 0;
    int rc;
    for (int k = 0; k < ext->n_keys; k++) {
        td_op_t* kop = strip_filter(ext->keys[k]);
        if (expr_elementwise(kop->opcode)) {
            ev[k] = eval_expr(x, tbl, kop);
            if (TD_IS_ERR(ev[k])) { td_t* e = ev[k]; ev[k] = NULL; return e; }
            key_src[k] = ev[k];
        } else if (kop->opcode == OP_SCAN) {
            key_src[k] = scan_col(g, kop, tbl);
        } else {
            return TD_ERR_PTR(TD_ERR_NYI);
        }
        if (!key_src[k]) return TD_ERR_PTR(TD_ERR_SCHEMA);
        if (key_src[k]->len != nrows) return TD_ERR_PTR(TD_ERR_LENGTH);
        if ((rc = dev_col(x, key_src[k], &S.keys[k]))) return TD_ERR_PTR(rc);
    }
    for (int a = 0; a < ext->n_aggs; a++) {
        td_op_t* aop = strip_filter(ext->agg_ins[a]);
        S.ops[a] = ext->agg_ops[a];
        agg_src[a] = NULL;
        aff[a].on = 0;
        td_t* base = NULL;
        if (aop->opcode == OP_SCAN) {
            agg_src[a] = scan_col(g, aop, tbl);
            if (!agg_src[a]) return TD_ERR_PTR(TD_ERR_SCHEMA);
            if (agg_src[a]->len != nrows) return TD_ERR_PTR(TD_ERR_LENGTH);
            if ((rc = dev_col(x, agg_src[a], &S.aggs[a]))) return TD_ERR_PTR(rc);
        } else if ((S.ops[a] == OP_SUM || S.ops[a] == OP_AVG) &&
                   (base = affine_input(g, tbl, aop, &aff[a]))) {
            /* aggregate the column, add the bias to the results */
            if (base->len != nrows) return TD_ERR_PTR(TD_ERR_LENGTH);
            agg_src[a] = base;
            if ((rc = dev_col(x, base, &S.aggs[a]))) return TD_ERR_PTR(rc);
            if (S.ops[a] == OP_SUM) need_count = 1;
        } else if (expr_elementwise(aop->opcode)) {
            /* no keys and an integer-linear SUM / AVG input: i64 arithmetic */
            int lin = ext->n_keys == 0 && (S.ops[a] == OP_SUM || S.ops[a] == OP_AVG) &&
                      linear_i64(g, tbl, aop);
            td_t* e = eval_expr_mode(x, tbl, aop, lin);
            if (TD_IS_ERR(e)) return e;
            ev[TDD_MAX_KEYS + a] = agg_src[a] = e;
            if ((rc = dev_col(x, e, &S.aggs[a]))) return TD_ERR_PTR(rc);
        } else if (!(aop->opcode == OP_CONST && S.ops[a] == OP_COUNT)) {
            return TD_ERR_PTR(TD_ERR_NYI);
        }
    }

    /* Mask: the README shape's shared predicate, else the lazy selection. */
    td_t* sel = NULL;
    td_op_t* rp = readme_pred(ext);
    if (!use_sel) {
        if (rp) return TD_ERR_PTR(TD_ERR_NYI);
    } else if (rp) {
        tdd_pred_t P;
        rc = compile_pred(x, tbl, rp, &P);
        if (rc == TD_ERR_NYI) sel = pred_expr_sel(x, tbl, rp, g->selection);
        else if (rc) return TD_ERR_PTR(rc);
        else sel = eval_sel(x, &P, nrows, g->selection);
        if (TD_IS_ERR(sel)) return sel;
    } else if (g->selection) {
        if (g->selection->len != nrows) return TD_ERR_PTR(TD_ERR_LENGTH);
        sel = g->selection;
        td_retain(sel);
    }
    if (use_sel && g->selection) { td_release(g->selection); g->selection = NULL; }
    tdd_sel_t mv;
    if (sel && (rc = dev_sel(x, sel, &mv))) { td_release(sel); return TD_ERR_PTR(rc); }
    S.mask = sel ? &mv : NULL;
    /* an affine SUM needs the group counts: a hidden COUNT after the others */
    int hidden = -1;
    if (need_count) {
        if (S.n_aggs >= TDD_MAX_AGGS) { td_release(sel); return TD_ERR_PTR(TD_ERR_NYI); }
        hidden = S.n_aggs++;
        S.ops[hidden] = OP_COUNT;
        memset(&S.aggs[hidden], 0, sizeof(S.aggs[hidden]));
    }
    tdd_group_result_t R;
    rc = tdd_group(&S, &R);
    td_release(sel);
    if (rc) return TD_ERR_PTR(rc);
    for (int a = 0; a < ext->n_aggs && !rc; a++)
        if (aff[a].on)
            rc = tdd_affine_fix(R.vals[a], R.out_types[a], S.ops[a] == OP_AVG,
                                hidden >= 0 ? (const int64_t*)R.vals[hidden] : NULL, R.n_groups,
                                aff[a].bias_i, aff[a].bias_f);
    if (hidden >= 0) td__arena_free(R.vals[hidden]);
    if (rc) {
        for (int k = 0; k < ext->n_keys; k++) td__arena_free(R.keys[k]);
        for (int a = 0; a < ext->n_aggs; a++) td__arena_free(R.vals[a]);
        return TD_ERR_PTR(rc);
    }
    td__set_plan("group path=%d groups=%lld tries=%d part_rows=%lld", R.path,
                 (long long)R.n_groups, R.tries, (long long)R.part_rows);
    return group_table(g, ext, &R, key_src, agg_src);
}

/* ---------------------------------------------------------------- join -- */

static int join_is_right_key(td_graph_t* g, td_op_ext_t* je, int64_t name) {
    for (int k = 0; k < je->join.n_join_keys; k++) {
        td_op_ext_t* rk = td__find_ext(g, je->join.right_keys[k]->id);
        if (rk && rk->base.opcode == OP_SCAN && rk->sym == name) return 1;
    }
    return 0;
}

static int join_keys(xctx_t* x, td_op_ext_t* je, td_t* left, td_t* right, tdd_join_spec_t* J) {
    td_graph_t* g = x->g;
    J->n_keys = je->join.n_join_keys;
    if (J->n_keys < 1 || J->n_keys > TDD_MAX_KEYS) return TD_ERR_NYI;
    for (int k = 0; k < J->n_keys; k++) {
        td_t* lc = scan_col(g, je->join.left_keys[k], left);
        td_t* rc_ = scan_col(g, je->join.right_keys[k], right);
        if (!rc_) {
            td_op_ext_t* re = td__find_ext(g, je->join.right_keys[k]->id);
            if (re && re->base.opcode == OP_CONST && re->literal && td_is_vec(re->literal))
                rc_ = re->literal;
        }
        if (!lc || !rc_) return TD_ERR_SCHEMA;
        int rc = dev_col(x, lc, &J->lkeys[k]);
        if (!rc) rc = dev_col(x, rc_, &J->rkeys[k]);
        if (rc) return rc;
    }
    return TD_OK;
}

/* Column statistics of a join output column copied from `src`: the
 * source's bounds hold for its rows (a subset), widened to 0 when rows
 * without a partner are zero-filled. */
static void join_out_stats(td_t* out, const tdd_col_t* src, int zero_filled) {
    if (!src->ranged || TD_IS_ERR(out) || out->mmod != TD_MMOD_HBM) return;
    td_hbm_ref_t* r = td_hbm_ref(out);
    r->lo = zero_filled && src->lo > 0 ? 0 : src->lo;
    r->hi = zero_filled && src->hi < 0 ? 0 : src->hi;
    r->ranged = 1;
}

static td_t* gather_col(xctx_t* x, td_t* col, const int64_t* idx, int64_t n, int nullable) {
    tdd_col_t src;
    int rc = dev_col(x, col, &src);
    if (rc) return TD_ERR_PTR(rc);
    td_t* out = td__hbm_vec_new(col->type, col->attrs & TD_SYM_W_MASK, n, (size_t)n * src.esz);
    if (TD_IS_ERR(out)) return out;
    rc = tdd_gather(&src, idx, n, nullable, td_hbm_ref(out)->dptr);
    if (rc) { td_release(out); return TD_ERR_PTR(rc); }
    join_out_stats(out, &src, nullable);
    return out;
}

/* INNER / LEFT joins through the materialising radix join: the output
 * columns (left columns, then right non-key columns, 16 each at most,
 * exec.c:8607-8728) come back written.  NULL when that join does not take
 * the query (repeated build keys, oversized partitions, more columns than
 * it writes, no memory for its partitions): the caller joins through
 * row-index pairs instead. */
static td_t* join_materialised(xctx_t* x, td_op_ext_t* je, td_t* left, td_t* right,
                               const tdd_join_spec_t* J) {
    td_graph_t* g = x->g;
    tdd_join_mat_spec_t M;
    memset(&M, 0, sizeof(M));
    M.join = *J;
    td_t* lc[16];
    td_t* rc_[16];
    int64_t rname[16], lname[16];
    int64_t lnc = td_table_ncols(left), rnc = td_table_ncols(right);
    for (int64_t c = 0; c < lnc; c++) {
        td_t* col = td_table_get_col_idx(left, c);
        if (!col) continue;
        if (M.n_lcols == 16) return NULL;
        if (dev_col(x, col, &M.lcols[M.n_lcols])) return NULL;
        lname[M.n_lcols] = td_table_col_name(left, c);
        lc[M.n_lcols++] = col;
    }
    for (int64_t c = 0; c < rnc; c++) {
        td_t* col = td_table_get_col_idx(right, c);
        int64_t name = td_table_col_name(right, c);
        if (!col || join_is_right_key(g, je, name)) continue;
        if (M.n_rcols == 16) return NULL;
        if (dev_col(x, col, &M.rcols[M.n_rcols])) return NULL;
        rname[M.n_rcols] = name;
        rc_[M.n_rcols++] = col;
    }
    tdd_join_mat_result_t R;
    int rc = tdd_join_mat(&M, &R);
    /* declined, or out of memory for its partitions: the pair-producing
     * join needs less (the reference's fallback to its chained join) */
    if (rc == TD_ERR_NYI || (rc == TD_ERR_OOM && J->force_path == 0)) return NULL;
    if (rc) return TD_ERR_PTR(rc);
    td__set_plan("join path=%d radix_bits=%d rows=%lld", R.path, R.radix_bits,
                 (long long)R.n_rows);
    td_t* result = td_table_new(M.n_lcols + M.n_rcols);
    for (int c = 0; c < M.n_lcols + M.n_rcols; c++) {
        int right_side = c >= M.n_lcols;
        int i = right_side ? c - M.n_lcols : c;
        td_t* src = right_side ? rc_[i] : lc[i];
        void* d = right_side ? R.rcols[i] : R.lcols[i];
        if (TD_IS_ERR(result)) { td__arena_free(d); continue; }
        const tdd_col_t* dc = right_side ? &M.rcols[i] : &M.lcols[i];
        td_t* o = td__hbm_wrap(src->type, src->attrs & TD_SYM_W_MASK, R.n_rows, d,
                               (size_t)(R.n_rows > 0 ? R.n_rows : 1) * dc->esz);
        if (TD_IS_ERR(o)) { td_release(result); result = o; continue; }
        join_out_stats(o, dc, right_side && J->join_type == 1);   /* LEFT: zero-filled right rows */
        int64_t name = right_side ? rname[i] : lname[i];
        result = td_table_add_col(result, name, o);
        td_release(o);
    }
    return result;
}

static td_t* exec_join(xctx_t* x, td_op_t* op) {
    td_graph_t* g = x->g;
    td_op_ext_t* je = td__find_ext(g, op->id);
    if (!je) return TD_ERR_PTR(TD_ERR_NYI);
    td_t* left = NULL;
    td_t* right = NULL;
    left = xnode(x, op->inputs[0]);
    if (TD_IS_ERR(left)) return left;
    right = xnode(x, op->inputs[1]);
    if (TD_IS_ERR(right)) { td_release(left); return right; }
    td_t* result = NULL;
    if (left->type != TD_TABLE || right->type != TD_TABLE) {
        result = TD_ERR_PTR(TD_ERR_TYPE);
        goto done;
    }
    int64_t nl = td_table_nrows(left), nr = td_table_nrows(right);
    if (nl > INT32_MAX || nr > INT32_MAX) { result = TD_ERR_PTR(TD_ERR_NYI); goto done; }
    tdd_join_spec_t J;
    memset(&J, 0, sizeof(J));
    J.nl = nl;
    J.nr = nr;
    J.join_type = je->join.join_type;
    J.force_path = td__force_join_path();
    J.radix_bits = td__force_radix_bits();
    int rc = join_keys(x, je, left, right, &J);
    if (rc) { result = TD_ERR_PTR(rc); goto done; }
    td_t* sel = g->selection;
    tdd_sel_t sv;
    if (sel) {
        if (sel->len != nl) { result = TD_ERR_PTR(TD_ERR_LENGTH); goto done; }
        if ((rc = dev_sel(x, sel, &sv))) { result = TD_ERR_PTR(rc); goto done; }
        J.lsel = &sv;
    }
    if (J.join_type <= 1) {
        result = join_materialised(x, je, left, right, &J);
        if (result) {
            if (sel) { td_release(g->selection); g->selection = NULL; }
            goto done;
        }
    }
    tdd_join_result_t R;
    rc = tdd_join(&J, &R);
    if (sel) { td_release(g->selection); g->selection = NULL; }
    if (rc) { result = TD_ERR_PTR(rc); goto done; }
    td__set_plan("join path=%d radix_bits=%d pairs=%lld", R.path, R.radix_bits,
                 (long long)R.n_pairs);
    int64_t lnc = td_table_ncols(left), rnc = td_table_ncols(right);
    result = td_table_new(lnc + rnc);
    int ln = 0, rn = 0;
    for (int64_t c = 0; c < lnc && ln < 16 && !TD_IS_ERR(result); c++) {
        td_t* col = td_table_get_col_idx(left, c);
        if (!col) continue;
        td_t* o = gather_col(x, col, R.l_idx, R.n_pairs, J.join_type == 2);
        if (TD_IS_ERR(o)) { td_release(result); result = o; break; }
        result = td_table_add_col(result, td_table_col_name(left, c), o);
        td_release(o);
        ln++;
    }
    for (int64_t c = 0; c < rnc && rn < 16 && !TD_IS_ERR(result); c++) {
        td_t* col = td_table_get_col_idx(right, c);
        int64_t name = td_table_col_name(right, c);
        if (!col || join_is_right_key(g, je, name)) continue;
        td_t* o = gather_col(x, col, R.r_idx, R.n_pairs, J.join_type >= 1);
        if (TD_IS_ERR(o)) { td_release(result); result = o; break; }
        result = td_table_add_col(result, name, o);
        td_release(o);
        rn++;
    }
    td__arena_free(R.l_idx);
    td__arena_free(R.r_idx);
done:
    td_release(left);
    td_release(right);
    return result;
}

/* The CONST table under FILTER*( ) and the FILTER predicates on the way. */
static td_t* const_table_under(td_graph_t* g, td_op_t* op, td_op_t** preds, int* npred) {
    *npred = 0;
    while (op->opcode == OP_FILTER) {
        if (*npred >= 8) return NULL;
        preds[(*npred)++] = op->inputs[1];
        op = op->inputs[0];
    }
    if (op->opcode != OP_CONST) return NULL;
    td_op_ext_t* e = td__find_ext(g, op->id);
    return e && e->literal && e->literal->type == TD_TABLE ? e->literal : NULL;
}

/* Shapes the fused kernel does not take: materialise the join, then group
 * its output table. */
static td_t* join_then_group(xctx_t* x, td_op_ext_t* ext, td_op_t* jop) {
    td_t* jt = xnode(x, jop);
    if (TD_IS_ERR(jt)) return jt;
    td_t* r = group_over(x, ext, jt, 0);
    td_release(jt);
    return r;
}

/* The side (0 left, 1 right) whose columns every SCAN under `op` reads,
 * -1 when they span both, name a right key, or are missing. */
static int expr_side(td_graph_t* g, td_op_ext_t* je, td_op_t* op, td_t* left, td_t* right) {
    if (!op) return -1;
    if (op->opcode == OP_SCAN) {
        td_op_ext_t* se = td__find_ext(g, op->id);
        if (!se) return -1;
        if (td_table_get_col(left, se->sym)) return 0;
        if (!join_is_right_key(g, je, se->sym) && td_table_get_col(right, se->sym)) return 1;
        return -1;
    }
    if (op->opcode == OP_CONST) return 2;               /* either side */
    int s = 2;
    for (int i = 0; i < op->arity && i < 2; i++) {
        int t = expr_side(g, je, op->inputs[i], left, right);
        if (t < 0) return -1;
        if (t == 2) continue;
        if (s != 2 && s != t) return -1;
        s = t;
    }
    return s == 2 ? 0 : s;
}

/* Inputs of the fused join -> group-by: SCANs and expressions over one
 * side, and no input the reference aggregates through a rewrite (affine
 * SUM / AVG, exec.c:614-694), which the join-then-group form applies. */
static int join_group_inputs_ok(td_graph_t* g, td_op_ext_t* ext, td_op_ext_t* je, td_t* left,
                                td_t* right) {
    for (int a = 0; a < ext->n_aggs; a++) {
        td_op_t* in = ext->agg_ins[a];
        if (in->opcode == OP_SCAN || in->opcode == OP_CONST) continue;
        if (!expr_elementwise(in->opcode) || expr_side(g, je, in, left, right) < 0) return 0;
        uint16_t o = ext->agg_ops[a];
        affine_t A;
        td_t* side = expr_side(g, je, in, left, right) ? right : left;
        if ((o == OP_SUM || o == OP_AVG) && affine_input(g, side, in, &A)) return 0;
    }
    for (int k = 0; k < ext->n_keys; k++)
        if (ext->keys[k]->opcode != OP_SCAN) return 0;
    return 1;
}

static td_t* exec_join_group(xctx_t* x, td_op_t* op, td_op_ext_t* ext, td_op_t* jop) {
    td_graph_t* g = x->g;
    td_op_ext_t* je = td__find_ext(g, jop->id);
    if (!je) return TD_ERR_PTR(TD_ERR_NYI);
    td_op_t* preds[8];
    int npred = 0, nrp = 0;
    td_op_t* rpreds[8];
    td_t* left = const_table_under(g, jop->inputs[0], preds, &npred);
    td_t* right = const_table_under(g, jop->inputs[1], rpreds, &nrp);
    (void)op;
    if (je->join.join_type != 0 || !left || !right || nrp ||
        td_table_nrows(left) > INT32_MAX || td_table_nrows(right) > INT32_MAX ||
        !join_group_inputs_ok(g, ext, je, left, right))
        return join_then_group(x, ext, jop);
    int64_t nl = td_table_nrows(left), nr = td_table_nrows(right);
    tdd_join_group_spec_t S;
    memset(&S, 0, sizeof(S));
    S.join.nl = nl;
    S.join.nr = nr;
    S.join.join_type = 0;
    S.join.force_path = td__force_join_path();
    int rc = join_keys(x, je, left, right, &S.join);
    if (rc) return TD_ERR_PTR(rc);
    /* Filters on the left input: predicates over the bound table (whose rows
     * must be the left table's), all ANDed, plus any pending selection. */
    tdd_pred_t P;
    td_t* sel = g->selection;
    if (sel) td_retain(sel);
    if (npred) {
        if (td_table_nrows(g->table) != nl) { td_release(sel); return TD_ERR_PTR(TD_ERR_LENGTH); }
        for (int i = 0; i < npred; i++) {
            rc = compile_pred(x, g->table, preds[i], &P);
            if (rc == TD_ERR_NYI) {
                /* arithmetic in the predicate: an expression selection */
                td_t* s2 = pred_expr_sel(x, g->table, preds[i], sel);
                td_release(sel);
                if (TD_IS_ERR(s2)) return s2;
                sel = s2;
                rc = TD_OK;
                if (i == npred - 1) npred = 0;               /* nothing left to fuse */
                continue;
            }
            if (rc) { td_release(sel); return TD_ERR_PTR(rc); }
            if (npred == 1 && !sel) break;                  /* fused into the probe */
            td_t* s2 = eval_sel(x, &P, nl, sel);
            td_release(sel);
            if (TD_IS_ERR(s2)) return s2;
            sel = s2;
        }
    }
    if (g->selection) { td_release(g->selection); g->selection = NULL; }
    tdd_sel_t sv;
    if (sel) {
        if (sel->len != nl) { td_release(sel); return TD_ERR_PTR(TD_ERR_LENGTH); }
        if ((rc = dev_sel(x, sel, &sv))) { td_release(sel); return TD_ERR_PTR(rc); }
        S.join.lsel = &sv;
    }
    S.lpred = (npred == 1 && !sel) ? &P : NULL;
    /* Group columns: left table first, then right non-key columns.  An
     * aggregate input that is an expression over one side's columns is
     * evaluated over that side first and aggregated as a column of it. */
    S.n_keys = ext->n_keys;
    S.n_aggs = ext->n_aggs;
    td_t* key_src[TDD_MAX_KEYS];
    td_t* agg_src[TDD_MAX_AGGS];
    td_t* ev[TDD_MAX_AGGS] = {0};
    for (int i = 0; i < ext->n_keys + ext->n_aggs && !rc; i++) {
        int is_key = i < ext->n_keys;
        td_op_t* in = is_key ? ext->keys[i] : ext->agg_ins[i - ext->n_keys];
        td_t* col = NULL;
        int side = 0;
        if (in->opcode == OP_SCAN) {
            td_op_ext_t* se = td__find_ext(g, in->id);
            col = td_table_get_col(left, se->sym);
            if (!col && !join_is_right_key(g, je, se->sym)) { col = td_table_get_col(right, se->sym); side = 1; }
            if (!col) rc = TD_ERR_SCHEMA;
        } else if (!is_key && expr_elementwise(in->opcode) &&
                   (side = expr_side(g, je, in, left, right)) >= 0) {
            td_t* e = eval_expr(x, side ? right : left, in);
            if (TD_IS_ERR(e)) rc = TD_ERR_CODE(e);
            else col = ev[i - ext->n_keys] = e;
        } else if (is_key || !(in->opcode == OP_CONST && ext->agg_ops[i - ext->n_keys] == OP_COUNT)) {
            rc = TD_ERR_NYI;
        }
        if (rc) break;
        if (is_key) {
            key_src[i] = col;
            S.key_side[i] = (uint8_t)side;
            rc = dev_col(x, col, &S.keys[i]);
        } else {
            int a = i - ext->n_keys;
            agg_src[a] = col;
            S.ops[a] = ext->agg_ops[a];
            S.agg_side[a] = (uint8_t)side;
            if (col) rc = dev_col(x, col, &S.aggs[a]);
        }
    }
    tdd_group_result_t R;
    if (!rc) rc = tdd_join_group(&S, &R);
    td_release(sel);
    for (int a = 0; a < TDD_MAX_AGGS; a++)
        if (ev[a]) td_release(ev[a]);          /* the launches are ordered before any reuse */
    if (rc) return TD_ERR_PTR(rc);
    td__set_plan("join_group path=%d groups=%lld fused_pred=%d", R.path, (long long)R.n_groups,
                 S.lpred != NULL);
    return group_table(g, ext, &R, key_src, agg_src);
}

/* ----------------------------------------------------------- reductions -- */

/* Distinct values of a vector (exec_count_distinct, exec.c:1658-1727):
 * integer-like types by their widened value, F64 by its bits with every NaN
 * one value and -0.0 equal to 0.0.  Counted as the groups of a one-key
 * group-by on the device. */
static td_t* count_distinct(xctx_t* x, td_t* in) {
    switch (in->type) {
    case TD_BOOL: case TD_U8: case TD_CHAR: case TD_I16: case TD_I32: case TD_I64:
    case TD_F64: case TD_DATE: case TD_TIME: case TD_TIMESTAMP: case TD_SYM:
        break;
    default:
        return TD_ERR_PTR(TD_ERR_TYPE);
    }
    int64_t n = in->len;
    if (n == 0) return td_i64(0);
    tdd_group_spec_t S;
    memset(&S, 0, sizeof(S));
    S.nrows = n;
    S.n_keys = 1;
    S.n_aggs = 1;
    S.ops[0] = OP_COUNT;
    int rc = dev_col(x, in, &S.keys[0]);
    void* canon = NULL;
    if (!rc && in->type == TD_F64) {
        canon = td__arena_alloc((size_t)n * 8);
        rc = canon ? tdd_f64_canon((const double*)S.keys[0].data, n, canon) : TD_ERR_OOM;
        S.keys[0].data = canon;
        S.keys[0].type = TD_I64;
        S.keys[0].ranged = 0;
    }
    tdd_group_result_t R;
    if (!rc) rc = tdd_group(&S, &R);
    if (!rc) {
        td__arena_free(R.keys[0]);
        td__arena_free(R.vals[0]);
    }
    if (canon) td__arena_free(canon);
    return rc ? TD_ERR_PTR(rc) : td_i64(R.n_groups);
}

static td_t* exec_reduce(xctx_t* x, td_op_t* op) {
    td_t* in = xnode(x, op->inputs[0]);
    if (TD_IS_ERR(in)) return in;
    td_t* r;
    if (op->opcode == OP_COUNT) {
        /* the reference counts a table's columns (exec.c:1734-1739) */
        r = td_i64(in->type == TD_TABLE ? td_table_ncols(in) : in->len);
        td_release(in);
        return r;
    }
    if (in->type <= 0 || in->type == TD_TABLE) { td_release(in); return TD_ERR_PTR(TD_ERR_TYPE); }
    if (op->opcode == OP_PROD) {
        /* exec_reduction's product (exec.c:1779, 1812): f64 for an F64
         * input, else the wrapping i64 product of the widened values; 1 for
         * no rows */
        tdd_col_t c;
        uint64_t v = 0;
        int rc = dev_col(x, in, &c);
        if (!rc) rc = tdd_reduce_prod(&c, in->len, &v);
        int isf = in->type == TD_F64;
        td_release(in);
        if (rc) return TD_ERR_PTR(rc);
        if (isf) {
            double d;
            memcpy(&d, &v, 8);
            return td_f64(d);
        }
        return td_i64((int64_t)v);
    }
    if (op->opcode == OP_COUNT_DISTINCT) {
        r = count_distinct(x, in);
        td_release(in);
        return r;
    }
    /* One group keyed by a constant: a zero-filled U8 key column. */
    int64_t n = in->len;
    tdd_group_spec_t S;
    memset(&S, 0, sizeof(S));
    S.nrows = n;
    S.n_keys = 1;
    S.n_aggs = 1;
    S.ops[0] = op->opcode;
    int rc = dev_col(x, in, &S.aggs[0]);
    void* zk = NULL;
    if (!rc) {
        zk = td__arena_alloc((size_t)(n > 0 ? n : 1));
        rc = zk ? tdd_fill(zk, 0, (size_t)(n > 0 ? n : 1)) : TD_ERR_OOM;
        if (zk && keep_tmp(x, zk)) rc = TD_ERR_OOM;
    }
    S.keys[0].data = zk;
    S.keys[0].type = TD_U8;
    S.keys[0].esz = 1;
    S.force_path = 1;
    tdd_group_result_t R;
    if (!rc) rc = tdd_group(&S, &R);
    int8_t it = in->type;
    td_release(in);
    if (rc) return TD_ERR_PTR(rc);
    uint64_t v = 0;
    int8_t ot = R.out_types[0];
    if (R.n_groups > 0) {
        uint8_t esz = ot == TD_F64 || ot == TD_I64 ? 8 : td_type_sizes[ot > 0 && ot < TD_TYPE_COUNT ? ot : TD_I64];
        rc = tdd_d2h(&v, R.vals[0], esz);
    }
    td__arena_free(R.keys[0]);
    td__arena_free(R.vals[0]);
    if (rc) return TD_ERR_PTR(rc);
    if (ot == TD_F64) {
        double d;
        memcpy(&d, &v, 8);
        /* no rows: the reference's accumulator defaults (exec.c:1808-1832) --
         * 0 for SUM / MIN / MAX / AVG / FIRST / LAST, NaN for the variances */
        if (R.n_groups == 0)
            d = (op->opcode == OP_VAR || op->opcode == OP_VAR_POP || op->opcode == OP_STDDEV ||
                 op->opcode == OP_STDDEV_POP) ? __builtin_nan("") : 0.0;
        return td_f64(d);
    }
    int64_t i = (int64_t)v;
    if (ot == TD_I32 || ot == TD_DATE || ot == TD_TIME) i = (int32_t)(uint32_t)v;
    else if (ot == TD_I16) i = (int16_t)(uint16_t)v;
    else if (ot == TD_U8 || ot == TD_BOOL) i = (uint8_t)v;
    (void)it;
    return td_i64(i);
}

/* First / last n rows of a vector, as a device copy. */
static td_t* slice_vec(xctx_t* x, td_t* v, int64_t n, int tail) {
    tdd_col_t c;
    int rc = dev_col(x, v, &c);
    if (rc) return TD_ERR_PTR(rc);
    if (n > v->len) n = v->len;
    if (n < 0) n = 0;
    int64_t from = tail ? v->len - n : 0;
    td_t* out = td__hbm_vec_new(v->type, v->attrs & TD_SYM_W_MASK, n, (size_t)n * c.esz);
    if (TD_IS_ERR(out)) return out;
    rc = tdd_d2d(td_hbm_ref(out)->dptr, (const char*)c.data + from * c.esz, (size_t)n * c.esz);
    if (rc) { td_release(out); return TD_ERR_PTR(rc); }
    return out;
}

/* HEAD / TAIL of a table or vector (exec.c:11919-12140 without the fused
 * sort / group / filter variants). */
/* ---------------------------------------------------------------- sort -- */

/* Symbol ids 0..n_ids-1 ordered by their strings (bytes of the common
 * prefix, then length; ids without a string first), ties by id -- the
 * order build_enum_rank gives (exec.c:2876-2970); rank[id] = position.
 * A stable bottom-up merge sort. */
static int sym_str_cmp(uint32_t a, uint32_t b) {
    td_t* sa = td_sym_str(a);
    td_t* sb = td_sym_str(b);
    size_t la = sa ? td_str_len(sa) : 0, lb = sb ? td_str_len(sb) : 0;
    size_t ml = la < lb ? la : lb;
    int c = ml ? memcmp(td_str_ptr(sa), td_str_ptr(sb), ml) : 0;
    if (c == 0) c = (la > lb) - (la < lb);
    return c;
}

static uint32_t* sym_ranks(uint32_t n_ids) {
    uint32_t* ids = malloc(sizeof(uint32_t) * n_ids);
    uint32_t* tmp = malloc(sizeof(uint32_t) * n_ids);
    uint32_t* rank = malloc(sizeof(uint32_t) * n_ids);
    if (!ids || !tmp || !rank) { free(ids); free(tmp); free(rank); return NULL; }
    for (uint32_t i = 0; i < n_ids; i++) ids[i] = i;
    for (uint32_t w = 1; w < n_ids; w *= 2) {
        for (uint32_t lo = 0; lo < n_ids; lo += 2 * w) {
            uint32_t mid = lo + w < n_ids ? lo + w : n_ids;
            uint32_t hi = lo + 2 * w < n_ids ? lo + 2 * w : n_ids;
            uint32_t a = lo, b = mid, k = lo;
            while (a < mid && b < hi) tmp[k++] = sym_str_cmp(ids[a], ids[b]) <= 0 ? ids[a++] : ids[b++];
            while (a < mid) tmp[k++] = ids[a++];
            while (b < hi) tmp[k++] = ids[b++];
        }
        uint32_t* t = ids; ids = tmp; tmp = t;
    }
    for (uint32_t i = 0; i < n_ids; i++) rank[ids[i]] = i;
    free(ids);
    free(tmp);
    return rank;
}

/* SORT of a table by its SCAN keys (exec.c:3370-3862): a device sort
 * permutation, then every column gathered through it; limit >= 0 keeps the
 * first `limit` rows (HEAD over SORT, exec.c:11895-11917), -1 all. */
static td_t* exec_sort(xctx_t* x, td_op_t* op, int64_t limit) {
    td_graph_t* g = x->g;
    td_op_ext_t* ext = td__find_ext(g, op->id);
    if (!ext || ext->sort.n_cols < 1 || ext->sort.n_cols > 16) return TD_ERR_PTR(TD_ERR_NYI);
    td_t* in = xnode(x, op->inputs[0]);
    if (TD_IS_ERR(in)) return in;
    td_t* tbl = in->type == TD_TABLE ? in : g->table;
    if (!tbl) { td_release(in); return TD_ERR_PTR(TD_ERR_SCHEMA); }
    td_retain(tbl);
    td_release(in);
    if (g->selection && td_table_nrows(tbl) == g->selection->len) {
        td_t* c = compact_table(x, tbl, g->selection);
        td_release(g->selection);
        g->selection = NULL;
        td_release(tbl);
        if (TD_IS_ERR(c)) return c;
        tbl = c;
    }
    const int nk = ext->sort.n_cols;
    const int64_t n = td_table_nrows(tbl);
    tdd_sort_key_t K[16];
    void* ranks[16] = {0};
    memset(K, 0, sizeof(K));
    int rc = TD_OK;
    td_t* kev[16] = {0};                         /* expression keys, evaluated over the input */
    for (int k = 0; k < nk && !rc; k++) {
        td_op_t* kop = ext->sort.columns[k];
        td_op_ext_t* ke = td__find_ext(g, kop->id);
        td_t* col = (ke && ke->base.opcode == OP_SCAN) ? td_table_get_col(tbl, ke->sym) : NULL;
        if (!col && expr_elementwise(kop->opcode)) {
            /* exec.c:3395-3403: the key evaluated with the sort input as table */
            td_t* e = eval_expr(x, tbl, kop);
            if (TD_IS_ERR(e)) { rc = TD_ERR_CODE(e); break; }
            col = kev[k] = e;
        }
        if (!col) { rc = kop->opcode == OP_SCAN ? TD_ERR_SCHEMA : TD_ERR_NYI; break; }
        rc = dev_col(x, col, &K[k].col);
        K[k].desc = ext->sort.desc ? ext->sort.desc[k] : 0;
        K[k].nulls_first = ext->sort.nulls_first ? ext->sort.nulls_first[k] : K[k].desc;
        if (!rc && col->type == TD_SYM && n > 0) {
            int64_t mn, mx;
            rc = tdd_minmax(&K[k].col, n, NULL, &mn, &mx);
            if (!rc && (mx < 0 || mx >= (int64_t)UINT32_MAX - 1)) rc = TD_ERR_NYI;
            uint32_t* hr = rc ? NULL : sym_ranks((uint32_t)mx + 1);
            if (!rc && !hr) rc = TD_ERR_OOM;
            if (!rc) {
                ranks[k] = td__arena_alloc(sizeof(uint32_t) * ((size_t)mx + 1));
                if (!ranks[k]) rc = TD_ERR_OOM;
                else rc = tdd_h2d(ranks[k], hr, sizeof(uint32_t) * ((size_t)mx + 1));
                K[k].rank = ranks[k];
            }
            free(hr);
        }
    }
    int64_t* perm = rc ? NULL : td__arena_alloc((size_t)(n > 0 ? n : 1) * 8);
    if (!rc && !perm) rc = TD_ERR_OOM;
    if (!rc) rc = tdd_sort_perm(K, nk, n, perm);
    td_t* out = TD_ERR_PTR(rc ? rc : TD_ERR_OOM);
    if (!rc) {
        const int64_t m = limit >= 0 && limit < n ? limit : n;
        int64_t nc = td_table_ncols(tbl);
        out = td_table_new(nc);
        for (int64_t c = 0; c < nc && !TD_IS_ERR(out); c++) {
            td_t* v = gather_col(x, td_table_get_col_idx(tbl, c), perm, m, 0);
            if (TD_IS_ERR(v)) { td_release(out); out = v; break; }
            out = td_table_add_col(out, td_table_col_name(tbl, c), v);
            td_release(v);
        }
    }
    if (perm) { tdd_sync(); td__arena_free(perm); }
    for (int k = 0; k < nk; k++) {
        if (ranks[k]) td__arena_free(ranks[k]);
        if (kev[k]) td_release(kev[k]);
    }
    td_release(tbl);
    td__set_plan("sort keys=%d rows=%lld", nk, (long long)n);
    return out;
}

static td_t* exec_head_tail(xctx_t* x, td_op_t* op) {
    td_op_ext_t* e = td__find_ext(x->g, op->id);
    int64_t n = e ? e->sym : 10;
    if (op->opcode == OP_HEAD && op->inputs[0]->opcode == OP_SORT)
        return exec_sort(x, op->inputs[0], n > 0 ? n : 0);      /* top-N: a prefix */

    td_t* in = xnode(x, op->inputs[0]);
    if (TD_IS_ERR(in)) return in;
    int tail = op->opcode == OP_TAIL;
    td_t* out;
    if (in->type == TD_TABLE) {
        if (x->g->selection && td_table_nrows(in) == x->g->selection->len) {
            td_t* c = compact_table(x, in, x->g->selection);
            td_release(x->g->selection);
            x->g->selection = NULL;
            td_release(in);
            if (TD_IS_ERR(c)) return c;
            in = c;
        }
        int64_t nc = td_table_ncols(in);
        out = td_table_new(nc);
        for (int64_t c = 0; c < nc && !TD_IS_ERR(out); c++) {
            td_t* s = slice_vec(x, td_table_get_col_idx(in, c), n, tail);
            if (TD_IS_ERR(s)) { td_release(out); out = s; break; }
            out = td_table_add_col(out, td_table_col_name(in, c), s);
            td_release(s);
        }
    } else if (in->type > 0) {
        out = slice_vec(x, in, n, tail);
    } else {
        out = TD_ERR_PTR(TD_ERR_TYPE);
    }
    td_release(in);
    return out;
}

/* FILTER over GROUP (HAVING): the predicate's SCANs name columns of the
 * group result, and the groups that pass come out compacted, in group
 * order (exec.c:11674-11704 swaps the graph's table to the group output
 * while it evaluates the predicate). */
static td_t* exec_having(xctx_t* x, td_op_t* op) {
    td_t* gr = xnode(x, op->inputs[0]);
    if (TD_IS_ERR(gr)) return gr;
    if (gr->type != TD_TABLE) { td_release(gr); return TD_ERR_PTR(TD_ERR_TYPE); }
    tdd_pred_t P;
    int rc = compile_pred(x, gr, op->inputs[1], &P);
    td_t* s = rc == TD_ERR_NYI ? pred_expr_sel(x, gr, op->inputs[1], NULL)   /* arithmetic */
            : rc ? TD_ERR_PTR(rc) : eval_sel(x, &P, td_table_nrows(gr), NULL);
    if (TD_IS_ERR(s)) { td_release(gr); return s; }
    td_t* out = compact_table(x, gr, s);
    td_release(s);
    td_release(gr);
    return out;
}

/* ---------------------------------------------------------- dispatcher -- */

static td_t* xnode(xctx_t* x, td_op_t* op) {
    td_graph_t* g = x->g;
    if (!op) return TD_ERR_PTR(TD_ERR_NYI);
    if (td__cancel_poll()) return TD_ERR_PTR(TD_ERR_CANCEL);
    switch (op->opcode) {
    case OP_SCAN: {
        td_op_ext_t* e = td__find_ext(g, op->id);
        if (!e) return TD_ERR_PTR(TD_ERR_NYI);
        td_t* tbl = scan_table(g, e);
        if (!tbl) return TD_ERR_PTR(TD_ERR_SCHEMA);
        td_t* col = td_table_get_col(tbl, e->sym);
        if (!col) return TD_ERR_PTR(TD_ERR_SCHEMA);
        if (TD_IS_PARTED(col->type) || col->type == TD_MAPCOMMON) return TD_ERR_PTR(TD_ERR_NYI);
        if (col->mmod != TD_MMOD_HBM) x->all_resident = 0;
        td_retain(col);
        return col;
    }
    case OP_CONST: {
        td_op_ext_t* e = td__find_ext(g, op->id);
        if (!e || !e->literal) return TD_ERR_PTR(TD_ERR_NYI);
        if (e->literal->type == TD_TABLE) return flat_table(x, e->literal);
        td_retain(e->literal);
        return e->literal;
    }
    case OP_FILTER: {
        if (op->inputs[0]->opcode == OP_GROUP) return exec_having(x, op);
        td_t* input = xnode(x, op->inputs[0]);
        if (TD_IS_ERR(input)) return input;
        tdd_pred_t P;
        int rc;
        if (input->type == TD_TABLE) {
            td_t* tbl = g->table;
            if (!tbl) { td_release(input); return TD_ERR_PTR(TD_ERR_SCHEMA); }
            rc = compile_pred(x, tbl, op->inputs[1], &P);
            if (rc == TD_ERR_NYI) {
                /* arithmetic in the predicate: evaluate it as an expression */
                td_t* s = pred_expr_sel(x, tbl, op->inputs[1], g->selection);
                if (TD_IS_ERR(s)) { td_release(input); return s; }
                if (g->selection) td_release(g->selection);
                g->selection = s;
                return input;
            }
            if (rc) { td_release(input); return TD_ERR_PTR(rc); }
            td_t* s = eval_sel(x, &P, td_table_nrows(tbl), g->selection);
            if (TD_IS_ERR(s)) { td_release(input); return s; }
            if (g->selection) td_release(g->selection);
            g->selection = s;
            return input;
        }
        /* vector: eager compaction */
        if (!g->table || td_table_nrows(g->table) != input->len) {
            td_release(input);
            return TD_ERR_PTR(TD_ERR_LENGTH);
        }
        rc = compile_pred(x, g->table, op->inputs[1], &P);
        td_t* s = rc == TD_ERR_NYI ? pred_expr_sel(x, g->table, op->inputs[1], NULL)
                  : rc ? TD_ERR_PTR(rc) : eval_sel(x, &P, input->len, NULL);
        if (TD_IS_ERR(s)) { td_release(input); return s; }
        td_t* out = compact_vec(x, input, s);
        td_release(s);
        td_release(input);
        return out;
    }
    case OP_GROUP:
        return exec_group(x, op);
    case OP_JOIN:
        return exec_join(x, op);
    case OP_SUM: case OP_MIN: case OP_MAX: case OP_COUNT: case OP_AVG:
    case OP_FIRST: case OP_LAST: case OP_STDDEV: case OP_STDDEV_POP:
    case OP_VAR: case OP_VAR_POP: case OP_PROD: case OP_COUNT_DISTINCT:
        return exec_reduce(x, op);
    case OP_ALIAS: case OP_MATERIALIZE:
        return xnode(x, op->inputs[0]);
    case OP_HEAD: case OP_TAIL:
        return exec_head_tail(x, op);
    case OP_SORT:
        return exec_sort(x, op, -1);
    default:
        /* element-wise subtree over the graph's table, fused into one kernel
         * (exec.c:11620-11640) */
        if (expr_elementwise(op->opcode)) {
            if (!g->table) return TD_ERR_PTR(TD_ERR_SCHEMA);
            return eval_expr(x, g->table, op);
        }
        return TD_ERR_PTR(TD_ERR_NYI);
    }
}

td_t* td_execute(td_graph_t* g, td_op_t* root) {
    if (!g || !root) return TD_ERR_PTR(TD_ERR_NYI);
    if (!td__dev_ready()) return TD_ERR_PTR(TD_ERR_IO);     /* no GPU: fail loudly */
    td__cancel_reset();
    xctx_t x;
    memset(&x, 0, sizeof(x));
    x.g = g;
    x.all_resident = 1;
    td__set_plan("-");
    /* parted tables run as their flat forms */
    td_t* saved = g->table;
    td_t* saved_tabs[8];
    int nt = g->tables && g->n_tables <= 8 ? (int)g->n_tables : 0;
    td_t* r = NULL;
    if (g->table) {
        td_t* f = flat_table(&x, g->table);
        if (TD_IS_ERR(f)) r = f;
        else { td_release(f); g->table = f; }     /* held by the execution's cache */
    }
    for (int i = 0; i < nt; i++) {
        saved_tabs[i] = g->tables[i];
        if (r || !g->tables[i]) continue;
        td_t* f = flat_table(&x, g->tables[i]);
        if (TD_IS_ERR(f)) r = f;
        else { td_release(f); g->tables[i] = f; }
    }
    td_t* sel_in = g->selection;
    if (sel_in) td_retain(sel_in);
    if (!r) r = xnode(&x, root);
    if (TD_IS_ERR(r) && g->selection != sel_in) {
        /* a failed (e.g. cancelled) query drops the selection its FILTERs
         * made and leaves the graph's own as it was */
        if (g->selection) td_release(g->selection);
        g->selection = sel_in;
        sel_in = NULL;
    }
    if (sel_in) td_release(sel_in);
    if (!TD_IS_ERR(r) && g->selection && r->type == TD_TABLE) {
        td_t* c = compact_table(&x, r, g->selection);      /* final compaction */
        td_release(r);
        td_release(g->selection);
        g->selection = NULL;
        r = c;
    }
    if (!TD_IS_ERR(r) && !x.all_resident) r = td__to_host(r);
    if (!TD_IS_ERR(r) && tdd_sync() != TD_OK) { td_release(r); r = TD_ERR_PTR(TD_ERR_IO); }
    g->table = saved;
    for (int i = 0; i < nt; i++) g->tables[i] = saved_tabs[i];
    drop_tmps(&x);
    return r;
}

/* ------------------------------------------------------ batched queries -- */

typedef struct {
    tdd_join_group_spec_t S;
    tdd_pred_t   P;
    td_t*        key_src[TDD_MAX_KEYS];
    td_t*        agg_src[TDD_MAX_AGGS];
    td_op_ext_t* ext;
    int          ok;
} jg_item_t;

/* The spec of a DAG the batch can take: root GROUP over the graph's inner
 * JOIN of FILTER*(CONST fact) (one predicate at most) with CONST dim, no
 * pending selection, every referenced column resident (nothing uploaded). */
static int jg_plan(td_graph_t* g, td_op_t* root, jg_item_t* it) {
    memset(it, 0, sizeof(*it));
    if (!g || !root || root->opcode != OP_GROUP || g->selection || !g->table) return TD_ERR_NYI;
    td_op_ext_t* ext = td__find_ext(g, root->id);
    if (!ext || ext->n_keys < 1 || ext->n_keys > TDD_MAX_KEYS || ext->n_aggs > TDD_MAX_AGGS)
        return TD_ERR_NYI;
    td_op_t* jop = NULL;
    for (uint32_t i = 0; i < g->node_count && !jop; i++)
        if (g->nodes[i].opcode == OP_JOIN) jop = &g->nodes[i];
    if (!jop) return TD_ERR_NYI;
    td_op_ext_t* je = td__find_ext(g, jop->id);
    td_op_t* preds[8];
    td_op_t* rpreds[8];
    int npred = 0, nrp = 0;
    td_t* left = const_table_under(g, jop->inputs[0], preds, &npred);
    td_t* right = const_table_under(g, jop->inputs[1], rpreds, &nrp);
    if (!je || je->join.join_type != 0 || !left || !right || nrp || npred > 1)
        return TD_ERR_NYI;
    if (npred && td_table_nrows(g->table) != td_table_nrows(left)) return TD_ERR_NYI;
    xctx_t x;
    memset(&x, 0, sizeof(x));
    x.g = g;
    x.all_resident = 1;
    tdd_join_group_spec_t* S = &it->S;
    S->join.nl = td_table_nrows(left);
    S->join.nr = td_table_nrows(right);
    int rc = join_keys(&x, je, left, right, &S->join);
    if (!rc && npred) rc = compile_pred(&x, g->table, preds[0], &it->P);
    if (!rc && npred) S->lpred = &it->P;
    S->n_keys = ext->n_keys;
    S->n_aggs = ext->n_aggs;
    for (int i = 0; i < ext->n_keys + ext->n_aggs && !rc; i++) {
        int is_key = i < ext->n_keys;
        td_op_t* in = is_key ? ext->keys[i] : ext->agg_ins[i - ext->n_keys];
        td_t* col = NULL;
        int side = 0;
        if (in->opcode == OP_SCAN) {
            td_op_ext_t* se = td__find_ext(g, in->id);
            col = td_table_get_col(left, se->sym);
            if (!col && !join_is_right_key(g, je, se->sym)) { col = td_table_get_col(right, se->sym); side = 1; }
            if (!col) rc = TD_ERR_SCHEMA;
        } else if (is_key || !(in->opcode == OP_CONST && ext->agg_ops[i - ext->n_keys] == OP_COUNT)) {
            rc = TD_ERR_NYI;
        }
        if (rc) break;
        if (is_key) {
            it->key_src[i] = col;
            S->key_side[i] = (uint8_t)side;
            rc = dev_col(&x, col, &S->keys[i]);
        } else {
            int a = i - ext->n_keys;
            it->agg_src[a] = col;
            S->ops[a] = ext->agg_ops[a];
            S->agg_side[a] = (uint8_t)side;
            if (col) rc = dev_col(&x, col, &S->aggs[a]);
        }
    }
    int uploaded = x.ntmp > 0 || x.nup > 0 || !x.all_resident;
    drop_tmps(&x);
    if (!rc && uploaded) rc = TD_ERR_NYI;
    it->ext = ext;
    it->ok = rc == 0;
    return rc;
}

/* Result table of batched query i: views into the batch's shared output
 * blocks (owners) at rows [off, off + n). */
static td_t* jg_view_table(td_graph_t* g, const jg_item_t* it, td_t* const* owners, int64_t off,
                           int64_t n) {
    td_op_ext_t* ext = it->ext;
    int nk = ext->n_keys, na = ext->n_aggs;
    td_t* out = td_table_new(nk + na);
    for (int c = 0; c < nk + na && !TD_IS_ERR(out); c++) {
        td_t* own = owners[c];
        td_t* src = c < nk ? it->key_src[c] : it->agg_src[c - nk];
        int8_t t = own->type;
        uint8_t attrs = own->attrs;
        size_t esz = td__esz(own);
        td_t* v = td__hbm_view(own, t, attrs, n, (size_t)off * esz, (size_t)n * esz);
        (void)src;
        if (TD_IS_ERR(v)) { td_release(out); return v; }
        int64_t name = c < nk ? key_name(g, ext->keys[c], c)
                              : agg_name(g, ext->agg_ins[c - nk], c - nk, ext->agg_ops[c - nk]);
        out = td_table_add_col(out, name, v);
        td_release(v);
    }
    return out;
}

struct td_dev_batch {
    int64_t          n;
    td_graph_t**     graphs;
    td_op_t**        roots;
    jg_item_t*       items;
    int*             bidx;        /* query -> batch index, -1: runs alone */
    int              nb;
    tdd_jg_batch_t*  dev;
};

td_dev_batch_t* td_dev_batch_prepare(td_graph_t* const* graphs, td_op_t* const* roots,
                                     int64_t n) {
    if (n <= 0 || n > INT32_MAX || !td__dev_ready()) return NULL;
    td_dev_batch_t* b = calloc(1, sizeof(*b));
    if (!b) return NULL;
    b->n = n;
    b->graphs = malloc(sizeof(td_graph_t*) * (size_t)n);
    b->roots = malloc(sizeof(td_op_t*) * (size_t)n);
    b->items = calloc((size_t)n, sizeof(jg_item_t));
    b->bidx = malloc(sizeof(int) * (size_t)n);
    const tdd_join_group_spec_t** specs = malloc(sizeof(void*) * (size_t)n);
    if (!b->graphs || !b->roots || !b->items || !b->bidx || !specs) {
        free(specs);
        td_dev_batch_free(b);
        return NULL;
    }
    for (int64_t i = 0; i < n; i++) {
        b->graphs[i] = graphs[i];
        b->roots[i] = roots[i];
        b->bidx[i] = -1;
        if (jg_plan(graphs[i], roots[i], &b->items[i]) == TD_OK) {
            b->bidx[i] = b->nb;
            specs[b->nb++] = &b->items[i].S;
        }
    }
    /* lpred points into the item: fix up after the items array is final */
    for (int64_t i = 0; i < n; i++)
        if (b->items[i].ok && b->items[i].S.lpred) b->items[i].S.lpred = &b->items[i].P;
    if (b->nb > 0 && tdd_jg_batch_prepare(specs, b->nb, 1, &b->dev) != TD_OK) b->dev = NULL;
    free(specs);
    if (b->dev) {
        for (int64_t i = 0; i < n; i++)
            if (b->bidx[i] >= 0 && tdd_jg_batch_status(b->dev, b->bidx[i]) != 0) b->bidx[i] = -1;
    } else {
        for (int64_t i = 0; i < n; i++) b->bidx[i] = -1;
    }
    return b;
}

int64_t td_dev_batch_info(const td_dev_batch_t* b, int64_t* tile_rows, int* block) {
    *tile_rows = 0;
    *block = 0;
    if (!b) return 0;
    int64_t c = 0;
    for (int64_t i = 0; i < b->n; i++) c += b->bidx[i] >= 0;
    if (b->dev) tdd_jg_batch_shape(b->dev, tile_rows, block);
    return c;
}

int64_t td_dev_batch_radix(const td_dev_batch_t* b, int64_t* part_rows, int* max_bits) {
    *part_rows = 0;
    *max_bits = 0;
    if (!b || !b->dev) return 0;
    return tdd_jg_batch_radix(b->dev, part_rows, max_bits);
}

static td_err_t batch_run(td_dev_batch_t* b, int use_graph, td_t** results) {
    td_err_t err = TD_OK;
    for (int64_t i = 0; i < b->n; i++) results[i] = NULL;
    if (b->dev) {
        tdd_group_result_t R;
        int64_t* offs = malloc(sizeof(int64_t) * (size_t)(b->nb + 1));
        if (!offs) return TD_ERR_OOM;
        int rc = tdd_jg_batch_run(b->dev, use_graph, &R, offs);
        if (rc) { free(offs); return (td_err_t)rc; }
        /* owners of the shared output blocks, typed like the first batched
         * query's columns */
        const jg_item_t* it0 = NULL;
        for (int64_t i = 0; i < b->n && !it0; i++) if (b->bidx[i] >= 0) it0 = &b->items[i];
        td_t* owners[TDD_MAX_KEYS + TDD_MAX_AGGS] = {0};
        int nk = it0 ? it0->ext->n_keys : 0, na = it0 ? it0->ext->n_aggs : 0;
        int64_t tot = offs[b->nb];
        for (int c = 0; c < nk + na; c++) {
            td_t* s = c < nk ? it0->key_src[c] : it0->agg_src[c - nk];
            int8_t t;
            uint8_t attrs = 0;
            size_t esz = 8;
            if (c < nk) {
                t = s->type;
                attrs = s->attrs & TD_SYM_W_MASK;
                esz = td__esz(s);
            } else {
                t = R.out_types[c - nk];
                if (t != TD_F64 && t != TD_I64 && s) { attrs = s->attrs & TD_SYM_W_MASK; esz = td__esz(s); }
            }
            void* blk = c < nk ? R.keys[c] : R.vals[c - nk];
            owners[c] = td__hbm_wrap(t, attrs, tot, blk, (size_t)(tot > 0 ? tot : 1) * esz);
            if (TD_IS_ERR(owners[c])) { err = TD_ERR_OOM; owners[c] = NULL; }
        }
        for (int64_t i = 0; i < b->n && !err; i++) {
            int q = b->bidx[i];
            if (q < 0) continue;
            results[i] = jg_view_table(b->graphs[i], &b->items[i], owners, offs[q],
                                       offs[q + 1] - offs[q]);
            if (TD_IS_ERR(results[i])) err = TD_ERR_OOM;
        }
        for (int c = 0; c < nk + na; c++) if (owners[c]) td_release(owners[c]);
        free(offs);
    }
    for (int64_t i = 0; i < b->n; i++) {
        if (results[i]) continue;
        results[i] = td_execute(b->graphs[i], b->roots[i]);
        if (TD_IS_ERR(results[i]) && !err) err = (td_err_t)(intptr_t)results[i];
    }
    td__set_plan("batch queries=%lld batched=%d", (long long)b->n, b->dev ? b->nb : 0);
    return err;
}

td_err_t td_dev_batch_run(td_dev_batch_t* b, td_t** results) {
    if (!b || !results) return TD_ERR_NYI;
    return batch_run(b, 1, results);
}

void td_dev_batch_free(td_dev_batch_t* b) {
    if (!b) return;
    if (b->dev) tdd_jg_batch_free(b->dev);
    free(b->graphs);
    free(b->roots);
    free(b->items);
    free(b->bidx);
    free(b);
}

td_err_t td_dev_execute_batch(td_graph_t* const* graphs, td_op_t* const* roots, int64_t n,
                              td_t** results) {
    if (!td__dev_ready()) return TD_ERR_IO;
    td_dev_batch_t* b = td_dev_batch_prepare(graphs, roots, n);
    if (!b) return TD_ERR_OOM;
    td_err_t e = batch_run(b, 0, results);
    td_dev_batch_free(b);
    return e;
}

/* ------------------------------------------------------ served requests -- */
/* A request loop over a resident fact / dim pair: DAGs of the served family
 * (see tdd_srv_* in tdd.h) share the device scan; others run through
 * td_execute when admitted.  Requests are admitted in arrival order while
 * request slots and the device-memory budget allow, join the running scan
 * at the next step and finish after one pass over every chunk. */

typedef struct {
    int64_t        id;
    td_graph_t*    g;
    td_op_t*       root;
    jg_item_t      it;
    tdd_srv_req_t  rq;
    int            shared;       /* 1: answered by the shared scan */
    int64_t        bytes;
    int64_t        steps_left;
    tdd_srv_acc_t* acc;
} srv_item_t;

typedef struct { int64_t id; td_t* result; } srv_done_t;

struct td_dev_server {
    tdd_srv_t*      dev;
    tdd_srv_spec_t  spec;
    td_t*           fact;
    td_t*           dim;
    uint64_t        budget, base, held;
    int             max_active;
    int64_t         chunks, next_id, steps, n_shared, n_alone;
    srv_item_t**    queue;       /* ring of pending requests */
    int64_t         q_head, q_len, q_cap;
    srv_item_t**    active;
    int             n_active;
    srv_done_t*     done;        /* ring of finished results */
    int64_t         d_head, d_len, d_cap;
};

static uint64_t table_dev_bytes(td_t* t) {
    uint64_t b = 0;
    for (int64_t c = 0; c < td_table_ncols(t); c++) {
        td_t* col = td_table_get_col_idx(t, c);
        if (col && !TD_IS_ERR(col)) b += (uint64_t)col->len * td__esz(col);
    }
    return b;
}

static int srv_col(td_t* tbl, int64_t sym, tdd_col_t* out) {
    td_t* col = td_table_get_col(tbl, sym);
    if (!col || col->mmod != TD_MMOD_HBM) return TD_ERR_SCHEMA;
    xctx_t x;
    memset(&x, 0, sizeof(x));
    x.all_resident = 1;
    int rc = dev_col(&x, col, out);
    drop_tmps(&x);
    return rc;
}

td_dev_server_t* td_dev_serve_open(const td_dev_serve_cfg_t* cfg) {
    if (!cfg || !td__dev_ready() || !cfg->fact || !cfg->dim || !td_dev_is_resident(cfg->fact) ||
        !td_dev_is_resident(cfg->dim) || cfg->n_vals < 0 || cfg->n_vals > TDD_SRV_MAX_VALS ||
        cfg->n_keys < 1 || cfg->n_keys > TDD_SRV_MAX_KEYS)
        return NULL;
    td_dev_server_t* s = calloc(1, sizeof(*s));
    if (!s) return NULL;
    tdd_srv_spec_t* S = &s->spec;
    S->nf = td_table_nrows(cfg->fact);
    S->nd = td_table_nrows(cfg->dim);
    S->n_vals = cfg->n_vals;
    S->n_keys = cfg->n_keys;
    S->chunk_rows = cfg->chunk_rows;
    int rc = srv_col(cfg->fact, cfg->fk, &S->fk);
    if (!rc) rc = srv_col(cfg->dim, cfg->pk, &S->pk);
    if (!rc) rc = srv_col(cfg->fact, cfg->pcol, &S->pcol);
    for (int j = 0; j < cfg->n_vals && !rc; j++) rc = srv_col(cfg->fact, cfg->vals[j], &S->vals[j]);
    for (int k = 0; k < cfg->n_keys && !rc; k++) rc = srv_col(cfg->dim, cfg->keys[k], &S->keys[k]);
    if (!rc) rc = tdd_srv_open(S, &s->dev);
    s->q_cap = 64;
    s->d_cap = 64;
    s->max_active = cfg->max_active > 0 ? cfg->max_active : 4096;
    s->queue = malloc(sizeof(srv_item_t*) * (size_t)s->q_cap);
    s->active = malloc(sizeof(srv_item_t*) * (size_t)s->max_active);
    s->done = malloc(sizeof(srv_done_t) * (size_t)s->d_cap);
    if (rc || !s->queue || !s->active || !s->done) {
        if (s->dev) tdd_srv_close(s->dev);
        free(s->queue);
        free(s->active);
        free(s->done);
        free(s);
        return NULL;
    }
    td_retain(cfg->fact);
    td_retain(cfg->dim);
    s->fact = cfg->fact;
    s->dim = cfg->dim;
    s->chunks = tdd_srv_chunks(s->dev);
    s->base = table_dev_bytes(cfg->fact) + table_dev_bytes(cfg->dim) +
              (uint64_t)tdd_srv_bytes(s->dev);
    s->held = s->base;
    s->budget = cfg->budget;
    return s;
}

/* The served-family request of a planned DAG, or -1. */
static int srv_match(const td_dev_server_t* s, srv_item_t* r) {
    const tdd_srv_spec_t* S = &s->spec;
    const tdd_join_group_spec_t* J = &r->it.S;
    tdd_srv_req_t* q = &r->rq;
    memset(q, 0, sizeof(*q));
    if (J->join.n_keys != 1 || J->join.nl != S->nf || J->join.nr != S->nd ||
        J->join.lkeys[0].data != S->fk.data || J->join.rkeys[0].data != S->pk.data)
        return -1;
    if (J->lpred) {
        const tdd_pred_t* P = J->lpred;
        if (P->n_ins == 1 && P->ins[0].kind == TDD_P_CMP_CONST && !P->ins[0].as_f64 &&
            P->cols[P->ins[0].a].data == S->pcol.data) {
            q->cmp = P->ins[0].cmp;             /* a range of the cube's axis */
            q->t = P->ins[0].ci;
        } else {
            q->pred = P;                        /* its own predicate on the shared chunks */
        }
    } else {
        q->cmp = OP_GE;
        q->t = INT64_MIN;
    }
    if (J->n_keys != 1 || J->key_side[0] != 1) return -1;
    q->key = -1;
    for (int k = 0; k < S->n_keys; k++)
        if (J->keys[0].data == S->keys[k].data) q->key = k;
    if (q->key < 0) return -1;
    q->n_aggs = J->n_aggs;
    for (int a = 0; a < J->n_aggs; a++) {
        q->ops[a] = J->ops[a];
        if (J->ops[a] == OP_COUNT) { q->src[a] = -1; continue; }
        if (J->agg_side[a] != 0) return -1;
        q->src[a] = -3;
        if (J->aggs[a].data == S->pcol.data) q->src[a] = -2;
        for (int j = 0; j < S->n_vals; j++)
            if (J->aggs[a].data == S->vals[j].data) q->src[a] = (int8_t)j;
        if (q->src[a] == -3) return -1;
    }
    return tdd_srv_req_bytes(s->dev, q, &r->bytes) == TD_OK ? 0 : -1;
}

static int srv_push_done(td_dev_server_t* s, int64_t id, td_t* res) {
    if (s->d_len == s->d_cap) {
        srv_done_t* n = malloc(sizeof(srv_done_t) * (size_t)s->d_cap * 2);
        if (!n) return TD_ERR_OOM;
        for (int64_t i = 0; i < s->d_len; i++) n[i] = s->done[(s->d_head + i) % s->d_cap];
        free(s->done);
        s->done = n;
        s->d_head = 0;
        s->d_cap *= 2;
    }
    s->done[(s->d_head + s->d_len++) % s->d_cap] = (srv_done_t){id, res};
    return TD_OK;
}

int64_t td_dev_serve_submit(td_dev_server_t* s, td_graph_t* g, td_op_t* root) {
    if (!s || !g || !root) return -TD_ERR_NYI;
    if (s->q_len == s->q_cap) {
        srv_item_t** n = malloc(sizeof(srv_item_t*) * (size_t)s->q_cap * 2);
        if (!n) return -TD_ERR_OOM;
        for (int64_t i = 0; i < s->q_len; i++) n[i] = s->queue[(s->q_head + i) % s->q_cap];
        free(s->queue);
        s->queue = n;
        s->q_head = 0;
        s->q_cap *= 2;
    }
    srv_item_t* r = calloc(1, sizeof(*r));
    if (!r) return -TD_ERR_OOM;
    r->id = s->next_id++;
    r->g = g;
    r->root = root;
    if (jg_plan(g, root, &r->it) == TD_OK) {
        if (r->it.S.lpred) r->it.S.lpred = &r->it.P;
        r->shared = srv_match(s, r) == 0;
    }
    s->queue[(s->q_head + s->q_len++) % s->q_cap] = r;
    return r->id;
}

/* Finished shared requests -> result tables. */
static int srv_finish(td_dev_server_t* s, srv_item_t** fin, int n) {
    if (n <= 0) return TD_OK;
    tdd_group_result_t* R = calloc((size_t)n, sizeof(tdd_group_result_t));
    tdd_srv_acc_t** acc = malloc(sizeof(tdd_srv_acc_t*) * (size_t)n);
    if (!R || !acc) { free(R); free(acc); return TD_ERR_OOM; }
    for (int i = 0; i < n; i++) acc[i] = fin[i]->acc;
    int rc = tdd_srv_finish(s->dev, acc, n, R);
    for (int i = 0; i < n; i++) {
        srv_item_t* r = fin[i];
        td_t* res;
        if (rc) {
            for (int k = 0; k < TDD_MAX_KEYS; k++) td__arena_free(R[i].keys[k]);
            for (int a = 0; a < TDD_MAX_AGGS; a++) td__arena_free(R[i].vals[a]);
            res = TD_ERR_PTR(rc);
        } else {
            res = group_table(r->g, r->it.ext, &R[i], r->it.key_src, r->it.agg_src);
        }
        tdd_srv_acc_free(r->acc);
        s->held -= (uint64_t)r->bytes;
        int e = srv_push_done(s, r->id, res);
        if (e && !rc) rc = e;
        free(r);
    }
    free(R);
    free(acc);
    return rc;
}

int td_dev_serve_step(td_dev_server_t* s) {
    if (!s) return -TD_ERR_NYI;
    int finished = 0, rc = TD_OK;
    /* admission, in arrival order */
    while (s->q_len > 0) {
        srv_item_t* r = s->queue[s->q_head];
        if (!r->shared) {
            td_t* res = td_execute(r->g, r->root);
            s->q_head = (s->q_head + 1) % s->q_cap;
            s->q_len--;
            s->n_alone++;
            finished++;
            rc = srv_push_done(s, r->id, res);
            free(r);
            if (rc) return -rc;
            continue;
        }
        if (s->n_active >= s->max_active) break;
        int fits = !s->budget || s->held + (uint64_t)r->bytes <= s->budget;
        int e = fits ? tdd_srv_begin(s->dev, &r->rq, &r->acc) : TD_ERR_OOM;
        if (e) {
            if (s->n_active > 0) break;        /* wait for active requests to free memory */
            /* nothing left to wait for: the request cannot fit at all */
            s->q_head = (s->q_head + 1) % s->q_cap;
            s->q_len--;
            finished++;
            rc = srv_push_done(s, r->id, TD_ERR_PTR(e));
            free(r);
            if (rc) return -rc;
            continue;
        }
        s->held += (uint64_t)r->bytes;
        r->steps_left = s->chunks;
        s->active[s->n_active++] = r;
        s->q_head = (s->q_head + 1) % s->q_cap;
        s->q_len--;
        s->n_shared++;
    }
    if (s->n_active == 0) return finished;
    tdd_srv_acc_t* accs[256];
    tdd_srv_acc_t** A = s->n_active <= 256 ? accs : malloc(sizeof(void*) * (size_t)s->n_active);
    if (!A) return -TD_ERR_OOM;
    for (int i = 0; i < s->n_active; i++) A[i] = s->active[i]->acc;
    rc = tdd_srv_step(s->dev, A, s->n_active);
    if (A != accs) free(A);
    if (rc) return -rc;
    s->steps++;
    /* requests that have now seen every chunk */
    srv_item_t* fin_buf[256];
    srv_item_t** fin = s->n_active <= 256 ? fin_buf : malloc(sizeof(void*) * (size_t)s->n_active);
    if (!fin) return -TD_ERR_OOM;
    int nf = 0, keep = 0;
    for (int i = 0; i < s->n_active; i++) {
        srv_item_t* r = s->active[i];
        if (--r->steps_left == 0) fin[nf++] = r;
        else s->active[keep++] = r;
    }
    s->n_active = keep;
    rc = srv_finish(s, fin, nf);
    if (fin != fin_buf) free(fin);
    if (rc) return -rc;
    return finished + nf;
}

int td_dev_serve_poll(td_dev_server_t* s, int64_t* id, td_t** result) {
    if (!s || s->d_len == 0) return 0;
    srv_done_t d = s->done[s->d_head];
    s->d_head = (s->d_head + 1) % s->d_cap;
    s->d_len--;
    *id = d.id;
    *result = d.result;
    return 1;
}

void td_dev_serve_info(const td_dev_server_t* s, td_dev_serve_info_t* out) {
    memset(out, 0, sizeof(*out));
    if (!s) return;
    out->queued = s->q_len;
    out->active = s->n_active;
    out->ready = s->d_len;
    out->steps = s->steps;
    out->chunks = s->chunks;
    out->shared = s->n_shared;
    out->alone = s->n_alone;
    out->held = s->held;
    out->base = s->base;
    out->budget = s->budget;
}

void td_dev_serve_close(td_dev_server_t* s) {
    if (!s) return;
    for (int i = 0; i < s->n_active; i++) {
        tdd_srv_acc_free(s->active[i]->acc);
        free(s->active[i]);
    }
    for (int64_t i = 0; i < s->q_len; i++) free(s->queue[(s->q_head + i) % s->q_cap]);
    for (int64_t i = 0; i < s->d_len; i++) {
        td_t* r = s->done[(s->d_head + i) % s->d_cap].result;
        if (r && !TD_IS_ERR(r)) td_release(r);
    }
    tdd_srv_close(s->dev);
    td_release(s->fact);
    td_release(s->dim);
    free(s->queue);
    free(s->active);
    free(s->done);
    free(s);
}
