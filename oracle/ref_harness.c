/*
 * ref_harness.c -- test infrastructure only.
 *
 * Drives the reference library (oracle/_ref/libteide_ref.so, compiled from
 * the reference's own sources) through its public td.h API on the benchmark
 * workloads, to (a) write golden result tables for the parity tests and
 * (b) time the reference's multithreaded CPU path for the benchmark's
 * cpu_baseline.  Data comes from the shared recipes in datagen.c, so the
 * reference sees the same bytes as the GPU path.
 *
 * The DAG shapes follow the reference's own idioms: a FILTER over a
 * CONST_TABLE executed to a compacted table (exec.c:11706-11737 +
 * td_execute's final compaction, exec.c:12312-12319), JOIN with a filtered
 * left side (exec.c:11840-11857), and GROUP on a table bound to a new graph
 * (exec.c:11759-11838).
 *
 * Usage (all sizes in rows):
 *   ref_harness fg     N SEED OUT             filter flag==0, group region
 *   ref_harness fgsel  N SEED OUT             same, GROUP with g->selection set
 *   ref_harness jg     NFACT NDIM NCAT SEED OUT [JOINOUT]
 *   ref_harness hc     N NVALUES NSYM SEED OUTKEY OUTSYM
 *   ref_harness join   NL NR KEYRANGE JTYPE SEED OUT
 *   ref_harness having N NVALUES NSYM SEED MINCOUNT MAXSUM TOP OUT OUTTOP
 *   ref_harness sort   N NVALUES NSYM SEED TOP OUT OUTTOP
 *   ref_harness save   N NVALUES NSYM SEED DIR SYMFILE   splayed table + symbols
 *   ref_harness bench  WORKLOAD ARGS... REPS  -> one JSON line (median ms)
 * Output files use the TDT1 column format read by tests/tdtio.py.
 */
#define _POSIX_C_SOURCE 200809L
#include <teide/td.h>
#include "datagen.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <errno.h>
#include <sys/stat.h>

static unsigned g_ref_threads;           /* pool workers + the caller */

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

static void die(const char* what) {
    fprintf(stderr, "ref_harness: %s\n", what);
    exit(2);
}

static td_t* check(td_t* v, const char* what) {
    if (!v || TD_IS_ERR(v)) {
        fprintf(stderr, "ref_harness: %s failed (err %d)\n", what,
                v ? (int)TD_ERR_CODE(v) : -1);
        exit(3);
    }
    return v;
}

static td_t* vec_of(int8_t type, int64_t n) {
    td_t* v = check(td_vec_new(type, n), "td_vec_new");
    v->len = n;
    return v;
}

static td_t* add(td_t* tbl, const char* name, td_t* col) {
    tbl = check(td_table_add_col(tbl, td_sym_intern(name, strlen(name)), col), "add_col");
    td_release(col);
    return tbl;
}

/* ------------------------------------------------------------- TDT1 out -- */

static void write_table(const char* path, td_t* tbl) {
    FILE* f = fopen(path, "wb");
    if (!f) die("cannot open output");
    uint32_t nc = (uint32_t)td_table_ncols(tbl);
    fwrite("TDT1", 1, 4, f);
    fwrite(&nc, 4, 1, f);
    for (uint32_t c = 0; c < nc; c++) {
        td_t* col = td_table_get_col_idx(tbl, c);
        td_t* nm = td_sym_str(td_table_col_name(tbl, c));
        uint32_t nl = nm ? (uint32_t)td_str_len(nm) : 0;
        fwrite(&nl, 4, 1, f);
        if (nl) fwrite(td_str_ptr(nm), 1, nl, f);
        int8_t t = col->type;
        uint8_t a = col->attrs;
        uint16_t pad = 0;
        int64_t len = col->len;
        fwrite(&t, 1, 1, f);
        fwrite(&a, 1, 1, f);
        fwrite(&pad, 2, 1, f);
        fwrite(&len, 8, 1, f);
        size_t esz = td_sym_elem_size(t, a);
        if (len) fwrite(td_data(col), esz, (size_t)len, f);
    }
    fclose(f);
}

/* ----------------------------------------------------------- workloads -- */

static int g_threads = 8;

static td_t* make_fg(int64_t n, uint64_t seed) {
    td_t* flag = vec_of(TD_I64, n);
    td_t* region = vec_of(TD_I64, n);
    td_t* amount = vec_of(TD_F64, n);
    tdg_recipe_fg(n, seed, g_threads, (int64_t*)td_data(flag),
                  (int64_t*)td_data(region), (double*)td_data(amount));
    td_t* t = check(td_table_new(3), "table");
    t = add(t, "flag", flag);
    t = add(t, "region", region);
    t = add(t, "amount", amount);
    return t;
}

/* GROUP(region; SUM amount, COUNT amount) over tbl, optionally with sel. */
static td_t* run_group_fg(td_t* tbl, td_t* sel) {
    td_graph_t* g = td_graph_new(tbl);
    if (sel) { td_retain(sel); g->selection = sel; }
    td_op_t* keys[1] = { td_scan(g, "region") };
    td_op_t* ins[2] = { td_scan(g, "amount"), td_scan(g, "amount") };
    uint16_t ops[2] = { OP_SUM, OP_COUNT };
    td_op_t* root = td_group(g, keys, 1, ops, ins, 2);
    root = td_optimize(g, root);
    td_t* out = check(td_execute(g, root), "group");
    td_graph_free(g);
    return out;
}

static td_t* run_filter_table(td_t* tbl, const char* col, int op, int64_t c) {
    td_graph_t* g = td_graph_new(tbl);
    td_op_t* t = td_const_table(g, tbl);
    td_op_t* s = td_scan(g, col);
    td_op_t* k = td_const_i64(g, c);
    td_op_t* p = op == OP_EQ ? td_eq(g, s, k) : td_gt(g, s, k);
    td_op_t* root = td_filter(g, t, p);
    root = td_optimize(g, root);
    td_t* out = check(td_execute(g, root), "filter");
    td_graph_free(g);
    return out;
}

static td_t* wl_fg(int64_t n, uint64_t seed) {
    td_t* t = make_fg(n, seed);
    td_t* filtered = run_filter_table(t, "flag", OP_EQ, 0);
    td_t* out = run_group_fg(filtered, NULL);
    td_release(filtered);
    td_release(t);
    return out;
}

static td_t* wl_fgsel(int64_t n, uint64_t seed) {
    td_t* t = make_fg(n, seed);
    td_t* flag = td_table_get_col(t, td_sym_intern("flag", 4));
    td_t* pred = vec_of(TD_BOOL, n);
    for (int64_t i = 0; i < n; i++)
        ((uint8_t*)td_data(pred))[i] = ((int64_t*)td_data(flag))[i] == 0;
    td_t* sel = check(td_sel_from_pred(pred), "sel");
    td_release(pred);
    td_t* out = run_group_fg(t, sel);
    td_release(sel);
    td_release(t);
    return out;
}

static void make_jg(int64_t nf, int64_t nd, int64_t ncat, uint64_t seed,
                    td_t** fact_out, td_t** dim_out) {
    td_t* pk = vec_of(TD_I64, nd);
    td_t* cat = vec_of(TD_I32, nd);
    tdg_recipe_dim(nd, ncat, seed, 200, g_threads, (int64_t*)td_data(pk), (int32_t*)td_data(cat));
    td_t* fk = vec_of(TD_I64, nf);
    td_t* qty = vec_of(TD_I64, nf);
    td_t* price = vec_of(TD_F64, nf);
    tdg_recipe_fact(nf, nd, seed, 200, g_threads, (int64_t*)td_data(fk),
                    (int64_t*)td_data(qty), (double*)td_data(price));
    td_t* dim = check(td_table_new(2), "dim");
    dim = add(dim, "pk", pk);
    dim = add(dim, "cat", cat);
    td_t* fact = check(td_table_new(3), "fact");
    fact = add(fact, "fk", fk);
    fact = add(fact, "qty", qty);
    fact = add(fact, "price", price);
    *fact_out = fact;
    *dim_out = dim;
}

/* JOIN(FILTER(CONST_TABLE fact, GT(SCAN qty, 50)), [SCAN fk],
 *      CONST_TABLE dim, [SCAN pk], inner) */
static td_t* run_jg_join(td_t* fact, td_t* dim) {
    td_graph_t* g = td_graph_new(fact);
    td_op_t* ft = td_const_table(g, fact);
    td_op_t* dt = td_const_table(g, dim);
    td_op_t* pred = td_gt(g, td_scan(g, "qty"), td_const_i64(g, 50));
    td_op_t* flt = td_filter(g, ft, pred);
    td_op_t* lk[1] = { td_scan(g, "fk") };
    td_op_t* rk[1] = { td_scan(g, "pk") };
    td_op_t* root = td_join(g, flt, lk, dt, rk, 1, 0);
    root = td_optimize(g, root);
    td_t* out = check(td_execute(g, root), "join");
    td_graph_free(g);
    return out;
}

static td_t* run_jg_group(td_t* joined) {
    td_graph_t* g = td_graph_new(joined);
    td_op_t* keys[1] = { td_scan(g, "cat") };
    td_op_t* ins[3] = { td_scan(g, "price"), td_scan(g, "qty"), td_scan(g, "price") };
    uint16_t ops[3] = { OP_SUM, OP_SUM, OP_COUNT };
    td_op_t* root = td_group(g, keys, 1, ops, ins, 3);
    root = td_optimize(g, root);
    td_t* out = check(td_execute(g, root), "group");
    td_graph_free(g);
    return out;
}

static void make_hc(int64_t n, int64_t nv, uint32_t nsym, uint64_t seed, td_t** out) {
    td_t* key = vec_of(TD_I64, n);
    td_t* sym = check(td_sym_vec_new(TD_SYM_W32, n), "sym");
    sym->len = n;
    td_t* vi = vec_of(TD_I64, n);
    td_t* vf = vec_of(TD_F64, n);
    tdg_recipe_hc(n, nv, 1.1, nsym, seed, g_threads, (int64_t*)td_data(key),
                  (uint32_t*)td_data(sym), (int64_t*)td_data(vi), (double*)td_data(vf));
    td_t* t = check(td_table_new(4), "hc");
    t = add(t, "key", key);
    t = add(t, "sym", sym);
    t = add(t, "v_i64", vi);
    t = add(t, "v_f64", vf);
    *out = t;
}

/* GROUP(cat) of expression inputs over the join output: products of fact
 * columns, an affine AVG, a fact x dim product and a MAX of an integer
 * expression (the group runs on the joined table, exec.c:5700-5712). */
static td_t* run_jg_expr_group(td_t* joined) {
    td_graph_t* g = td_graph_new(joined);
    td_op_t* price = td_scan(g, "price");
    td_op_t* qty = td_scan(g, "qty");
    td_op_t* keys[1] = { td_scan(g, "cat") };
    td_op_t* ins[5] = {
        td_mul(g, price, td_cast(g, qty, TD_F64)),
        td_mul(g, price, td_const_f64(g, 0.9)),
        td_add(g, qty, td_const_i64(g, 1)),
        td_mul(g, price, td_cast(g, td_scan(g, "cat"), TD_F64)),
        td_sub(g, td_mul(g, qty, td_const_i64(g, 2)), td_const_i64(g, 1)),
    };
    uint16_t ops[5] = { OP_SUM, OP_SUM, OP_AVG, OP_SUM, OP_MAX };
    td_op_t* root = td_group(g, keys, 1, ops, ins, 5);
    td_t* out = check(td_execute(g, root), "jg expr group");
    td_graph_free(g);
    return out;
}

/* HAVING on the group output (exec.c:11674-11704): GROUP(sym; SUM v_f64,
 * COUNT v_i64, MAX v_i64) filtered to v_i64_count >= MINC and
 * v_f64_sum < MAXS; with TOP > 0, HEAD(SORT(that, v_f64_sum desc), TOP). */
static td_t* run_having(td_t* t, int64_t minc, double maxs, int64_t top) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* keys[1] = { td_scan(g, "sym") };
    td_op_t* ins[3] = { td_scan(g, "v_f64"), td_scan(g, "v_i64"), td_scan(g, "v_i64") };
    uint16_t ops[3] = { OP_SUM, OP_COUNT, OP_MAX };
    td_op_t* grp = td_group(g, keys, 1, ops, ins, 3);
    td_op_t* pred = td_and(g, td_ge(g, td_scan(g, "v_i64_count"), td_const_i64(g, minc)),
                           td_lt(g, td_scan(g, "v_f64_sum"), td_const_f64(g, maxs)));
    td_op_t* root = td_filter(g, grp, pred);
    if (top > 0) {
        td_op_t* sk[1] = { td_scan(g, "v_f64_sum") };
        uint8_t desc[1] = { 1 }, nf[1] = { 1 };
        root = td_head(g, td_sort_op(g, root, sk, desc, nf, 1), top);
    }
    td_t* out = check(td_execute(g, root), "having");
    td_graph_free(g);
    return out;
}

/* SORT of the table by (key asc, v_i64 desc); with TOP > 0, HEAD(SORT(t,
 * v_f64 desc), TOP) -- the reference's heap top-N (exec.c:3455-3460). */
static td_t* run_sort(td_t* t, int64_t top) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* tn = td_const_table(g, t);
    td_op_t* root;
    if (top > 0) {
        td_op_t* sk[1] = { td_scan(g, "v_f64") };
        uint8_t desc[1] = { 1 }, nf[1] = { 1 };
        root = td_head(g, td_sort_op(g, tn, sk, desc, nf, 1), top);
    } else {
        td_op_t* sk[2] = { td_scan(g, "key"), td_scan(g, "v_i64") };
        uint8_t desc[2] = { 0, 1 }, nf[2] = { 0, 1 };
        root = td_sort_op(g, tn, sk, desc, nf, 2);
    }
    td_t* out = check(td_execute(g, root), "sort");
    td_graph_free(g);
    return out;
}

/* The high-cardinality recipe's table saved by the reference as a splayed
 * directory (store/splay.c:46-101) with its symbol file. */
static void run_save(td_t* t, const char* dir, const char* sym_path) {
    td_err_t e = td_splay_save(t, dir, sym_path);
    if (e != TD_OK) {
        fprintf(stderr, "ref_harness: td_splay_save failed (err %d)\n", (int)e);
        exit(3);
    }
}

/* GROUP(k; SUM/MIN/MAX v_i64, SUM/MIN/MAX v_f64, COUNT v_i64) */
static td_t* run_hc_group(td_t* t, const char* k) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* keys[1] = { td_scan(g, k) };
    td_op_t* vi = td_scan(g, "v_i64");
    td_op_t* vf = td_scan(g, "v_f64");
    td_op_t* ins[7] = { vi, vi, vi, vf, vf, vf, vi };
    uint16_t ops[7] = { OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_MIN, OP_MAX, OP_COUNT };
    td_op_t* root = td_group(g, keys, 1, ops, ins, 7);
    root = td_optimize(g, root);
    td_t* out = check(td_execute(g, root), "hc group");
    td_graph_free(g);
    return out;
}

/* Element-wise expressions (exec.c:760-1260).  MODE "grp": the table
 * filtered by an arithmetic predicate, then GROUP(sym) of expression inputs;
 * MODE "vec": a table of expression vectors over every row. */
static td_t* run_expr_grp(td_t* t) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* vi = td_scan(g, "v_i64");
    td_op_t* vf = td_scan(g, "v_f64");
    td_op_t* p1 = td_gt(g, td_add(g, td_mod(g, vi, td_const_i64(g, 7)), td_scan(g, "sym")),
                        td_const_i64(g, 1000));
    td_op_t* p2 = td_lt(g, td_sub(g, td_mul(g, vf, td_const_f64(g, 2.5)), td_const_f64(g, 1.0)),
                        td_add(g, td_abs(g, vf), td_const_f64(g, 1.0)));
    td_op_t* root = td_filter(g, td_const_table(g, t), td_and(g, p1, p2));
    root = td_optimize(g, root);
    td_t* f = check(td_execute(g, root), "expr filter");
    td_graph_free(g);

    g = td_graph_new(f);
    vi = td_scan(g, "v_i64");
    vf = td_scan(g, "v_f64");
    td_op_t* key = td_scan(g, "key");
    td_op_t* keys[1] = { td_scan(g, "sym") };
    td_op_t* ins[6] = {
        td_sub(g, td_mul(g, vi, td_const_i64(g, 3)), key),
        td_mul(g, vf, td_cast(g, td_mod(g, vi, td_const_i64(g, 100)), TD_F64)),
        td_mod(g, key, td_const_i64(g, 1000003)),
        td_div(g, vf, td_add(g, td_abs(g, td_mod(g, key, td_const_i64(g, 1000))),
                             td_const_i64(g, 1))),
        td_sqrt_op(g, td_abs(g, vf)),
        vi,
    };
    uint16_t ops[6] = { OP_SUM, OP_SUM, OP_MAX, OP_MIN, OP_AVG, OP_COUNT };
    root = td_group(g, keys, 1, ops, ins, 6);
    root = td_optimize(g, root);
    td_t* out = check(td_execute(g, root), "expr group");
    td_graph_free(g);
    td_release(f);
    return out;
}

/* GROUP(sym) of affine aggregate inputs, SUM / AVG(col +- c), which the
 * reference aggregates as the column plus a bias (exec.c:614-694), beside
 * one that is not affine (a fractional bias on an integer column). */
static td_t* run_expr_aff(td_t* t) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* vi = td_scan(g, "v_i64");
    td_op_t* vf = td_scan(g, "v_f64");
    td_op_t* keys[1] = { td_scan(g, "sym") };
    td_op_t* ins[7] = {
        td_add(g, vf, td_const_f64(g, 2.5)),
        td_sub(g, vi, td_const_i64(g, 7)),
        td_add(g, vi, td_const_f64(g, 3.0)),
        td_add(g, td_const_i64(g, 10), vf),
        td_sub(g, vf, td_const_f64(g, 1.0)),
        td_add(g, vi, td_const_f64(g, 0.5)),
        vi,
    };
    uint16_t ops[7] = { OP_SUM, OP_AVG, OP_SUM, OP_SUM, OP_AVG, OP_SUM, OP_COUNT };
    td_op_t* root = td_group(g, keys, 1, ops, ins, 7);
    td_t* out = check(td_execute(g, root), "expr aff");
    td_graph_free(g);
    return out;
}

/* GROUP with no keys (the scalar aggregate, exec.c:5766-5940): plain,
 * affine and integer-linear inputs (an integral f64 constant keeps the
 * linear sum in i64), and one that is neither. */
static td_t* run_expr_scalar(td_t* t) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* vi = td_scan(g, "v_i64");
    td_op_t* vf = td_scan(g, "v_f64");
    td_op_t* key = td_scan(g, "key");
    td_op_t* ins[7] = {
        vi, vi, vf,
        td_add(g, vf, td_const_f64(g, 1.5)),
        td_add(g, td_mul(g, vi, td_const_f64(g, 2.0)), key),
        td_sub(g, vi, td_mul(g, td_const_i64(g, 3), key)),
        td_mul(g, vi, td_const_f64(g, 0.5)),
    };
    uint16_t ops[7] = { OP_SUM, OP_COUNT, OP_AVG, OP_SUM, OP_SUM, OP_AVG, OP_SUM };
    td_op_t* root = td_group(g, NULL, 0, ops, ins, 7);
    td_t* out = check(td_execute(g, root), "expr scalar");
    td_graph_free(g);
    return out;
}

static td_op_t* expr_vec_op(td_graph_t* g, int i) {
    td_op_t* vi = td_scan(g, "v_i64");
    td_op_t* vf = td_scan(g, "v_f64");
    td_op_t* key = td_scan(g, "key");
    td_op_t* sym = td_scan(g, "sym");
    switch (i) {
    case 0:   /* f64 transcendental mix */
        return td_add(g, td_sqrt_op(g, td_abs(g, vf)),
                      td_mul(g, td_log_op(g, td_add(g, td_abs(g, vf), td_const_f64(g, 1.0))),
                             td_exp_op(g, td_div(g, vf, td_const_f64(g, 4.0)))));
    case 1:   /* wrapping i64 arithmetic */
        return td_sub(g, td_mod(g, td_add(g, td_mul(g, key, td_const_i64(g, 31)), vi),
                                td_const_i64(g, 1000003)),
                      td_max2(g, vi, td_min2(g, key, td_const_i64(g, 0))));
    case 2:   /* f64 vs f64 comparison */
        return td_gt(g, vf, td_div(g, vi, td_const_f64(g, 1e9)));
    case 3:   /* mixed f64 / i64 comparison */
        return td_le(g, td_mul(g, vf, td_const_f64(g, 1e9)), vi);
    case 4:   /* f64 -> i64 cast, clamped */
        return td_cast(g, td_mul(g, vf, td_const_f64(g, 3e18)), TD_I64);
    case 5:
        return td_neg(g, td_abs(g, td_min2(g, vi, key)));
    case 6:
        return td_add(g, td_floor_op(g, td_mul(g, vf, td_const_f64(g, 10.0))),
                      td_ceil_op(g, vf));
    default:  /* i64 MOD by zero gives 0 */
        return td_mod(g, vi, td_mod(g, sym, td_const_i64(g, 3)));
    }
}

/* The pricing expression of a line-item scan, price * (1 - disc) * (1 + tax),
 * over three f64 columns: price in [900, 105000), disc in [0, 0.1), tax in
 * [0, 0.08). */
static td_t* make_charge(int64_t n, uint64_t seed) {
    td_t* price = vec_of(TD_F64, n);
    td_t* disc = vec_of(TD_F64, n);
    td_t* tax = vec_of(TD_F64, n);
    tdg_f64_uniform((double*)td_data(price), n, 900.0, 105000.0, seed, 701, g_threads);
    tdg_f64_uniform((double*)td_data(disc), n, 0.0, 0.1, seed, 702, g_threads);
    tdg_f64_uniform((double*)td_data(tax), n, 0.0, 0.08, seed, 703, g_threads);
    td_t* t = check(td_table_new(3), "charge");
    t = add(t, "price", price);
    t = add(t, "disc", disc);
    t = add(t, "tax", tax);
    return t;
}

static td_t* run_charge(td_t* t) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* e = td_mul(g, td_mul(g, td_scan(g, "price"),
                                  td_sub(g, td_const_f64(g, 1.0), td_scan(g, "disc"))),
                        td_add(g, td_const_f64(g, 1.0), td_scan(g, "tax")));
    td_t* out = check(td_execute(g, td_optimize(g, e)), "charge");
    td_graph_free(g);
    return out;
}

static td_t* run_expr_vec(td_t* t) {
    td_t* out = check(td_table_new(8), "expr vec");
    for (int i = 0; i < 8; i++) {
        td_graph_t* g = td_graph_new(t);
        td_op_t* root = td_optimize(g, expr_vec_op(g, i));
        td_t* v = check(td_execute(g, root), "expr vec");
        td_graph_free(g);
        char name[8];
        snprintf(name, sizeof(name), "e%d", i);
        out = add(out, name, v);
    }
    return out;
}

/* A date-partitioned database (store/part.c layout) of the high-cardinality
 * recipe's rows: NP partitions 2024.01.01, 2024.01.02, ... of table `t`,
 * partition p holding the next (p + 1) / (NP (NP + 1) / 2) share of the
 * rows, and the symbol file `sym`. */
static void run_partsave(td_t* t, int np, const char* root) {
    int64_t n = td_table_nrows(t), at = 0, tri = (int64_t)np * (np + 1) / 2;
    if (mkdir(root, 0755) != 0 && errno != EEXIST) die("partsave: mkdir");
    char dir[1024];
    for (int p = 0; p < np; p++) {
        int64_t len = p == np - 1 ? n - at : n * (p + 1) / tri;
        snprintf(dir, sizeof(dir), "%s/2024.01.%02d", root, p + 1);
        if (mkdir(dir, 0755) != 0 && errno != EEXIST) die("partsave: mkdir");
        td_t* part = check(td_table_new(4), "part");
        for (int64_t c = 0; c < td_table_ncols(t); c++) {
            td_t* col = td_table_get_col_idx(t, c);
            uint8_t esz = td_sym_elem_size(col->type, col->attrs);
            td_t* v = col->type == TD_SYM ? check(td_sym_vec_new(col->attrs & TD_SYM_W_MASK, len), "seg")
                                          : vec_of(col->type, len);
            v->len = len;
            memcpy(td_data(v), (char*)td_data(col) + at * esz, (size_t)len * esz);
            part = check(td_table_add_col(part, td_table_col_name(t, c), v), "seg");
            td_release(v);
        }
        snprintf(dir, sizeof(dir), "%s/2024.01.%02d/t", root, p + 1);
        if (td_splay_save(part, dir, NULL) != TD_OK) die("partsave: splay");
        td_release(part);
        at += len;
    }
    snprintf(dir, sizeof(dir), "%s/sym", root);
    if (td_sym_save(dir) != TD_OK) die("partsave: sym");
}

/* GROUP over td_read_parted of that database: by the partition key (date),
 * by a parted column (sym), and by both. */
static td_t* run_partgroup(td_t* t, int which) {
    td_graph_t* g = td_graph_new(t);
    td_op_t* vi = td_scan(g, "v_i64");
    td_op_t* vf = td_scan(g, "v_f64");
    td_op_t* root;
    if (which == 0) {
        td_op_t* keys[1] = { td_scan(g, "date") };
        td_op_t* ins[5] = { vi, vf, vf, vi, vi };
        uint16_t ops[5] = { OP_SUM, OP_SUM, OP_MIN, OP_MAX, OP_COUNT };
        root = td_group(g, keys, 1, ops, ins, 5);
    } else if (which == 1) {
        td_op_t* keys[1] = { td_scan(g, "sym") };
        td_op_t* ins[3] = { vi, vf, vi };
        uint16_t ops[3] = { OP_SUM, OP_SUM, OP_COUNT };
        root = td_group(g, keys, 1, ops, ins, 3);
    } else {
        td_op_t* keys[2] = { td_scan(g, "date"), td_scan(g, "sym") };
        td_op_t* ins[2] = { vi, vi };
        uint16_t ops[2] = { OP_COUNT, OP_SUM };
        root = td_group(g, keys, 2, ops, ins, 2);
    }
    td_t* out = check(td_execute(g, td_optimize(g, root)), "partgroup");
    td_graph_free(g);
    return out;
}

/* General equi-join on one i64 key: left (k, lv), right (k, rv). */
static td_t* run_join(int64_t nl, int64_t nr, int64_t krange, int jtype, uint64_t seed) {
    td_t* lk = vec_of(TD_I64, nl);
    td_t* lv = vec_of(TD_I64, nl);
    td_t* rk = vec_of(TD_I64, nr);
    td_t* rv = vec_of(TD_F64, nr);
    tdg_i64_uniform((int64_t*)td_data(lk), nl, 0, krange - 1, seed, 601, g_threads);
    tdg_i64_uniform((int64_t*)td_data(lv), nl, -1000000, 1000000, seed, 602, g_threads);
    tdg_i64_uniform((int64_t*)td_data(rk), nr, 0, krange - 1, seed, 603, g_threads);
    tdg_f64_uniform((double*)td_data(rv), nr, -1.0, 1.0, seed, 604, g_threads);
    td_t* l = check(td_table_new(2), "l");
    l = add(l, "k", lk);
    l = add(l, "lv", lv);
    td_t* r = check(td_table_new(2), "r");
    r = add(r, "k", rk);
    r = add(r, "rv", rv);
    td_graph_t* g = td_graph_new(l);
    td_op_t* lt = td_const_table(g, l);
    td_op_t* rt = td_const_table(g, r);
    td_op_t* key[1] = { td_scan(g, "k") };
    td_op_t* root = td_join(g, lt, key, rt, key, 1, (uint8_t)jtype);
    root = td_optimize(g, root);
    td_t* out = check(td_execute(g, root), "join");
    td_graph_free(g);
    td_release(l);
    td_release(r);
    return out;
}

/* ---------------------------------------------------------------- bench -- */

static int cmp_d(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}

static td_t* run_partgroup(td_t* t, int which);

static void bench(int argc, char** argv) {
    if (argc < 4) die("bench: WORKLOAD ARGS... REPS");
    const char* wl = argv[2];
    int reps = atoi(argv[argc - 1]);
    if (reps < 1) reps = 1;
    if (reps > 64) reps = 64;
    double t[64];
    int64_t rows = 0;
    if (!strcmp(wl, "jg")) {
        int64_t nf = atoll(argv[3]), nd = atoll(argv[4]), ncat = atoll(argv[5]);
        td_t *fact, *dim;
        make_jg(nf, nd, ncat, 42, &fact, &dim);
        rows = nf;
        for (int r = -1; r < reps; r++) {             /* r = -1 is the warm-up */
            double t0 = now_ms();
            td_t* j = run_jg_join(fact, dim);
            td_t* o = run_jg_group(j);
            double t1 = now_ms();
            td_release(j);
            td_release(o);
            if (r >= 0) t[r] = t1 - t0;
        }
        td_release(fact);
        td_release(dim);
    } else if (!strcmp(wl, "fg")) {
        int64_t n = atoll(argv[3]);
        td_t* tb = make_fg(n, 42);
        rows = n;
        for (int r = -1; r < reps; r++) {
            double t0 = now_ms();
            td_t* f = run_filter_table(tb, "flag", OP_EQ, 0);
            td_t* o = run_group_fg(f, NULL);
            double t1 = now_ms();
            td_release(f);
            td_release(o);
            if (r >= 0) t[r] = t1 - t0;
        }
        td_release(tb);
    } else if (!strcmp(wl, "hc")) {
        int64_t n = atoll(argv[3]), nv = atoll(argv[4]);
        uint32_t nsym = (uint32_t)atoll(argv[5]);
        td_t* tb;
        make_hc(n, nv, nsym, 42, &tb);
        rows = n;
        for (int r = -1; r < reps; r++) {
            double t0 = now_ms();
            td_t* a = run_hc_group(tb, "key");
            td_t* b = run_hc_group(tb, "sym");
            double t1 = now_ms();
            td_release(a);
            td_release(b);
            if (r >= 0) t[r] = t1 - t0;
        }
        td_release(tb);
    } else if (!strcmp(wl, "partgroup")) {
        /* a partitioned database on disk (argv[3]): td_read_parted once,
         * then per repetition GROUP(date) and GROUP(sym) as run_partgroup */
        td_t* tb = check(td_read_parted(argv[3], "t"), "read_parted");
        rows = td_table_nrows(tb);
        for (int r = -1; r < reps; r++) {
            double t0 = now_ms();
            td_t* a = run_partgroup(tb, 0);
            td_t* b = run_partgroup(tb, 1);
            double t1 = now_ms();
            td_release(a);
            td_release(b);
            if (r >= 0) t[r] = t1 - t0;
        }
        td_release(tb);
    } else if (!strcmp(wl, "charge")) {
        int64_t n = atoll(argv[3]);
        td_t* tb = make_charge(n, 42);
        rows = n;
        for (int r = -1; r < reps; r++) {
            double t0 = now_ms();
            td_t* v = run_charge(tb);
            double t1 = now_ms();
            td_release(v);
            if (r >= 0) t[r] = t1 - t0;
        }
        td_release(tb);
    } else {
        die("bench: unknown workload");
    }
    qsort(t, (size_t)reps, sizeof(double), cmp_d);
    double med = t[reps / 2];
    printf("{\"workload\": \"%s\", \"rows\": %lld, \"reps\": %d, \"median_ms\": %.4f, "
           "\"min_ms\": %.4f, \"threads\": %u}\n",
           wl, (long long)rows, reps, med, t[0], g_ref_threads);
}

int main(int argc, char** argv) {
    if (argc < 2) die("usage: ref_harness CMD ...");
    td_heap_init();
    td_sym_init();
    /* TD_REF_THREADS=T (T >= 2): T-1 pool workers + the caller; otherwise
     * the reference's default of nproc-1 workers (pool.c:107-110). */
    const char* tenv = getenv("TD_REF_THREADS");
    int treq = tenv ? atoi(tenv) : 0;
    if (treq >= 2) {
        td_pool_init((uint32_t)(treq - 1));
        g_ref_threads = (unsigned)treq;
    } else {
        td_pool_init(0);
        g_ref_threads = td_thread_count();
    }
    const char* cmd = argv[1];
    if (!strcmp(cmd, "fg") && argc == 5) {
        td_t* o = wl_fg(atoll(argv[2]), strtoull(argv[3], NULL, 10));
        write_table(argv[4], o);
        td_release(o);
    } else if (!strcmp(cmd, "fgsel") && argc == 5) {
        td_t* o = wl_fgsel(atoll(argv[2]), strtoull(argv[3], NULL, 10));
        write_table(argv[4], o);
        td_release(o);
    } else if (!strcmp(cmd, "jg") && (argc == 7 || argc == 8)) {
        td_t *fact, *dim;
        make_jg(atoll(argv[2]), atoll(argv[3]), atoll(argv[4]),
                strtoull(argv[5], NULL, 10), &fact, &dim);
        td_t* j = run_jg_join(fact, dim);
        if (argc == 8) write_table(argv[7], j);
        td_t* o = run_jg_group(j);
        write_table(argv[6], o);
        td_release(o);
        td_release(j);
        td_release(fact);
        td_release(dim);
    } else if (!strcmp(cmd, "jgexpr") && argc == 7) {
        td_t *fact, *dim;
        make_jg(atoll(argv[2]), atoll(argv[3]), atoll(argv[4]), strtoull(argv[5], NULL, 10),
                &fact, &dim);
        td_t* j = run_jg_join(fact, dim);
        td_t* o = run_jg_expr_group(j);
        write_table(argv[6], o);
        td_release(o);
        td_release(j);
        td_release(fact);
        td_release(dim);
    } else if (!strcmp(cmd, "hc") && argc == 8) {
        td_t* t;
        make_hc(atoll(argv[2]), atoll(argv[3]), (uint32_t)atoll(argv[4]),
                strtoull(argv[5], NULL, 10), &t);
        td_t* a = run_hc_group(t, "key");
        write_table(argv[6], a);
        td_t* b = run_hc_group(t, "sym");
        write_table(argv[7], b);
        td_release(a);
        td_release(b);
        td_release(t);
    } else if (!strcmp(cmd, "having") && argc == 11) {
        td_t* t;
        make_hc(atoll(argv[2]), atoll(argv[3]), (uint32_t)atoll(argv[4]),
                strtoull(argv[5], NULL, 10), &t);
        td_t* a = run_having(t, atoll(argv[6]), atof(argv[7]), 0);
        write_table(argv[9], a);
        td_t* b = run_having(t, atoll(argv[6]), atof(argv[7]), atoll(argv[8]));
        write_table(argv[10], b);
        td_release(a);
        td_release(b);
        td_release(t);
    } else if (!strcmp(cmd, "sort") && argc == 9) {
        td_t* t;
        make_hc(atoll(argv[2]), atoll(argv[3]), (uint32_t)atoll(argv[4]),
                strtoull(argv[5], NULL, 10), &t);
        td_t* a = run_sort(t, 0);
        write_table(argv[7], a);
        td_t* b = run_sort(t, atoll(argv[6]));
        write_table(argv[8], b);
        td_release(a);
        td_release(b);
        td_release(t);
    } else if (!strcmp(cmd, "expr") && argc == 8) {
        td_t* t;
        make_hc(atoll(argv[2]), atoll(argv[3]), (uint32_t)atoll(argv[4]),
                strtoull(argv[5], NULL, 10), &t);
        td_t* o = !strcmp(argv[6], "grp") ? run_expr_grp(t)
                : !strcmp(argv[6], "aff") ? run_expr_aff(t)
                : !strcmp(argv[6], "scalar") ? run_expr_scalar(t) : run_expr_vec(t);
        write_table(argv[7], o);
        td_release(o);
        td_release(t);
    } else if (!strcmp(cmd, "partsave") && argc == 8) {
        /* the sym column's ids name real symbols (the reference's parted
         * group-by sizes a bitmap by the symbol count) */
        char nm[32];
        for (int64_t i = 0; i < atoll(argv[4]); i++) {
            int k = snprintf(nm, sizeof(nm), "s%05lld", (long long)i);
            td_sym_intern(nm, (size_t)k);
        }
        td_t* t;
        make_hc(atoll(argv[2]), atoll(argv[3]), (uint32_t)atoll(argv[4]),
                strtoull(argv[5], NULL, 10), &t);
        run_partsave(t, atoi(argv[6]), argv[7]);
        td_release(t);
    } else if (!strcmp(cmd, "partgroup") && argc == 6) {
        td_t* t = check(td_read_parted(argv[2], "t"), "read_parted");
        for (int w = 0; w < 3; w++) {
            td_t* o = run_partgroup(t, w);
            write_table(argv[3 + w], o);
            td_release(o);
        }
        td_release(t);
    } else if (!strcmp(cmd, "save") && argc == 8) {
        td_t* t;
        make_hc(atoll(argv[2]), atoll(argv[3]), (uint32_t)atoll(argv[4]),
                strtoull(argv[5], NULL, 10), &t);
        run_save(t, argv[6], argv[7]);
        td_release(t);
    } else if (!strcmp(cmd, "join") && argc == 8) {
        td_t* o = run_join(atoll(argv[2]), atoll(argv[3]), atoll(argv[4]), atoi(argv[5]),
                           strtoull(argv[6], NULL, 10));
        write_table(argv[7], o);
        td_release(o);
    } else if (!strcmp(cmd, "bench")) {
        bench(argc, argv);
    } else {
        die("bad command or argument count");
    }
    td_pool_destroy();
    return 0;
}
