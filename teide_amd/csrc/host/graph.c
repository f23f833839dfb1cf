/*
 * graph.c -- query DAG builders.
 *
 * A graph owns an array of 32-byte td_op_t nodes plus heap-allocated
 * td_op_ext_t records for nodes that carry a payload (SCAN name, CONST
 * literal, GROUP/JOIN/SORT specs ...).  The contract matches the reference's
 * src/ops/graph.c: node pointers handed out point into g->nodes, node ids are
 * array indices, the ext record's base mirrors its node, the nodes array may
 * move when it grows (every stored input pointer is rebased), SCAN keeps
 * table_id + 1 in pad[0..1] (graph.c:1036-1061), IF keeps the else-node id
 * in literal (graph.c:453-482), and out_type follows the reference's
 * promotion rule (graph.c:403-415).
 */
#include "internal.h"

#include <stdlib.h>

#define INIT_NODES 4096

td_graph_t* td_graph_new(td_t* tbl) {
    td_graph_t* g = calloc(1, sizeof(td_graph_t));
    if (!g) return NULL;
    g->nodes = calloc(INIT_NODES, sizeof(td_op_t));
    if (!g->nodes) { free(g); return NULL; }
    g->node_cap = INIT_NODES;
    g->table = tbl;
    if (tbl) td_retain(tbl);
    return g;
}

void td_graph_free(td_graph_t* g) {
    if (!g) return;
    for (uint32_t i = 0; i < g->ext_count; i++) {
        td_op_ext_t* e = g->ext_nodes[i];
        if (!e) continue;
        uint16_t oc = g->nodes[e->base.id].opcode;
        if (oc == OP_CONST && e->literal) td_release(e->literal);
        free(e);
    }
    free(g->ext_nodes);
    free(g->nodes);
    if (g->table) td_release(g->table);
    for (uint16_t i = 0; i < g->n_tables; i++) if (g->tables[i]) td_release(g->tables[i]);
    free(g->tables);
    if (g->selection) td_release(g->selection);
    free(g);
}

/* ---------------------------------------------------------- node arrays -- */

static td_op_t* rebase(td_op_t* p, ptrdiff_t d) { return p ? (td_op_t*)((char*)p + d) : NULL; }

static void rebase_all(td_graph_t* g, ptrdiff_t d) {
    if (!d) return;
    for (uint32_t i = 0; i < g->node_count; i++) {
        g->nodes[i].inputs[0] = rebase(g->nodes[i].inputs[0], d);
        g->nodes[i].inputs[1] = rebase(g->nodes[i].inputs[1], d);
    }
    for (uint32_t i = 0; i < g->ext_count; i++) {
        td_op_ext_t* e = g->ext_nodes[i];
        if (!e) continue;
        e->base.inputs[0] = rebase(e->base.inputs[0], d);
        e->base.inputs[1] = rebase(e->base.inputs[1], d);
        switch (e->base.opcode) {
        case OP_GROUP:
            for (int k = 0; k < e->n_keys; k++) e->keys[k] = rebase(e->keys[k], d);
            for (int a = 0; a < e->n_aggs; a++) e->agg_ins[a] = rebase(e->agg_ins[a], d);
            break;
        case OP_JOIN:
            for (int k = 0; k < e->join.n_join_keys; k++) {
                e->join.left_keys[k] = rebase(e->join.left_keys[k], d);
                e->join.right_keys[k] = rebase(e->join.right_keys[k], d);
            }
            break;
        case OP_SORT: case OP_SELECT:
            for (int k = 0; k < e->sort.n_cols; k++) e->sort.columns[k] = rebase(e->sort.columns[k], d);
            break;
        case OP_WINDOW_JOIN:
            e->asof.time_key = rebase(e->asof.time_key, d);
            for (int k = 0; k < e->asof.n_eq_keys; k++) e->asof.eq_keys[k] = rebase(e->asof.eq_keys[k], d);
            break;
        case OP_WINDOW:
            for (int k = 0; k < e->window.n_part_keys; k++)
                e->window.part_keys[k] = rebase(e->window.part_keys[k], d);
            for (int k = 0; k < e->window.n_order_keys; k++)
                e->window.order_keys[k] = rebase(e->window.order_keys[k], d);
            for (int k = 0; k < e->window.n_funcs; k++)
                e->window.func_inputs[k] = rebase(e->window.func_inputs[k], d);
            break;
        default: break;
        }
    }
}

static int grow_nodes(td_graph_t* g) {
    if (g->node_count < g->node_cap) return 1;
    if (g->node_cap > UINT32_MAX / 2) return 0;
    uintptr_t old = (uintptr_t)g->nodes;
    td_op_t* n = realloc(g->nodes, (size_t)g->node_cap * 2 * sizeof(td_op_t));
    if (!n) return 0;
    g->nodes = n;
    g->node_cap *= 2;
    rebase_all(g, (ptrdiff_t)((uintptr_t)n - old));
    return 1;
}

static td_op_t* new_node(td_graph_t* g) {
    if (!grow_nodes(g)) return NULL;
    td_op_t* n = &g->nodes[g->node_count];
    memset(n, 0, sizeof(*n));
    n->id = g->node_count++;
    return n;
}

/* A fresh node for the optimizer's rewrites (may move g->nodes). */
td_op_t* td__graph_new_node(td_graph_t* g) { return new_node(g); }

/* An ext record plus `extra` trailing bytes for its arrays; its node slot is
 * reserved now and filled by publish(). */
static td_op_ext_t* new_ext(td_graph_t* g, size_t extra) {
    if (!grow_nodes(g)) return NULL;
    if (g->ext_count == g->ext_cap) {
        uint32_t nc = g->ext_cap ? g->ext_cap * 2 : 16;
        td_op_ext_t** ne = realloc(g->ext_nodes, nc * sizeof(td_op_ext_t*));
        if (!ne) return NULL;
        g->ext_nodes = ne;
        g->ext_cap = nc;
    }
    td_op_ext_t* e = calloc(1, sizeof(td_op_ext_t) + extra);
    if (!e) return NULL;
    e->base.id = g->node_count;
    memset(&g->nodes[g->node_count], 0, sizeof(td_op_t));
    g->nodes[g->node_count].id = g->node_count;
    g->node_count++;
    g->ext_nodes[g->ext_count++] = e;
    return e;
}

static td_op_t* publish(td_graph_t* g, td_op_ext_t* e) {
    g->nodes[e->base.id] = e->base;
    return &g->nodes[e->base.id];
}

#define TRAIL(e) ((char*)((e) + 1))

/* -------------------------------------------------------------- sources -- */

static td_t* table_for_scan(td_graph_t* g, uint16_t stored) {
    if (stored && g->tables && stored - 1 < g->n_tables) return g->tables[stored - 1];
    return g->table;
}

static td_op_t* scan_impl(td_graph_t* g, uint16_t stored, const char* name) {
    td_op_ext_t* e = new_ext(g, 0);
    if (!e) return NULL;
    e->base.opcode = OP_SCAN;
    e->sym = td_sym_intern(name, strlen(name));
    memcpy(e->base.pad, &stored, sizeof(stored));
    td_t* tbl = table_for_scan(g, stored);
    td_t* col = tbl ? td_table_get_col(tbl, e->sym) : NULL;
    if (col) {
        e->base.out_type = col->type;
        e->base.est_rows = (uint32_t)col->len;
    }
    return publish(g, e);
}

td_op_t* td_scan(td_graph_t* g, const char* col_name) { return scan_impl(g, 0, col_name); }

td_op_t* td_scan_table(td_graph_t* g, uint16_t table_id, const char* col_name) {
    if (table_id >= g->n_tables || !g->tables[table_id]) return NULL;
    return scan_impl(g, (uint16_t)(table_id + 1), col_name);
}

uint16_t td_graph_add_table(td_graph_t* g, td_t* table) {
    td_t** nt = realloc(g->tables, ((size_t)g->n_tables + 1) * sizeof(td_t*));
    if (!nt) return UINT16_MAX;
    g->tables = nt;
    g->tables[g->n_tables] = table;
    td_retain(table);
    return g->n_tables++;
}

static td_op_t* const_impl(td_graph_t* g, td_t* lit, int8_t out_type, uint32_t est) {
    td_op_ext_t* e = new_ext(g, 0);
    if (!e) { if (lit && !TD_IS_ERR(lit)) td_release(lit); return NULL; }
    e->base.opcode = OP_CONST;
    e->base.out_type = out_type;
    e->base.est_rows = est;
    e->literal = (lit && !TD_IS_ERR(lit)) ? lit : NULL;
    return publish(g, e);
}

td_op_t* td_const_f64(td_graph_t* g, double v)  { return const_impl(g, td_f64(v), TD_F64, 0); }
td_op_t* td_const_i64(td_graph_t* g, int64_t v) { return const_impl(g, td_i64(v), TD_I64, 0); }
td_op_t* td_const_bool(td_graph_t* g, bool v)   { return const_impl(g, td_bool(v), TD_BOOL, 0); }
td_op_t* td_const_str(td_graph_t* g, const char* s) {
    return const_impl(g, td_str(s, strlen(s)), TD_SYM, 0);
}
td_op_t* td_const_vec(td_graph_t* g, td_t* v) {
    td_retain(v);
    return const_impl(g, v, v->type, (uint32_t)v->len);
}
td_op_t* td_const_table(td_graph_t* g, td_t* t) {
    td_retain(t);
    return const_impl(g, t, TD_TABLE, 0);
}

/* ------------------------------------------------------- element-wise ops -- */

static td_op_t* unary(td_graph_t* g, uint16_t op, td_op_t* a, int8_t t) {
    uint32_t aid = a->id, est = a->est_rows;
    td_op_t* n = new_node(g);
    if (!n) return NULL;
    n->opcode = op;
    n->arity = 1;
    n->inputs[0] = &g->nodes[aid];
    n->out_type = t;
    n->est_rows = est;
    return n;
}

static td_op_t* binary(td_graph_t* g, uint16_t op, td_op_t* a, td_op_t* b, int8_t t) {
    uint32_t aid = a->id, bid = b->id;
    uint32_t est = a->est_rows > b->est_rows ? a->est_rows : b->est_rows;
    td_op_t* n = new_node(g);
    if (!n) return NULL;
    n->opcode = op;
    n->arity = 2;
    n->inputs[0] = &g->nodes[aid];
    n->inputs[1] = &g->nodes[bid];
    n->out_type = t;
    n->est_rows = est;
    return n;
}

/* BOOL < U8 < I16 < I32 < I64 < F64; SYM and TIMESTAMP count as I64, DATE
 * and TIME as I32. */
static int8_t promote(int8_t a, int8_t b) {
#define ANY(x) (a == (x) || b == (x))
    if (ANY(TD_F64)) return TD_F64;
    if (ANY(TD_I64) || ANY(TD_SYM) || ANY(TD_TIMESTAMP)) return TD_I64;
    if (ANY(TD_I32) || ANY(TD_DATE) || ANY(TD_TIME)) return TD_I32;
    if (ANY(TD_I16)) return TD_I16;
    if (ANY(TD_U8)) return TD_U8;
    return TD_BOOL;
#undef ANY
}

td_op_t* td_neg(td_graph_t* g, td_op_t* a)      { return unary(g, OP_NEG, a, a->out_type); }
td_op_t* td_abs(td_graph_t* g, td_op_t* a)      { return unary(g, OP_ABS, a, a->out_type); }
td_op_t* td_not(td_graph_t* g, td_op_t* a)      { return unary(g, OP_NOT, a, TD_BOOL); }
td_op_t* td_sqrt_op(td_graph_t* g, td_op_t* a)  { return unary(g, OP_SQRT, a, TD_F64); }
td_op_t* td_log_op(td_graph_t* g, td_op_t* a)   { return unary(g, OP_LOG, a, TD_F64); }
td_op_t* td_exp_op(td_graph_t* g, td_op_t* a)   { return unary(g, OP_EXP, a, TD_F64); }
td_op_t* td_ceil_op(td_graph_t* g, td_op_t* a)  { return unary(g, OP_CEIL, a, a->out_type); }
td_op_t* td_floor_op(td_graph_t* g, td_op_t* a) { return unary(g, OP_FLOOR, a, a->out_type); }
td_op_t* td_isnull(td_graph_t* g, td_op_t* a)   { return unary(g, OP_ISNULL, a, TD_BOOL); }
td_op_t* td_cast(td_graph_t* g, td_op_t* a, int8_t t) { return unary(g, OP_CAST, a, t); }

#define BIN_P(fn, op) \
    td_op_t* fn(td_graph_t* g, td_op_t* a, td_op_t* b) { \
        return binary(g, op, a, b, promote(a->out_type, b->out_type)); }
#define BIN_T(fn, op, t) \
    td_op_t* fn(td_graph_t* g, td_op_t* a, td_op_t* b) { return binary(g, op, a, b, t); }
BIN_P(td_add, OP_ADD)
BIN_P(td_sub, OP_SUB)
BIN_P(td_mul, OP_MUL)
BIN_T(td_div, OP_DIV, TD_F64)
BIN_P(td_mod, OP_MOD)
BIN_T(td_eq, OP_EQ, TD_BOOL)
BIN_T(td_ne, OP_NE, TD_BOOL)
BIN_T(td_lt, OP_LT, TD_BOOL)
BIN_T(td_le, OP_LE, TD_BOOL)
BIN_T(td_gt, OP_GT, TD_BOOL)
BIN_T(td_ge, OP_GE, TD_BOOL)
BIN_T(td_and, OP_AND, TD_BOOL)
BIN_T(td_or, OP_OR, TD_BOOL)
BIN_P(td_min2, OP_MIN2)
BIN_P(td_max2, OP_MAX2)
BIN_T(td_like, OP_LIKE, TD_BOOL)
BIN_T(td_ilike, OP_ILIKE, TD_BOOL)

td_op_t* td_if(td_graph_t* g, td_op_t* c, td_op_t* t, td_op_t* f) {
    uint32_t cid = c->id, tid = t->id, fid = f->id, est = c->est_rows;
    int8_t ot = promote(t->out_type, f->out_type);
    if (t->out_type == TD_SYM || f->out_type == TD_SYM) ot = TD_SYM;
    td_op_ext_t* e = new_ext(g, 0);
    if (!e) return NULL;
    e->base.opcode = OP_IF;
    e->base.arity = 2;
    e->base.inputs[0] = &g->nodes[cid];
    e->base.inputs[1] = &g->nodes[tid];
    e->base.out_type = ot;
    e->base.est_rows = est;
    e->literal = (td_t*)(uintptr_t)fid;
    return publish(g, e);
}

td_op_t* td_upper(td_graph_t* g, td_op_t* a)   { return unary(g, OP_UPPER, a, TD_SYM); }
td_op_t* td_lower(td_graph_t* g, td_op_t* a)   { return unary(g, OP_LOWER, a, TD_SYM); }
td_op_t* td_strlen(td_graph_t* g, td_op_t* a)  { return unary(g, OP_STRLEN, a, TD_I64); }
td_op_t* td_trim_op(td_graph_t* g, td_op_t* a) { return unary(g, OP_TRIM, a, TD_SYM); }

static td_op_t* ternary_ext(td_graph_t* g, uint16_t op, td_op_t* a, td_op_t* b, td_op_t* c, int8_t t) {
    uint32_t aid = a->id, bid = b->id, cid = c->id, est = a->est_rows;
    td_op_ext_t* e = new_ext(g, 0);
    if (!e) return NULL;
    e->base.opcode = op;
    e->base.arity = 2;
    e->base.inputs[0] = &g->nodes[aid];
    e->base.inputs[1] = &g->nodes[bid];
    e->base.out_type = t;
    e->base.est_rows = est;
    e->literal = (td_t*)(uintptr_t)cid;
    return publish(g, e);
}
td_op_t* td_substr(td_graph_t* g, td_op_t* s, td_op_t* st, td_op_t* l) {
    return ternary_ext(g, OP_SUBSTR, s, st, l, TD_SYM);
}
td_op_t* td_replace(td_graph_t* g, td_op_t* s, td_op_t* f, td_op_t* t) {
    return ternary_ext(g, OP_REPLACE, s, f, t, TD_SYM);
}

td_op_t* td_concat(td_graph_t* g, td_op_t** args, int n) {
    if (!args || n < 2 || n > 256) return NULL;
    uint32_t ids[256];
    for (int i = 0; i < n; i++) ids[i] = args[i]->id;
    td_op_ext_t* e = new_ext(g, (size_t)(n - 2) * sizeof(uint32_t));
    if (!e) return NULL;
    e->base.opcode = OP_CONCAT;
    e->base.arity = 2;
    e->base.inputs[0] = &g->nodes[ids[0]];
    e->base.inputs[1] = &g->nodes[ids[1]];
    e->base.out_type = TD_SYM;
    e->base.est_rows = g->nodes[ids[0]].est_rows;
    e->sym = n;
    for (int i = 2; i < n; i++) ((uint32_t*)TRAIL(e))[i - 2] = ids[i];
    return publish(g, e);
}

static td_op_t* unary_ext(td_graph_t* g, uint16_t op, td_op_t* a, int8_t t, int64_t sym) {
    uint32_t aid = a->id, est = a->est_rows;
    td_op_ext_t* e = new_ext(g, 0);
    if (!e) return NULL;
    e->base.opcode = op;
    e->base.arity = 1;
    e->base.inputs[0] = &g->nodes[aid];
    e->base.out_type = t;
    e->base.est_rows = est;
    e->sym = sym;
    return publish(g, e);
}
td_op_t* td_extract(td_graph_t* g, td_op_t* c, int64_t f)    { return unary_ext(g, OP_EXTRACT, c, TD_I64, f); }
td_op_t* td_date_trunc(td_graph_t* g, td_op_t* c, int64_t f) { return unary_ext(g, OP_DATE_TRUNC, c, TD_TIMESTAMP, f); }

/* ----------------------------------------------------------- reductions -- */

#define RED(fn, op, t) td_op_t* fn(td_graph_t* g, td_op_t* a) { return unary(g, op, a, t); }
td_op_t* td_sum(td_graph_t* g, td_op_t* a)  { return unary(g, OP_SUM, a, a->out_type == TD_F64 ? TD_F64 : TD_I64); }
td_op_t* td_prod(td_graph_t* g, td_op_t* a) { return unary(g, OP_PROD, a, a->out_type == TD_F64 ? TD_F64 : TD_I64); }
td_op_t* td_min_op(td_graph_t* g, td_op_t* a) { return unary(g, OP_MIN, a, a->out_type); }
td_op_t* td_max_op(td_graph_t* g, td_op_t* a) { return unary(g, OP_MAX, a, a->out_type); }
td_op_t* td_first(td_graph_t* g, td_op_t* a)  { return unary(g, OP_FIRST, a, a->out_type); }
td_op_t* td_last(td_graph_t* g, td_op_t* a)   { return unary(g, OP_LAST, a, a->out_type); }
RED(td_count, OP_COUNT, TD_I64)
RED(td_avg, OP_AVG, TD_F64)
RED(td_count_distinct, OP_COUNT_DISTINCT, TD_I64)
RED(td_stddev, OP_STDDEV, TD_F64)
RED(td_stddev_pop, OP_STDDEV_POP, TD_F64)
RED(td_var, OP_VAR, TD_F64)
RED(td_var_pop, OP_VAR_POP, TD_F64)

/* ------------------------------------------------------------ structural -- */

td_op_t* td_filter(td_graph_t* g, td_op_t* input, td_op_t* pred) {
    uint32_t iid = input->id, pid = pred->id, est = input->est_rows / 2;
    td_op_t* n = new_node(g);
    if (!n) return NULL;
    n->opcode = OP_FILTER;
    n->arity = 2;
    n->inputs[0] = &g->nodes[iid];
    n->inputs[1] = &g->nodes[pid];
    n->out_type = g->nodes[iid].out_type;
    n->est_rows = est;
    return n;
}

/* GROUP node: after the ext header come the key node pointers, the
 * aggregate opcodes and (pointer-aligned) the aggregate input pointers.
 * Inputs are remembered by id, since adding a node may move g->nodes. */
td_op_t* td_group(td_graph_t* g, td_op_t** keys, uint8_t n_keys,
                  uint16_t* agg_ops, td_op_t** agg_ins, uint8_t n_aggs) {
    uint32_t ids[512];
    for (int i = 0; i < n_keys + n_aggs; i++)
        ids[i] = i < n_keys ? keys[i]->id : agg_ins[i - n_keys]->id;
    const size_t at_ops = (size_t)n_keys * sizeof(td_op_t*);
    size_t at_ins = at_ops + (size_t)n_aggs * sizeof(uint16_t);
    at_ins += (sizeof(td_op_t*) - at_ins % sizeof(td_op_t*)) % sizeof(td_op_t*);
    td_op_ext_t* e = new_ext(g, at_ins + (size_t)n_aggs * sizeof(td_op_t*));
    if (!e) return NULL;
    char* trail = TRAIL(e);
    e->keys = (td_op_t**)trail;
    e->agg_ops = (uint16_t*)(trail + at_ops);
    e->agg_ins = (td_op_t**)(trail + at_ins);
    e->n_keys = n_keys;
    e->n_aggs = n_aggs;
    for (int i = 0; i < n_keys; i++) e->keys[i] = &g->nodes[ids[i]];
    for (int a = 0; a < n_aggs; a++) {
        e->agg_ops[a] = agg_ops[a];
        e->agg_ins[a] = &g->nodes[ids[n_keys + a]];
    }
    e->base.opcode = OP_GROUP;
    e->base.out_type = TD_TABLE;
    if (n_keys > 0) {
        /* a tenth of the first key's rows: the planner's group-count guess */
        e->base.inputs[0] = e->keys[0];
        e->base.est_rows = e->keys[0]->est_rows / 10;
    }
    return publish(g, e);
}

td_op_t* td_distinct(td_graph_t* g, td_op_t** keys, uint8_t n_keys) {
    return td_group(g, keys, n_keys, NULL, NULL, 0);
}

td_op_t* td_join(td_graph_t* g, td_op_t* lt, td_op_t** lkeys, td_op_t* rt,
                 td_op_t** rkeys, uint8_t n_keys, uint8_t join_type) {
    uint32_t lid = lt->id, rid = rt->id, lk[256], rk[256];
    for (int i = 0; i < n_keys; i++) { lk[i] = lkeys[i]->id; rk[i] = rkeys[i]->id; }
    size_t ksz = (size_t)n_keys * sizeof(td_op_t*);
    td_op_ext_t* e = new_ext(g, 2 * ksz);
    if (!e) return NULL;
    e->base.opcode = OP_JOIN;
    e->base.arity = 2;
    e->base.inputs[0] = &g->nodes[lid];
    e->base.inputs[1] = &g->nodes[rid];
    e->base.out_type = TD_TABLE;
    e->base.est_rows = g->nodes[lid].est_rows;
    e->join.left_keys = (td_op_t**)TRAIL(e);
    e->join.right_keys = (td_op_t**)(TRAIL(e) + ksz);
    for (int i = 0; i < n_keys; i++) {
        e->join.left_keys[i] = &g->nodes[lk[i]];
        e->join.right_keys[i] = &g->nodes[rk[i]];
    }
    e->join.n_join_keys = n_keys;
    e->join.join_type = join_type;
    return publish(g, e);
}

td_op_t* td_asof_join(td_graph_t* g, td_op_t* lt, td_op_t* rt, td_op_t* time_key,
                      td_op_t** eq_keys, uint8_t n_eq, uint8_t join_type) {
    uint32_t lid = lt->id, rid = rt->id, tid = time_key->id, ek[256];
    for (int i = 0; i < n_eq; i++) ek[i] = eq_keys[i]->id;
    td_op_ext_t* e = new_ext(g, (size_t)n_eq * sizeof(td_op_t*));
    if (!e) return NULL;
    e->base.opcode = OP_WINDOW_JOIN;
    e->base.arity = 2;
    e->base.inputs[0] = &g->nodes[lid];
    e->base.inputs[1] = &g->nodes[rid];
    e->base.out_type = TD_TABLE;
    e->base.est_rows = g->nodes[lid].est_rows;
    e->asof.time_key = &g->nodes[tid];
    e->asof.eq_keys = (td_op_t**)TRAIL(e);
    for (int i = 0; i < n_eq; i++) e->asof.eq_keys[i] = &g->nodes[ek[i]];
    e->asof.n_eq_keys = n_eq;
    e->asof.join_type = join_type;
    return publish(g, e);
}

td_op_t* td_sort_op(td_graph_t* g, td_op_t* tn, td_op_t** keys, uint8_t* descs,
                    uint8_t* nulls_first, uint8_t n_cols) {
    uint32_t tid = tn->id, kid[256];
    for (int i = 0; i < n_cols; i++) kid[i] = keys[i]->id;
    size_t ksz = (size_t)n_cols * sizeof(td_op_t*);
    td_op_ext_t* e = new_ext(g, ksz + 2 * (size_t)n_cols);
    if (!e) return NULL;
    e->base.opcode = OP_SORT;
    e->base.arity = 1;
    e->base.inputs[0] = &g->nodes[tid];
    e->base.out_type = TD_TABLE;
    e->base.est_rows = g->nodes[tid].est_rows;
    e->sort.columns = (td_op_t**)TRAIL(e);
    for (int i = 0; i < n_cols; i++) e->sort.columns[i] = &g->nodes[kid[i]];
    e->sort.desc = (uint8_t*)TRAIL(e) + ksz;
    if (n_cols) memcpy(e->sort.desc, descs, n_cols);
    e->sort.nulls_first = (uint8_t*)TRAIL(e) + ksz + n_cols;
    for (int i = 0; i < n_cols; i++)
        e->sort.nulls_first[i] = nulls_first ? nulls_first[i] : (descs[i] ? 1 : 0);
    e->sort.n_cols = n_cols;
    return publish(g, e);
}

td_op_t* td_window_op(td_graph_t* g, td_op_t* tn, td_op_t** part_keys, uint8_t n_part,
                      td_op_t** order_keys, uint8_t* order_descs, uint8_t n_order,
                      uint8_t* func_kinds, td_op_t** func_inputs, int64_t* func_params,
                      uint8_t n_funcs, uint8_t frame_type, uint8_t frame_start,
                      uint8_t frame_end, int64_t frame_start_n, int64_t frame_end_n) {
    uint32_t tid = tn->id, pk[256], ok[256], fi[256];
    for (int i = 0; i < n_part; i++) pk[i] = part_keys[i]->id;
    for (int i = 0; i < n_order; i++) ok[i] = order_keys[i]->id;
    for (int i = 0; i < n_funcs; i++) fi[i] = func_inputs[i]->id;
    size_t P = sizeof(td_op_t*);
    size_t o_ok = (size_t)n_part * P, o_od = o_ok + (size_t)n_order * P;
    size_t o_fi = (o_od + n_order + P - 1) & ~(P - 1);
    size_t o_fk = o_fi + (size_t)n_funcs * P;
    size_t o_fp = (o_fk + n_funcs + 7) & ~(size_t)7;
    td_op_ext_t* e = new_ext(g, o_fp + (size_t)n_funcs * 8);
    if (!e) return NULL;
    e->base.opcode = OP_WINDOW;
    e->base.arity = 1;
    e->base.inputs[0] = &g->nodes[tid];
    e->base.out_type = TD_TABLE;
    e->base.est_rows = g->nodes[tid].est_rows;
    char* t = TRAIL(e);
    e->window.part_keys = (td_op_t**)t;
    for (int i = 0; i < n_part; i++) e->window.part_keys[i] = &g->nodes[pk[i]];
    e->window.order_keys = (td_op_t**)(t + o_ok);
    for (int i = 0; i < n_order; i++) e->window.order_keys[i] = &g->nodes[ok[i]];
    e->window.order_descs = (uint8_t*)(t + o_od);
    if (n_order) memcpy(e->window.order_descs, order_descs, n_order);
    e->window.func_inputs = (td_op_t**)(t + o_fi);
    for (int i = 0; i < n_funcs; i++) e->window.func_inputs[i] = &g->nodes[fi[i]];
    e->window.func_kinds = (uint8_t*)(t + o_fk);
    if (n_funcs) memcpy(e->window.func_kinds, func_kinds, n_funcs);
    e->window.func_params = (int64_t*)(t + o_fp);
    if (n_funcs) memcpy(e->window.func_params, func_params, (size_t)n_funcs * 8);
    e->window.n_part_keys = n_part;
    e->window.n_order_keys = n_order;
    e->window.n_funcs = n_funcs;
    e->window.frame_type = frame_type;
    e->window.frame_start = frame_start;
    e->window.frame_end = frame_end;
    e->window.frame_start_n = frame_start_n;
    e->window.frame_end_n = frame_end_n;
    return publish(g, e);
}

td_op_t* td_select(td_graph_t* g, td_op_t* input, td_op_t** cols, uint8_t n_cols) {
    uint32_t iid = input->id, cid[256];
    for (int i = 0; i < n_cols; i++) cid[i] = cols[i]->id;
    td_op_ext_t* e = new_ext(g, (size_t)n_cols * sizeof(td_op_t*));
    if (!e) return NULL;
    e->base.opcode = OP_SELECT;
    e->base.arity = 1;
    e->base.inputs[0] = &g->nodes[iid];
    e->base.out_type = TD_TABLE;
    e->base.est_rows = g->nodes[iid].est_rows;
    e->sort.columns = (td_op_t**)TRAIL(e);
    for (int i = 0; i < n_cols; i++) e->sort.columns[i] = &g->nodes[cid[i]];
    e->sort.n_cols = n_cols;
    return publish(g, e);
}

static td_op_t* count_op(td_graph_t* g, uint16_t op, td_op_t* input, int64_t n) {
    uint32_t iid = input->id;
    td_op_ext_t* e = new_ext(g, 0);
    if (!e) return NULL;
    e->base.opcode = op;
    e->base.arity = 1;
    e->base.inputs[0] = &g->nodes[iid];
    e->base.out_type = g->nodes[iid].out_type;
    e->base.est_rows = (uint32_t)n;
    e->sym = n;
    return publish(g, e);
}
td_op_t* td_head(td_graph_t* g, td_op_t* input, int64_t n) { return count_op(g, OP_HEAD, input, n); }
td_op_t* td_tail(td_graph_t* g, td_op_t* input, int64_t n) { return count_op(g, OP_TAIL, input, n); }

td_op_t* td_alias(td_graph_t* g, td_op_t* input, const char* name) {
    td_op_t* n = count_op(g, OP_ALIAS, input, 0);
    if (!n) return NULL;
    td_op_ext_t* e = td__find_ext(g, n->id);
    e->sym = td_sym_intern(name, strlen(name));
    e->base.est_rows = g->nodes[n->id].inputs[0]->est_rows;
    return publish(g, e);
}

td_op_t* td_materialize(td_graph_t* g, td_op_t* input) {
    return unary(g, OP_MATERIALIZE, input, input->out_type);
}

/* Graph traversal operators build their nodes so that DAGs written for the
 * reference still construct; executing them reports TD_ERR_NYI. */
static td_op_t* graph_op(td_graph_t* g, uint16_t op, td_op_t* a, td_op_t* b, void* rel,
                         uint8_t dir, uint8_t mind, uint8_t maxd, uint8_t track) {
    uint32_t aid = a ? a->id : 0, bid = b ? b->id : 0;
    td_op_ext_t* e = new_ext(g, 0);
    if (!e) return NULL;
    e->base.opcode = op;
    e->base.arity = (uint8_t)((a != NULL) + (b != NULL));
    if (a) e->base.inputs[0] = &g->nodes[aid];
    if (b) e->base.inputs[1] = &g->nodes[bid];
    e->base.out_type = TD_TABLE;
    e->graph.rel = rel;
    e->graph.direction = dir;
    e->graph.min_depth = mind;
    e->graph.max_depth = maxd;
    e->graph.path_tracking = track;
    return publish(g, e);
}
td_op_t* td_expand(td_graph_t* g, td_op_t* src, td_rel_t* rel, uint8_t dir) {
    return graph_op(g, OP_EXPAND, src, NULL, rel, dir, 1, 1, 0);
}
td_op_t* td_var_expand(td_graph_t* g, td_op_t* s, td_rel_t* rel, uint8_t dir,
                       uint8_t mind, uint8_t maxd, bool track) {
    return graph_op(g, OP_VAR_EXPAND, s, NULL, rel, dir, mind, maxd, track);
}
td_op_t* td_shortest_path(td_graph_t* g, td_op_t* s, td_op_t* d, td_rel_t* rel, uint8_t maxd) {
    return graph_op(g, OP_SHORTEST_PATH, s, d, rel, 0, 0, maxd, 0);
}
td_op_t* td_wco_join(td_graph_t* g, td_rel_t** rels, uint8_t n_rels, uint8_t n_vars) {
    td_op_ext_t* e = new_ext(g, (size_t)n_rels * sizeof(td_rel_t*));
    if (!e) return NULL;
    e->base.opcode = OP_WCO_JOIN;
    e->base.out_type = TD_TABLE;
    e->wco.rels = (void**)TRAIL(e);
    if (n_rels) memcpy(e->wco.rels, rels, (size_t)n_rels * sizeof(td_rel_t*));
    e->wco.n_rels = n_rels;
    e->wco.n_vars = n_vars;
    return publish(g, e);
}

td_op_ext_t* td__find_ext(td_graph_t* g, uint32_t id) {
    for (uint32_t i = 0; i < g->ext_count; i++)
        if (g->ext_nodes[i] && g->ext_nodes[i]->base.id == id) return g->ext_nodes[i];
    return NULL;
}
