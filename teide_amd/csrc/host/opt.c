/*
 * opt.c -- td_optimize: the rewrites that matter on the filter -> join ->
 * group-by path, plus the plan printer.
 *
 * Of the reference's ten passes (src/ops/opt.c:1641-1675) three change
 * this path's DAGs: constant folding of element-wise nodes over constants
 * (opt.c:192-507), whose arithmetic differs from the fused evaluator's in
 * places (IEEE division by zero, fmin / fmax), so results depend on it;
 * folding a FILTER whose predicate is a constant into MATERIALIZE (keep all
 * rows) or HEAD 0 (keep none) (opt.c:461-494); and splitting every
 * FILTER(AND(a, b)) into FILTER(FILTER(x, a), b) with each chain of filters
 * ordered cheapest-first (opt.c:1258-1459), which the device evaluates as
 * one AND-ed selection anyway.  The others (type
 * inference, SIP, factorisation, pushdown into SELECT/EXPAND, projection
 * marking, partition pruning, fusion flags, DCE) do not change the result
 * of the supported shapes and are not run.
 */
#include "internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static int const_bool(td_graph_t* g, td_op_t* op, int* out) {
    if (op->opcode != OP_CONST) return 0;
    td_op_ext_t* e = td__find_ext(g, op->id);
    if (!e || !e->literal || !td_is_atom(e->literal)) return 0;
    td_t* a = e->literal;
    switch (a->type) {
    case TD_ATOM_BOOL: case TD_ATOM_U8: *out = a->u8 != 0; return 1;
    case TD_ATOM_I64: *out = a->i64 != 0; return 1;
    case TD_ATOM_I32: *out = a->i32 != 0; return 1;
    case TD_ATOM_F64: *out = a->f64 != 0.0; return 1;
    default: return 0;
    }
}

static td_op_ext_t* ensure_ext(td_graph_t* g, td_op_t* node) {
    td_op_ext_t* e = td__find_ext(g, node->id);
    if (e) return e;
    if (g->ext_count == g->ext_cap) {
        uint32_t nc = g->ext_cap ? g->ext_cap * 2 : 16;
        td_op_ext_t** ne = realloc(g->ext_nodes, nc * sizeof(td_op_ext_t*));
        if (!ne) return NULL;
        g->ext_nodes = ne;
        g->ext_cap = nc;
    }
    e = calloc(1, sizeof(td_op_ext_t));
    if (!e) return NULL;
    e->base = *node;
    g->ext_nodes[g->ext_count++] = e;
    return e;
}

/* ---------------------------------------------------- constant folding -- */

/* Numeric value of a CONST node's atom: integer atoms widen to i64, f64
 * atoms also give their truncation (opt.c:242-276). */
static int const_num(td_graph_t* g, td_op_t* op, int64_t* i, double* f, int* isf) {
    if (!op || op->opcode != OP_CONST) return 0;
    td_op_ext_t* e = td__find_ext(g, op->id);
    if (!e || !e->literal || !td_is_atom(e->literal)) return 0;
    td_t* a = e->literal;
    *isf = 0;
    switch (a->type) {
    case TD_ATOM_F64:
        *f = a->f64;
        *i = (int64_t)a->f64;
        *isf = 1;
        return 1;
    case TD_ATOM_I64: case TD_ATOM_SYM: case TD_ATOM_DATE: case TD_ATOM_TIME:
    case TD_ATOM_TIMESTAMP:
        *i = a->i64; break;
    case TD_ATOM_I32: *i = a->i32; break;
    case TD_ATOM_I16: *i = a->i16; break;
    case TD_ATOM_U8: case TD_ATOM_BOOL: *i = a->u8; break;
    default: return 0;
    }
    *f = (double)*i;
    return 1;
}

/* The value of a unary element-wise op over a constant, as the reference's
 * optimizer computes it (opt.c:305-358); NULL when it leaves the node. */
static td_t* fold_unary(uint16_t op, int64_t i, double f, int isf) {
    double x = isf ? f : (double)i;
    switch (op) {
    case OP_NEG:
        if (isf) return td_f64(-f);
        return i == INT64_MIN ? NULL : td_i64(-i);
    case OP_ABS:
        if (isf) return td_f64(fabs(f));
        return i == INT64_MIN ? NULL : td_i64(i < 0 ? -i : i);
    case OP_NOT: return td_bool(isf ? f == 0.0 : i == 0);
    case OP_SQRT: return td_f64(sqrt(x));
    case OP_LOG: return td_f64(log(x));
    case OP_EXP: return td_f64(exp(x));
    case OP_CEIL: return isf ? td_f64(ceil(f)) : td_i64(i);
    case OP_FLOOR: return isf ? td_f64(floor(f)) : td_i64(i);
    default: return NULL;
    }
}

/* A binary element-wise op over two constants in the node's declared type
 * (opt.c:360-447): f64 with IEEE division and fmod, fmin / fmax; i64
 * wrapping with DIV / MOD by zero giving 0; bool comparisons in f64. */
static td_t* fold_binary(uint16_t op, int8_t t, int64_t li, double lf, int lfl, int64_t ri,
                         double rf, int rfl) {
    double a = lfl ? lf : (double)li, b = rfl ? rf : (double)ri;
    if (t == TD_F64) {
        switch (op) {
        case OP_ADD: return td_f64(a + b);
        case OP_SUB: return td_f64(a - b);
        case OP_MUL: return td_f64(a * b);
        case OP_DIV: return td_f64(a / b);
        case OP_MOD: return td_f64(fmod(a, b));
        case OP_MIN2: return td_f64(fmin(a, b));
        case OP_MAX2: return td_f64(fmax(a, b));
        default: return NULL;
        }
    }
    if (t == TD_I64) {
        int64_t x = lfl ? (int64_t)lf : li, y = rfl ? (int64_t)rf : ri;
        int ok = y != 0 && !(x == INT64_MIN && y == -1);
        switch (op) {
        case OP_ADD: return td_i64((int64_t)((uint64_t)x + (uint64_t)y));
        case OP_SUB: return td_i64((int64_t)((uint64_t)x - (uint64_t)y));
        case OP_MUL: return td_i64((int64_t)((uint64_t)x * (uint64_t)y));
        case OP_DIV: return td_i64(ok ? x / y : 0);
        case OP_MOD: return td_i64(ok ? x % y : 0);
        case OP_MIN2: return td_i64(x < y ? x : y);
        case OP_MAX2: return td_i64(x > y ? x : y);
        default: return NULL;
        }
    }
    if (t == TD_BOOL) {
        switch (op) {
        case OP_EQ: return td_bool(a == b);
        case OP_NE: return td_bool(a != b);
        case OP_LT: return td_bool(a < b);
        case OP_LE: return td_bool(a <= b);
        case OP_GT: return td_bool(a > b);
        case OP_GE: return td_bool(a >= b);
        case OP_AND: return td_bool(a != 0.0 && b != 0.0);
        case OP_OR: return td_bool(a != 0.0 || b != 0.0);
        default: return NULL;
        }
    }
    return NULL;
}

/* Element-wise nodes whose inputs are all constants become constants,
 * children first, so whole constant subtrees collapse (opt.c:496-507).
 * Nodes that already carry an ext record are left alone. */
static void fold_consts(td_graph_t* g, td_op_t* op, uint8_t* seen) {
    if (!op || seen[op->id]) return;
    seen[op->id] = 1;
    for (int i = 0; i < op->arity && i < 2; i++) fold_consts(g, op->inputs[i], seen);
    td_op_ext_t* e = td__find_ext(g, op->id);
    if (e && op->opcode == OP_GROUP) {
        for (int k = 0; k < e->n_keys; k++) fold_consts(g, e->keys[k], seen);
        for (int a = 0; a < e->n_aggs; a++) fold_consts(g, e->agg_ins[a], seen);
        return;
    }
    if (e && op->opcode == OP_SORT) {
        for (int k = 0; k < e->sort.n_cols; k++) fold_consts(g, e->sort.columns[k], seen);
        return;
    }
    if (e) return;
    td_t* v = NULL;
    int64_t li, ri;
    double lf, rf;
    int lfl, rfl;
    if (op->arity == 1 && op->opcode >= OP_NEG && op->opcode <= OP_FLOOR &&
        const_num(g, op->inputs[0], &li, &lf, &lfl)) {
        v = fold_unary(op->opcode, li, lf, lfl);
    } else if (op->arity == 2 && op->opcode >= OP_ADD && op->opcode <= OP_MAX2 &&
               const_num(g, op->inputs[0], &li, &lf, &lfl) &&
               const_num(g, op->inputs[1], &ri, &rf, &rfl)) {
        v = fold_binary(op->opcode, op->out_type, li, lf, lfl, ri, rf, rfl);
    }
    if (!v || TD_IS_ERR(v)) return;
    td_op_ext_t* x = ensure_ext(g, op);
    if (!x) { td_release(v); return; }
    op->opcode = OP_CONST;
    op->arity = 0;
    op->inputs[0] = op->inputs[1] = NULL;
    op->out_type = (int8_t)(v->type < 0 ? -v->type : v->type);
    x->base = *op;
    x->literal = v;
}

static void fold_filters(td_graph_t* g, td_op_t* op, uint8_t* seen) {
    if (!op || seen[op->id]) return;
    seen[op->id] = 1;
    for (int i = 0; i < op->arity && i < 2; i++) fold_filters(g, op->inputs[i], seen);
    td_op_ext_t* e = td__find_ext(g, op->id);
    if (e && op->opcode == OP_GROUP) {
        for (int k = 0; k < e->n_keys; k++) fold_filters(g, e->keys[k], seen);
        for (int a = 0; a < e->n_aggs; a++) fold_filters(g, e->agg_ins[a], seen);
    }
    int keep;
    if (op->opcode == OP_FILTER && op->arity == 2 && const_bool(g, op->inputs[1], &keep)) {
        if (keep) {
            op->opcode = OP_MATERIALIZE;
            op->arity = 1;
            op->inputs[1] = NULL;
        } else {
            td_op_ext_t* x = ensure_ext(g, op);
            if (!x) return;
            op->opcode = OP_HEAD;
            op->arity = 1;
            op->inputs[1] = NULL;
            op->est_rows = 0;
            x->base = *op;
            x->sym = 0;
        }
    }
}

/* Every FILTER(x, AND(a, b)) in the graph becomes FILTER(FILTER(x, a), b),
 * until no AND predicate is left under a FILTER (the split of
 * pass_filter_reorder, opt.c:1304-1395).  The rewritten node keeps its id as
 * the outer filter on b and a new node filters x on a, so whatever consumed
 * the node still sees the outer filter and needs no redirection. */
static void split_ands(td_graph_t* g) {
    for (int round = 0; round < 16; round++) {
        int changed = 0;
        uint32_t nc = g->node_count;
        for (uint32_t i = 0; i < nc; i++) {
            td_op_t* n = &g->nodes[i];
            if (n->opcode != OP_FILTER || n->arity != 2 || !n->inputs[0] || !n->inputs[1] ||
                n->inputs[1]->opcode != OP_AND || n->inputs[1]->arity != 2)
                continue;
            uint32_t xid = n->inputs[0]->id, aid = n->inputs[1]->inputs[0]->id,
                     bid = n->inputs[1]->inputs[1]->id;
            td_op_t* inner = td__graph_new_node(g);          /* may move g->nodes */
            if (!inner) return;
            n = &g->nodes[i];
            inner->opcode = OP_FILTER;
            inner->arity = 2;
            inner->inputs[0] = &g->nodes[xid];
            inner->inputs[1] = &g->nodes[aid];
            inner->out_type = n->out_type;
            inner->est_rows = n->est_rows;
            n->inputs[0] = inner;
            n->inputs[1] = &g->nodes[bid];
            changed = 1;
        }
        if (!changed) return;
    }
}

/* Cost of a filter predicate, lower runs first (filter_cost,
 * opt.c:1258-1299): +4 when neither input is a constant, +0..3 by the
 * first input's width (BOOL / U8, I16, I32 / DATE / TIME, wider), +0 for
 * EQ / NE, +2 for ordered comparisons, +1 otherwise. */
static int pred_cost(td_op_t* p) {
    if (!p) return 99;
    int c = 0, has_const = 0;
    for (int i = 0; i < p->arity && i < 2; i++)
        if (p->inputs[i] && p->inputs[i]->opcode == OP_CONST) has_const = 1;
    if (!has_const) c += 4;
    int8_t t = p->arity >= 1 && p->inputs[0] ? p->inputs[0]->out_type : p->out_type;
    c += t == TD_BOOL || t == TD_U8 ? 0 : t == TD_I16 ? 1 :
         (t == TD_I32 || t == TD_DATE || t == TD_TIME) ? 2 : 3;
    switch (p->opcode) {
    case OP_EQ: case OP_NE: break;
    case OP_LT: case OP_LE: case OP_GT: case OP_GE: c += 2; break;
    case OP_LIKE: case OP_ILIKE: c += 4; break;
    default: c += 1; break;
    }
    return c;
}

/* Nodes and ext records that take node `id` as an input. */
static int consumers(td_graph_t* g, uint32_t id) {
    int n = 0;
    for (uint32_t i = 0; i < g->node_count; i++)
        for (int k = 0; k < g->nodes[i].arity && k < 2; k++)
            if (g->nodes[i].inputs[k] && g->nodes[i].inputs[k]->id == id) n++;
    for (uint32_t e = 0; e < g->ext_count; e++) {
        td_op_ext_t* x = g->ext_nodes[e];
        if (!x || x->base.opcode != OP_GROUP) continue;
        for (int k = 0; k < x->n_keys; k++) if (x->keys[k] && x->keys[k]->id == id) n++;
        for (int a = 0; a < x->n_aggs; a++) if (x->agg_ins[a] && x->agg_ins[a]->id == id) n++;
    }
    return n;
}

/* Within each chain of stacked FILTERs that nothing else shares, order the
 * predicates by cost: the cheapest on the innermost filter (evaluated
 * first), ties in their original order (pass_filter_reorder's stable
 * insertion sort, opt.c:1397-1459).  Predicates move between the chain's
 * nodes; the nodes stay where they are. */
static void reorder_chains(td_graph_t* g) {
    uint8_t* done = calloc(g->node_count + 1, 1);
    if (!done) return;
    for (uint32_t i = 0; i < g->node_count; i++) {
        td_op_t* n = &g->nodes[i];
        if (n->opcode != OP_FILTER || done[i]) continue;
        td_op_t* chain[64];
        int cost[64], len = 0, shared = 0;
        for (td_op_t* c = n; c && c->opcode == OP_FILTER && c->arity == 2 && len < 64;
             c = c->inputs[0]) {
            chain[len++] = c;
            done[c->id] = 1;
        }
        if (len < 2) continue;
        for (int c = 0; c < len && !shared; c++) shared = consumers(g, chain[c]->id) > 1;
        if (shared) continue;
        for (int c = 0; c < len; c++) cost[c] = pred_cost(chain[c]->inputs[1]);
        for (int c = 1; c < len; c++) {          /* descending cost, outermost first */
            td_op_t* p = chain[c]->inputs[1];
            int k = cost[c], j = c - 1;
            while (j >= 0 && cost[j] < k) {
                chain[j + 1]->inputs[1] = chain[j]->inputs[1];
                cost[j + 1] = cost[j];
                j--;
            }
            chain[j + 1]->inputs[1] = p;
            cost[j + 1] = k;
        }
    }
    free(done);
}

td_op_t* td_optimize(td_graph_t* g, td_op_t* root) {
    if (!g || !root) return root;
    uint32_t rid = root->id;
    uint8_t* seen = calloc(g->node_count + 1, 1);
    if (!seen) return root;
    fold_consts(g, root, seen);
    memset(seen, 0, g->node_count + 1);
    fold_filters(g, root, seen);
    free(seen);
    split_ands(g);
    reorder_chains(g);
    return &g->nodes[rid];
}

void td_fuse_pass(td_graph_t* g, td_op_t* root) { (void)g; (void)root; }

static const char* op_name(uint16_t op) {
    switch (op) {
    case OP_SCAN: return "SCAN"; case OP_CONST: return "CONST"; case OP_FILTER: return "FILTER";
    case OP_GROUP: return "GROUP"; case OP_JOIN: return "JOIN"; case OP_EQ: return "EQ";
    case OP_NE: return "NE"; case OP_LT: return "LT"; case OP_LE: return "LE"; case OP_GT: return "GT";
    case OP_GE: return "GE"; case OP_AND: return "AND"; case OP_OR: return "OR"; case OP_NOT: return "NOT";
    case OP_SUM: return "SUM"; case OP_COUNT: return "COUNT"; case OP_MIN: return "MIN";
    case OP_MAX: return "MAX"; case OP_AVG: return "AVG"; case OP_HEAD: return "HEAD";
    case OP_MATERIALIZE: return "MATERIALIZE"; case OP_ALIAS: return "ALIAS";
    default: return "OP";
    }
}

static void dump(td_graph_t* g, td_op_t* op, FILE* f, int depth) {
    if (!op || depth > 64) return;
    fprintf(f, "%*s%s#%u", depth * 2, "", op_name(op->opcode), op->id);
    td_op_ext_t* e = td__find_ext(g, op->id);
    if (e && op->opcode == OP_SCAN) {
        size_t n;
        const char* s = td__sym_cstr(e->sym, &n);
        fprintf(f, " %.*s", (int)n, s ? s : "?");
    }
    fprintf(f, "\n");
    for (int i = 0; i < op->arity && i < 2; i++) dump(g, op->inputs[i], f, depth + 1);
    if (e && op->opcode == OP_GROUP) {
        for (int k = 0; k < e->n_keys; k++) dump(g, e->keys[k], f, depth + 1);
        for (int a = 0; a < e->n_aggs; a++) dump(g, e->agg_ins[a], f, depth + 1);
    }
}

void td_graph_dump(td_graph_t* g, td_op_t* root, void* out) {
    if (!g) return;
    dump(g, root, out ? (FILE*)out : stderr, 0);
}
