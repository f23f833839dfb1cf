/*
 * opt.c -- td_optimize: the rewrites that matter on the filter -> join ->
 * group-by path, plus the plan printer.
 *
 * Of the reference's ten passes (src/ops/opt.c:1641-1675) only two change
 * this path's DAGs: folding a FILTER whose predicate is a constant into
 * MATERIALIZE (keep all rows) or HEAD 0 (keep none) (opt.c:461-494), and
 * splitting FILTER(AND(a, b)) into FILTER(FILTER(x, a), b) (opt.c:1304-1342),
 * which the device evaluates as one AND-ed selection anyway.  The others
 * (type inference, SIP, factorisation, pushdown into SELECT/EXPAND,
 * projection marking, partition pruning, fusion flags, DCE) do not change
 * the result of the supported shapes and are not run.
 */
#include "internal.h"

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

/* FILTER(x, AND(a, b)) -> FILTER(FILTER(x, a), b), repeatedly. */
static td_op_t* split_ands(td_graph_t* g, td_op_t* root) {
    if (!root || root->opcode != OP_FILTER || root->arity != 2) return root;
    td_op_t* p = root->inputs[1];
    if (!p || p->opcode != OP_AND) return root;
    uint32_t rid = root->id, aid = p->inputs[0]->id, bid = p->inputs[1]->id;
    td_op_t* outer = NULL;
    if (g->node_count < g->node_cap) {
        outer = &g->nodes[g->node_count];
        memset(outer, 0, sizeof(*outer));
        outer->id = g->node_count++;
    } else {
        return root;                       /* no room to split: leave as is */
    }
    root = &g->nodes[rid];
    root->inputs[1] = &g->nodes[aid];
    outer->opcode = OP_FILTER;
    outer->arity = 2;
    outer->inputs[0] = root;
    outer->inputs[1] = &g->nodes[bid];
    outer->out_type = root->out_type;
    outer->est_rows = root->est_rows;
    split_ands(g, root);
    return split_ands(g, outer);
}

td_op_t* td_optimize(td_graph_t* g, td_op_t* root) {
    if (!g || !root) return root;
    uint8_t* seen = calloc(g->node_count + 1, 1);
    if (!seen) return root;
    fold_filters(g, root, seen);
    free(seen);
    return split_ands(g, root);
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
