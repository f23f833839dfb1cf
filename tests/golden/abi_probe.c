/* Prints the byte layout of the td.h ABI types, one "name value" per line.
 * Compiled once against this repository's header and once against the
 * reference header; the two outputs must be identical. */
#include <stdio.h>
#include <stddef.h>
#include <teide/td.h>

#define P(name, expr) printf("%s %zu\n", name, (size_t)(expr))
int main(void) {
    P("sizeof_td_t", sizeof(td_t));
    P("alignof_td_t", _Alignof(td_t));
    P("off_mmod", offsetof(td_t, mmod));
    P("off_order", offsetof(td_t, order));
    P("off_type", offsetof(td_t, type));
    P("off_attrs", offsetof(td_t, attrs));
    P("off_rc", offsetof(td_t, rc));
    P("off_len", offsetof(td_t, len));
    P("off_data", offsetof(td_t, data));
    P("off_sym_dict", offsetof(td_t, sym_dict));
    P("off_slice_offset", offsetof(td_t, slice_offset));
    P("sizeof_td_op_t", sizeof(td_op_t));
    P("off_op_id", offsetof(td_op_t, id));
    P("off_op_est", offsetof(td_op_t, est_rows));
    P("off_op_inputs", offsetof(td_op_t, inputs));
    P("sizeof_td_op_ext_t", sizeof(td_op_ext_t));
    P("off_ext_keys", offsetof(td_op_ext_t, keys));
    P("off_ext_n_keys", offsetof(td_op_ext_t, n_keys));
    P("off_ext_n_aggs", offsetof(td_op_ext_t, n_aggs));
    P("off_ext_agg_ops", offsetof(td_op_ext_t, agg_ops));
    P("off_ext_agg_ins", offsetof(td_op_ext_t, agg_ins));
    P("off_ext_join_nk", offsetof(td_op_ext_t, join.n_join_keys));
    P("off_ext_join_type", offsetof(td_op_ext_t, join.join_type));
    P("off_ext_join_rk", offsetof(td_op_ext_t, join.right_keys));
    P("off_ext_sort_ncols", offsetof(td_op_ext_t, sort.n_cols));
    P("off_ext_win_fen", offsetof(td_op_ext_t, window.frame_end_n));
    P("off_ext_graph_fact", offsetof(td_op_ext_t, graph.factorized));
    P("off_ext_wco_nvars", offsetof(td_op_ext_t, wco.n_vars));
    P("sizeof_td_graph_t", sizeof(td_graph_t));
    P("off_g_table", offsetof(td_graph_t, table));
    P("off_g_tables", offsetof(td_graph_t, tables));
    P("off_g_n_tables", offsetof(td_graph_t, n_tables));
    P("off_g_ext_nodes", offsetof(td_graph_t, ext_nodes));
    P("off_g_ext_count", offsetof(td_graph_t, ext_count));
    P("off_g_selection", offsetof(td_graph_t, selection));
    P("sizeof_td_morsel_t", sizeof(td_morsel_t));
    P("sizeof_td_sel_meta_t", sizeof(td_sel_meta_t));
    P("sizeof_td_pipe_t", sizeof(td_pipe_t));
    P("sizeof_td_mem_stats_t", sizeof(td_mem_stats_t));
    P("TD_TYPE_COUNT", TD_TYPE_COUNT);
    P("TD_ERR_CANCEL", TD_ERR_CANCEL);
    P("OP_WCO_JOIN", OP_WCO_JOIN);
    P("OP_VAR_POP", OP_VAR_POP);
    P("OP_MATERIALIZE", OP_MATERIALIZE);
    P("TD_PARALLEL_THRESHOLD", TD_PARALLEL_THRESHOLD);
    return 0;
}
