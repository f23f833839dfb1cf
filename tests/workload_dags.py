"""The benchmark workloads' query DAGs, written once against the td.h mirror
(teide_amd.td) so the same code drives the GPU engine and the reference
build.  Shapes follow oracle/ref_harness.c."""
import numpy as np

from teide_amd import workloads as W


def fg_table(e, n, seed=42):
    d = W.filter_group(n, seed)
    return e.table([("flag", d["flag"]), ("region", d["region"]), ("amount", d["amount"])]), d


def run_filter_table(e, tbl, col, op, c):
    """FILTER(CONST_TABLE t, col OP c) executed to a compacted table."""
    g = e.graph(tbl)
    pred = getattr(g, op)(g.scan(col), g.const_i64(c))
    root = g.optimize(g.filter(g.const_table(tbl), pred))
    out = g.execute(root)
    g.free()
    return out


def run_group_fg(e, tbl, sel=None):
    """GROUP(region; SUM amount, COUNT amount)."""
    g = e.graph(tbl)
    if sel:
        g.set_selection(sel)
    root = g.group([g.scan("region")], [("SUM", g.scan("amount")), ("COUNT", g.scan("amount"))])
    out = g.execute(g.optimize(root))
    g.free()
    return out


def fg_reference_form(e, tbl):
    """Filter to a compacted table, then group it on a new graph."""
    f = run_filter_table(e, tbl, "flag", "eq", 0)
    out = run_group_fg(e, f)
    e.release(f)
    return out


def fg_readme_form(e, tbl):
    """The README quick-start shape: keys / aggregates are FILTER(SCAN, P)."""
    g = e.graph(tbl)
    pred = g.eq(g.scan("flag"), g.const_i64(0))
    key = g.filter(g.scan("region"), pred)
    amt = g.filter(g.scan("amount"), pred)
    root = g.group([key], [("SUM", amt), ("COUNT", amt)])
    out = g.execute(root)
    g.free()
    return out


def jg_tables(e, n_fact, n_dim, n_cat, seed=42):
    dim = W.join_dim(n_dim, n_cat, seed)
    fact = W.join_fact(n_fact, n_dim, seed)
    td = e.table([("pk", dim["pk"]), ("cat", dim["cat"])])
    tf = e.table([("fk", fact["fk"]), ("qty", fact["qty"]), ("price", fact["price"])])
    return tf, td, fact, dim


def jg_join(e, fact, dim, thr=50):
    """JOIN(FILTER(CONST_TABLE fact, GT(SCAN qty, thr)), [SCAN fk],
            CONST_TABLE dim, [SCAN pk], inner)."""
    g = e.graph(fact)
    flt = g.filter(g.const_table(fact), g.gt(g.scan("qty"), g.const_i64(thr)))
    root = g.join(flt, [g.scan("fk")], g.const_table(dim), [g.scan("pk")], 0)
    out = g.execute(g.optimize(root))
    g.free()
    return out


def jg_group(e, joined):
    """GROUP(cat; SUM price, SUM qty, COUNT price) on the join result."""
    g = e.graph(joined)
    root = g.group([g.scan("cat")], [("SUM", g.scan("price")), ("SUM", g.scan("qty")),
                                     ("COUNT", g.scan("price"))])
    out = g.execute(g.optimize(root))
    g.free()
    return out


def jg_two_call(e, fact, dim, thr=50):
    """The reference's composition: two td_execute calls."""
    j = jg_join(e, fact, dim, thr)
    out = jg_group(e, j)
    e.release(j)
    return out


def jg_fused(e, fact, dim, thr=50):
    """One DAG: GROUP over the graph's JOIN output (fused on the device)."""
    g = e.graph(fact)
    flt = g.filter(g.const_table(fact), g.gt(g.scan("qty"), g.const_i64(thr)))
    g.join(flt, [g.scan("fk")], g.const_table(dim), [g.scan("pk")], 0)
    root = g.group([g.scan("cat")], [("SUM", g.scan("price")), ("SUM", g.scan("qty")),
                                     ("COUNT", g.scan("price"))])
    out = g.execute(root)
    g.free()
    return out


def hc_table(e, n, n_values, n_sym, seed=42):
    d = W.high_card(n, n_values, n_sym, 1.1, seed)
    t = e.table([("key", d["key"]), ("sym", d["sym"], 20, 2), ("v_i64", d["v_i64"]),
                 ("v_f64", d["v_f64"])])
    return t, d


HC_AGGS = [("SUM", "v_i64"), ("MIN", "v_i64"), ("MAX", "v_i64"), ("SUM", "v_f64"),
           ("MIN", "v_f64"), ("MAX", "v_f64"), ("COUNT", "v_i64")]


def hc_group(e, tbl, key):
    g = e.graph(tbl)
    root = g.group([g.scan(key)], [(op, g.scan(c)) for op, c in HC_AGGS])
    out = g.execute(g.optimize(root))
    g.free()
    return out


def cols(e, tbl):
    return e.read_table(tbl)
