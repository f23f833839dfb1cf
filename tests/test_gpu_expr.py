"""GPU parity of the fused element-wise expression path (k_expr): the trees
of tests/golden/expr_* against the reference's outputs, random typed trees
over every integer width, symbols and f64 against the oracle
(orc.expr_eval) and, where oracle/_ref is built, the reference library
itself; arithmetic predicates in FILTER (lazy selection and the shared
predicate of GROUP inputs), expression keys and aggregate inputs of GROUP,
SUM over an expression, and expressions too large for one kernel."""
import json
import os

import numpy as np
import pytest

import expr_cases as X
import orc
from parity import compare_groups
from tdtio import read_tdt
from teide_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))


def hc(name):
    argv = MANIFEST[name]["argv"]
    return W.high_card(int(argv[1]), int(argv[2]), int(argv[3]), 1.1, int(argv[4]))


def hc_table(eng, d, resident):
    t = eng.table([("key", d["key"]), ("sym", d["sym"], 20, 2), ("v_i64", d["v_i64"]),
                   ("v_f64", d["v_f64"])])
    if resident:
        r = eng.put(t)
        eng.release(t)
        t = r
    return t


def run_vec(eng, t, e):
    """The tree's value over table t: an array, or for a tree the
    optimizer folds to a constant, that constant."""
    g = eng.graph(t)
    try:
        v = g.execute(g.optimize(X.build(g, e)))
    finally:
        g.free()
    if eng.hdr(v).type < 0:
        out = eng.atom_value(v)
    else:
        out = eng.read_vec(v)
    eng.release(v)
    return out


def oracle(e, cols):
    """(type, value) of the tree as the reference computes it: constant
    subtrees folded by the optimizer's rules, the rest fused."""
    f = orc.expr_fold(e)
    if f[0] in ("i64", "f64", "bool"):
        return f[0], f[1]
    t, v = orc.expr_eval(X.bind(f, cols))
    n = len(next(iter(cols.values())))
    return t, np.broadcast_to(v, (n,))          # CAST of a constant stays a column


def check(got, typ, ref, what):
    if np.ndim(ref) == 0:                       # folded to a constant
        assert np.ndim(got) == 0, what
        if typ == "f64" and np.isnan(ref):
            assert np.isnan(got), what
        else:
            assert got == ref, (what, got, ref)
    else:
        X.check_vec(got, typ, ref, what)


@pytest.mark.parametrize("resident", [False, True])
def test_expr_vec_golden(eng, resident):
    d = hc("expr_vec_4k")
    t = hc_table(eng, d, resident)
    ref = read_tdt(os.path.join(GOLD, "expr_vec_4k.tdt"))
    for i, e in enumerate(X.VEC):
        got = run_vec(eng, t, e)
        t_out = {6: "i64", 7: "f64", 1: "bool"}[ref[i][1]]
        X.check_vec(got, t_out, ref[i][3], f"e{i}")
    eng.release(t)


@pytest.mark.parametrize("resident", [False, True])
def test_expr_grp_golden(eng, resident):
    d = hc("expr_grp_60k")
    t = hc_table(eng, d, resident)
    g = eng.graph(t)
    f = g.execute(g.optimize(g.filter(g.const_table(t), X.build(g, X.GRP_PRED))))
    g.free()
    g = eng.graph(f)
    root = g.group([g.scan("sym")], [(o, X.build(g, e)) for o, e in X.GRP_AGGS])
    out = eng.read_table(g.execute(g.optimize(root)))
    g.free()
    ref = read_tdt(os.path.join(GOLD, "expr_grp_60k.tdt"))
    assert [c[0] for c in out] == [c[0] for c in ref]
    assert [c[1] for c in out] == [c[1] for c in ref]
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]],
                   [ref[0][3].astype(np.int64)], [c[3] for c in ref[1:]],
                   [False, True, False, False, True, False], "expr grp")
    eng.release(f)
    eng.release(t)


# ------------------------------------------------------------ random trees

INT_COLS = {"a": np.int64, "b": np.int32, "c": np.int16, "d": np.uint8, "s": np.uint16,
            "h": np.int64}


def rand_cols(rng, n):
    cols = {
        "a": rng.integers(-1000, 1000, n),
        "b": rng.integers(-2 ** 31, 2 ** 31 - 1, n).astype(np.int32),
        "c": rng.integers(-5, 5, n).astype(np.int16),
        "d": rng.integers(0, 4, n).astype(np.uint8),
        "s": rng.integers(0, 9, n).astype(np.uint16),
        "h": rng.integers(-2 ** 63, 2 ** 63 - 1, n, dtype=np.int64),
        "f": np.round(rng.normal(size=n) * 3, 2),
        "z": rng.normal(size=n) * 1e10,
    }
    cols["f"][rng.integers(0, n, n // 50)] = 0.0
    return cols


def table_of(eng, cols):
    spec = []
    for k, v in cols.items():
        spec.append((k, v, 20, 1) if k == "s" else (k, v))
    return eng.table(spec)


def num_tree(rng, depth, top=True):
    """(tree, type) of a numeric tree whose every operation the reference's
    fused evaluator defines for its operand types.  SQRT / LOG / EXP only at
    the top: libm implementations differ in the last bit, which a MOD or a
    cancelling SUB above them would magnify past any fixed tolerance."""
    if depth == 0 or rng.random() < 0.25:
        r = rng.random()
        if r < 0.55:
            name = str(rng.choice(list(INT_COLS)))
            return ("col", name), "i64"
        if r < 0.75:
            return ("col", str(rng.choice(["f", "z"]))), "f64"
        if r < 0.9:
            return ("i64", int(rng.integers(-9, 10))), "i64"
        return ("f64", float(np.round(rng.normal() * 4, 2))), "f64"
    r = rng.random()
    if r < 0.6:
        op = str(rng.choice(["ADD", "SUB", "MUL", "DIV", "MOD", "MIN2", "MAX2"]))
        a, ta = num_tree(rng, depth - 1, False)
        b, tb = num_tree(rng, depth - 1, False)
        ot = "f64" if "f64" in (ta, tb) or op == "DIV" else "i64"
        return (op, a, b), ot
    a, ta = num_tree(rng, depth - 1, False)
    if r < 0.72:
        return (str(rng.choice(["NEG", "ABS"])), a), ta
    if r < 0.82 and top:
        return (str(rng.choice(["SQRT", "LOG", "EXP"])), ("DIV", a, ("f64", 64.0))), "f64"
    if r < 0.9 and ta == "f64":
        return (str(rng.choice(["CEIL", "FLOOR"])), a), "f64"
    return ("cast", a, "i64" if ta == "f64" else "f64"), "i64" if ta == "f64" else "f64"


def bool_tree(rng, depth):
    r = rng.random()
    if depth == 0 or r < 0.5:
        op = str(rng.choice(["EQ", "NE", "LT", "LE", "GT", "GE"]))
        return (op, num_tree(rng, max(depth - 1, 0))[0], num_tree(rng, max(depth - 1, 0))[0])
    if r < 0.85:
        return (str(rng.choice(["AND", "OR"])), bool_tree(rng, depth - 1),
                bool_tree(rng, depth - 1))
    return ("NOT", bool_tree(rng, depth - 1))


def regs(e):
    """(type, registers) the compiler uses for `e` (one per node plus the
    promotion casts), to keep random trees within one kernel."""
    if e[0] == "col":
        return ("f64" if e[1] in ("f", "z") else "i64"), 1
    if e[0] in ("i64", "f64", "bool"):
        return ("i64" if e[0] == "bool" else e[0]), 1
    if e[0] == "cast":
        t, n = regs(e[1])
        return e[2], n + 1
    op = e[0]
    t1, n1 = regs(e[1])
    t2, n2 = regs(e[2]) if len(e) > 2 else (t1, 0)
    n = n1 + n2 + 1
    if op in ("EQ", "NE", "LT", "LE", "GT", "GE", "AND", "OR", "NOT"):
        if len(e) > 2 and t1 != t2:
            n += 1
        return "bool", n
    if "f64" in (t1, t2) or op in ("DIV", "SQRT", "LOG", "EXP"):
        n += (t1 != "f64") + (len(e) > 2 and t2 != "f64")
        return "f64", n
    return "i64", n


BOOL_OPS = ("EQ", "NE", "LT", "LE", "GT", "GE", "AND", "OR", "NOT")


def const_only(e):
    return e[0] != "col" and all(const_only(c) for c in e[1:] if isinstance(c, tuple))


def const_bool_inside(e):
    """A comparison or logic subtree over constants alone, which folding
    turns into a bool constant that the reference's evaluator then reads
    as raw bytes of another type; such trees are left out."""
    if e[0] in BOOL_OPS and const_only(e):
        return True
    return any(const_bool_inside(c) for c in e[1:] if isinstance(c, tuple) and c[0] != "col")


def trees(seed, count, kind):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        e = num_tree(rng, 3)[0] if kind == "num" else bool_tree(rng, 3)
        if e[0] in ("col", "i64", "f64"):
            continue
        if not const_only(e) and const_bool_inside(e):
            continue
        if regs(orc.expr_fold(e))[1] <= 16:
            out.append(e)
    return out


@pytest.mark.parametrize("kind", ["num", "bool"])
def test_expr_random_vs_oracle(eng, kind):
    rng = np.random.default_rng(5)
    cols = rand_cols(rng, 70001)
    t = table_of(eng, cols)
    rt = eng.put(t)
    for i, e in enumerate(trees(11 if kind == "num" else 12, 40, kind)):
        typ, ref = oracle(e, cols)
        got = run_vec(eng, rt if i % 2 else t, e)
        check(got, typ, ref, f"{kind} tree {i}: {e}")
    eng.release(rt)
    eng.release(t)


def test_expr_random_vs_reference(eng, ref_engine):
    """The same random trees through the reference library (oracle/_ref)."""
    rng = np.random.default_rng(6)
    cols = rand_cols(rng, 20000)
    t = table_of(eng, cols)
    rtab = table_of(ref_engine, cols)
    for i, e in enumerate(trees(21, 30, "num") + trees(22, 30, "bool")):
        got = run_vec(eng, t, e)
        ref = run_vec(ref_engine, rtab, e)
        check(got, oracle(e, cols)[0], ref, f"tree {i}: {e}")
    ref_engine.release(rtab)
    eng.release(t)


# ----------------------------------------------------- in FILTER and GROUP

@pytest.mark.parametrize("resident", [False, True])
def test_arith_filter_then_group(eng, resident):
    """FILTER(table, arithmetic predicate) leaves a lazy selection that the
    next GROUP consumes; the README shape (FILTER on every group input with
    one shared predicate) takes the same path."""
    rng = np.random.default_rng(9)
    n = 200003
    k = rng.integers(0, 300, n)
    v = rng.integers(-10 ** 6, 10 ** 6, n)
    f = rng.normal(size=n)
    tt = eng.table([("k", k), ("v", v), ("f", f)])
    t = eng.put(tt) if resident else tt
    pe = ("GT", ("ADD", ("MUL", ("col", "f"), ("f64", 2.0)), ("MOD", ("col", "v"), ("i64", 3))),
          ("f64", 0.5))
    _, keep = orc.expr_eval(X.bind(pe, {"k": k, "v": v, "f": f}))
    ref = orc.group([k[keep]], ["SUM", "COUNT", "MAX"], [v[keep], v[keep], f[keep]])
    # README shape
    g = eng.graph(t)
    p = X.build(g, pe)
    root = g.group([g.filter(g.scan("k"), p)],
                   [("SUM", g.filter(g.scan("v"), p)), ("COUNT", g.filter(g.scan("v"), p)),
                    ("MAX", g.filter(g.scan("f"), p))])
    out = eng.read_table(g.execute(root))
    g.free()
    compare_groups([out[0][3]], [c[3] for c in out[1:]], ref["keys"], ref["vals"],
                   [False, False, False], "readme arith")
    # a FILTER of the table first: the group sees the selection
    g = eng.graph(t)
    flt = g.filter(g.const_table(t), X.build(g, pe))
    sel_tbl = g.execute(flt)
    g.free()
    g = eng.graph(sel_tbl)
    root = g.group([g.scan("k")], [("SUM", g.scan("v")), ("COUNT", g.scan("v")),
                                   ("MAX", g.scan("f"))])
    out = eng.read_table(g.execute(root))
    g.free()
    compare_groups([out[0][3]], [c[3] for c in out[1:]], ref["keys"], ref["vals"],
                   [False, False, False], "filter arith")
    eng.release(sel_tbl)
    if resident:
        eng.release(t)
    eng.release(tt)


def test_expression_keys_and_sum(eng):
    rng = np.random.default_rng(10)
    n = 150001
    a = rng.integers(-2 ** 62, 2 ** 62, n)
    b = rng.integers(0, 1000, n).astype(np.int32)
    f = rng.normal(size=n)
    t = eng.table([("a", a), ("b", b), ("f", f)])
    cols = {"a": a, "b": b, "f": f}
    ke = ("MOD", ("col", "a"), ("i64", 97))
    ve = ("SUB", ("MUL", ("col", "a"), ("col", "b")), ("i64", 5))
    fe = ("MUL", ("col", "f"), ("cast", ("col", "b"), "f64"))
    _, kv = orc.expr_eval(X.bind(ke, cols))
    _, vv = orc.expr_eval(X.bind(ve, cols))
    _, fv = orc.expr_eval(X.bind(fe, cols))
    g = eng.graph(t)
    root = g.group([X.build(g, ke)], [("SUM", X.build(g, ve)), ("MIN", X.build(g, fe)),
                                      ("COUNT", g.scan("a"))])
    out = eng.read_table(g.execute(root))
    g.free()
    ref = orc.group([kv], ["SUM", "MIN", "COUNT"], [vv, fv, vv])
    compare_groups([out[0][3]], [c[3] for c in out[1:]], ref["keys"], ref["vals"],
                   [False, False, False], "expr keys")
    # SUM over an expression (no group): wrapping i64 sum, exact
    g = eng.graph(t)
    s = g.execute(g.sum(X.build(g, ve)))
    g.free()
    assert eng.atom_value(s) == int(vv.astype(np.uint64).sum(dtype=np.uint64).astype(np.int64))
    eng.release(s)
    eng.release(t)


def test_large_expression_splits(eng):
    """A tree past one kernel's 16 registers is evaluated in pieces with the
    same typing and results."""
    rng = np.random.default_rng(12)
    n = 50000
    cols = {"x": rng.integers(-100, 100, n), "y": rng.normal(size=n)}
    t = eng.table([("x", cols["x"]), ("y", cols["y"])])
    e = ("col", "x")
    for i in range(14):
        e = ("ADD" if i % 3 else "MUL", e, ("MOD", ("col", "x"), ("i64", i + 2)))
    e = ("ADD", e, ("DIV", ("col", "y"), ("f64", 3.0)))
    assert regs(e)[1] > 16
    typ, ref = orc.expr_eval(X.bind(e, cols))
    X.check_vec(run_vec(eng, t, e), typ, ref, "split")
    eng.release(t)


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("optimize", [False, True])
def test_affine_aggregates_golden(eng, resident, optimize):
    """SUM / AVG(col +- c) aggregate the column and add the bias to the
    results, as the reference does; an integral bias on an integer column
    keeps the integer SUM, a fractional one is an f64 expression."""
    from test_oracle import AFF_CASES
    d = hc("expr_aff_60k")
    t = hc_table(eng, d, resident)
    g = eng.graph(t)
    ins = []
    for op, col, c in AFF_CASES:
        s = g.scan(col)
        if c is None:
            ins.append((op, s))
        elif col == "v_f64" and c == 10:
            ins.append((op, g.add(g.const_i64(10), s)))
        elif c < 0:
            ins.append((op, g.sub(s, g.const_f64(-c) if isinstance(c, float) else g.const_i64(-c))))
        else:
            ins.append((op, g.add(s, g.const_f64(c) if isinstance(c, float) else g.const_i64(c))))
    root = g.group([g.scan("sym")], ins)
    if optimize:
        root = g.optimize(root)
    out = eng.read_table(g.execute(root))
    g.free()
    ref = read_tdt(os.path.join(GOLD, "expr_aff_60k.tdt"))
    assert [c[0] for c in out] == [c[0] for c in ref]
    assert [c[1] for c in out] == [c[1] for c in ref]
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]],
                   [ref[0][3].astype(np.int64)], [c[3] for c in ref[1:]],
                   [True, True, False, True, True, True, False], "affine")
    eng.release(t)


@pytest.mark.parametrize("resident", [False, True])
def test_scalar_aggregates_golden(eng, resident):
    """GROUP with no keys: one row, plain / affine / integer-linear inputs
    typed and summed as the reference does (tests/golden/expr_scalar_60k),
    and no row for an empty table."""
    d = hc("expr_scalar_60k")
    t = hc_table(eng, d, resident)
    g = eng.graph(t)
    vi, vf, key = g.scan("v_i64"), g.scan("v_f64"), g.scan("key")
    aggs = [("SUM", vi), ("COUNT", vi), ("AVG", vf), ("SUM", g.add(vf, g.const_f64(1.5))),
            ("SUM", g.add(g.mul(vi, g.const_f64(2.0)), key)),
            ("AVG", g.sub(vi, g.mul(g.const_i64(3), key))), ("SUM", g.mul(vi, g.const_f64(0.5)))]
    out = eng.read_table(g.execute(g.group([], aggs)))
    g.free()
    ref = read_tdt(os.path.join(GOLD, "expr_scalar_60k.tdt"))
    assert [c[0] for c in out] == [c[0] for c in ref]
    assert [c[1] for c in out] == [c[1] for c in ref]
    for i, (a, b) in enumerate(zip(out, ref)):
        if a[1] == 7:
            np.testing.assert_allclose(a[3], b[3], rtol=1e-12, err_msg=a[0])
        else:
            np.testing.assert_array_equal(a[3], b[3], err_msg=a[0])
    eng.release(t)
    e0 = eng.table([("a", np.zeros(0, np.int64))])
    g = eng.graph(e0)
    out = eng.read_table(g.execute(g.group([], [("SUM", g.scan("a")), ("COUNT", g.scan("a"))])))
    g.free()
    eng.release(e0)
    assert [len(c[3]) for c in out] == [0, 0]


def _jg_expr_aggs(g, full=True):
    price, qty, cat = g.scan("price"), g.scan("qty"), g.scan("cat")
    aggs = [("SUM", g.mul(price, g.cast(qty, X.TD_F64))),
            ("SUM", g.mul(price, g.const_f64(0.9)))]
    if full:
        aggs += [("AVG", g.add(qty, g.const_i64(1))),
                 ("SUM", g.mul(price, g.cast(cat, X.TD_F64)))]
    aggs.append(("MAX", g.sub(g.mul(qty, g.const_i64(2)), g.const_i64(1))))
    return aggs


@pytest.mark.parametrize("resident", [False, True])
def test_join_group_expression_inputs(eng, resident):
    """GROUP over a JOIN with expression inputs, against the reference's
    group of its join output (tests/golden/jg_expr_400k): in one DAG --
    fact-side expressions evaluated over the fact table and fused into the
    join -> group-by; an affine or a fact x dim input makes the whole group
    run on the materialised join -- and in the reference's two-call form."""
    import workload_dags as D
    argv = MANIFEST["jg_expr_400k"]["argv"]
    tf, td, _, _ = D.jg_tables(eng, int(argv[1]), int(argv[2]), int(argv[3]), int(argv[4]))
    if resident:
        rf, rd = eng.put(tf), eng.put(td)
        eng.release(tf)
        eng.release(td)
        tf, td = rf, rd
    ref = read_tdt(os.path.join(GOLD, "jg_expr_400k.tdt"))
    rk = [ref[0][3].astype(np.int64)]
    f64 = [True, True, True, True, False]

    def one_dag(full):
        g = eng.graph(tf)
        flt = g.filter(g.const_table(tf), g.gt(g.scan("qty"), g.const_i64(50)))
        g.join(flt, [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
        out = eng.read_table(g.execute(g.group([g.scan("cat")], _jg_expr_aggs(g, full))))
        g.free()
        return out, eng.last_plan()

    out, plan = one_dag(True)
    assert [c[0] for c in out] == [c[0] for c in ref]
    assert [c[1] for c in out] == [c[1] for c in ref]
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]], rk,
                   [c[3] for c in ref[1:]], f64, "join-group expr (materialised)")
    out, plan = one_dag(False)
    assert plan.startswith("join_group"), plan                 # fused
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]], rk,
                   [ref[1][3], ref[2][3], ref[5][3]], [True, True, False], "join-group expr (fused)")
    j = D.jg_join(eng, tf, td)
    g = eng.graph(j)
    out = eng.read_table(g.execute(g.group([g.scan("cat")], _jg_expr_aggs(g, True))))
    g.free()
    eng.release(j)
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]], rk,
                   [c[3] for c in ref[1:]], f64, "join then group expr")
    eng.release(tf)
    eng.release(td)


def test_arith_predicates_before_join_and_in_having(eng):
    """An arithmetic FILTER under a fused join -> group-by, and a HAVING
    with arithmetic over the group's columns, against the oracle."""
    import workload_dags as D
    tf, td, fact, dim = D.jg_tables(eng, 300000, 40000, 200, 13)
    g = eng.graph(tf)
    pe = ("GT", ("MUL", ("col", "qty"), ("col", "price")), ("f64", 2000.0))
    flt = g.filter(g.const_table(tf), X.build(g, pe))
    g.join(flt, [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
    grp = g.group([g.scan("cat")], [("SUM", g.scan("price")), ("COUNT", g.scan("qty"))])
    having = g.filter(grp, g.gt(g.div(g.scan("price_sum"), g.scan("qty_count")),
                                g.const_f64(40.0)))
    out = eng.read_table(g.execute(having))
    g.free()
    _, keep = orc.expr_eval(X.bind(pe, fact))
    cat_of = np.zeros(int(dim["pk"].max()) + 1, np.int64)
    cat_of[dim["pk"]] = dim["cat"]
    ok = keep & np.isin(fact["fk"], dim["pk"])
    r = orc.group([cat_of[fact["fk"][ok]]], ["SUM", "COUNT"], [fact["price"][ok], fact["qty"][ok]])
    pass_ = r["vals"][0] / r["vals"][1] > 40.0
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]],
                   [r["keys"][0][pass_]], [v[pass_] for v in r["vals"]], [True, False],
                   "arith filter -> join -> group -> arith having")
    eng.release(tf)
    eng.release(td)


def test_sort_by_expression(eng):
    """SORT keys that are expressions, evaluated over the sort's input (a
    table, and a group result), against the oracle's sort of the evaluated
    key."""
    rng = np.random.default_rng(21)
    n = 30011
    a = rng.integers(-1000, 1000, n)
    f = np.round(rng.normal(size=n), 3)
    t = eng.table([("a", a), ("f", f)])
    g = eng.graph(t)
    ke = ("ADD", ("MUL", ("col", "f"), ("f64", 3.0)), ("MOD", ("col", "a"), ("i64", 7)))
    root = g.sort(g.const_table(t), [X.build(g, ke), g.scan("a")], [True, False])
    out = eng.read_table(g.execute(root))
    g.free()
    _, kv = orc.expr_eval(X.bind(ke, {"a": a, "f": f}))
    p = orc.sort_perm([kv, a], [True, False])
    np.testing.assert_array_equal(out[0][3], a[p])
    np.testing.assert_array_equal(out[1][3], f[p])
    # ORDER BY an expression of a group result's columns, top 5
    k = a % 50
    t2 = eng.table([("k", k), ("a", a), ("f", f)])
    g = eng.graph(t2)
    grp = g.group([g.scan("k")], [("SUM", g.scan("f")), ("COUNT", g.scan("a"))])
    key = g.div(g.scan("f_sum"), g.scan("a_count"))
    out = eng.read_table(g.execute(g.head(g.sort(grp, [key], [True]), 5)))
    g.free()
    r = orc.group([k], ["SUM", "COUNT"], [f, a])
    mean = r["vals"][0] / r["vals"][1]
    top = orc.sort_perm([mean], [True])[:5]
    np.testing.assert_array_equal(out[0][3], r["keys"][0][top])
    eng.release(t)
    eng.release(t2)
