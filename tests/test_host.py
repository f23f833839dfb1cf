"""CPU tests of the host side: the td.h ABI layout, the exported C ABI, and
the host API (vectors, tables, symbols, graph building) that needs no GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from teide_amd.td import LIB_PATH, TD_I64, Engine, TdError

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")
HEADERS = [os.path.join(REPO, "include", "teide", h) for h in ("td.h", "td_dev.h", "tdd.h")]


@pytest.fixture(scope="module")
def host():
    return Engine(LIB_PATH, device=True)


def declared_functions(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    src = re.sub(r"//[^\n]*", " ", src)
    src = re.sub(r"#ifdef __cplusplus.*?#endif", " ", src, flags=re.S)
    src = re.sub(r"^\s*#.*$", " ", src, flags=re.M)
    # keep file-scope text only; a brace block ends a statement
    top, depth = [], 0
    for ch in src:
        if ch == "{":
            depth += 1
            if depth == 1:
                top.append(";")
        elif ch == "}":
            depth -= 1
        elif depth == 0:
            top.append(ch)
    names = set()
    for stmt in "".join(top).split(";"):
        stmt = stmt.strip()
        if "static" in stmt or "typedef" in stmt or not stmt.endswith(")"):
            continue
        m = re.search(r"\b(td\w*)\s*\(", stmt)
        if m:
            names.add(m.group(1))
    return sorted(names)


def test_abi_layout_matches_reference(tmp_path):
    exe = tmp_path / "abi_probe"
    subprocess.run(["gcc", "-std=c17", "-I", os.path.join(REPO, "include"),
                    os.path.join(GOLD, "abi_probe.c"), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert got == open(os.path.join(GOLD, "abi_layout.txt")).read()


@pytest.mark.parametrize("header", HEADERS, ids=os.path.basename)
def test_library_exports_every_declared_function(header):
    lib = C.CDLL(LIB_PATH)
    names = declared_functions(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{os.path.basename(header)}: not exported: {missing}"


def test_tables_symbols_and_vectors(host):
    e = host
    t = e.table([("a", np.arange(10, dtype=np.int64)), ("b", np.linspace(0, 1, 10)),
                 ("s", np.arange(10, dtype=np.uint32) % 3, 20, 2)])
    assert e.L.td_table_ncols(t) == 3 and e.L.td_table_nrows(t) == 10
    cols = e.read_table(t)
    assert [c[0] for c in cols] == ["a", "b", "s"]
    np.testing.assert_array_equal(cols[0][3], np.arange(10))
    assert cols[2][3].dtype == np.uint32
    assert e.sym("a") == e.L.td_sym_find(b"a", 1)
    assert e.sym_name(e.sym("zeta")) == "zeta"
    e.release(t)


def test_graph_building_and_optimize(host):
    e = host
    t = e.table([("x", np.arange(5, dtype=np.int64))])
    g = e.graph(t)
    x = g.scan("x")
    # like graph.c:247-270, a SCAN of an unknown column still builds a node
    assert x and g.scan("missing")
    p = g.and_(g.gt(x, g.const_i64(1)), g.lt(x, g.const_i64(4)))
    root = g.optimize(g.filter(g.const_table(t), p))
    assert root
    g.free()
    e.release(t)


def test_execute_without_gpu_fails_loudly(host):
    e = host
    if e.L.td_dev_available() == 1:
        pytest.skip("a GPU is present")
    t = e.table([("x", np.arange(5, dtype=np.int64))])
    g = e.graph(t)
    with pytest.raises(TdError) as ei:
        g.execute(g.sum(g.scan("x")))
    assert ei.value.code == 8          # TD_ERR_IO: no device, no CPU fallback
    g.free()
    e.release(t)


def test_reference_engine_kat(ref_engine):
    """The reference build answers the test_graph.c group_sum KAT (sanity of
    the ctypes mirror used for both engines)."""
    e = ref_engine
    t = e.table([("id1", np.array([1, 1, 2, 2, 3, 3, 1, 2, 3, 1], np.int64)),
                 ("v1", np.array([10, 20, 30, 40, 50, 60, 70, 80, 90, 100], np.int64))])
    g = e.graph(t)
    out = e.read_table(g.execute(g.group([g.scan("id1")], [("SUM", g.scan("v1"))])))
    assert dict(zip(out[0][3].tolist(), out[1][3].tolist())) == {1: 200, 2: 150, 3: 200}
    g.free()


def plan_tree(e, g, root):
    """td_graph_dump's text as nested (op, children) tuples."""
    import tempfile
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    e.L.td_graph_dump.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    e.L.td_graph_dump.restype = None
    with tempfile.NamedTemporaryFile("r", suffix=".txt") as f:
        fp = libc.fopen(f.name.encode(), b"w")
        e.L.td_graph_dump(g.g, root, fp)
        libc.fclose(fp)
        lines = open(f.name).read().splitlines()
    stack, top = [], []
    for ln in lines:
        depth = (len(ln) - len(ln.lstrip(" "))) // 2
        node = (ln.strip().split("#")[0] + ("" if " " not in ln.strip() else
                                           " " + ln.strip().split(" ", 1)[1]), [])
        del stack[depth:]
        (stack[-1][1] if stack else top).append(node)
        stack.append(node)
    return top[0]


def filter_chain(node):
    """Predicates of a stack of FILTERs, outermost first, and what is below."""
    preds = []
    while node[0] == "FILTER":
        preds.append(node[1][1])
        node = node[1][0]
    return preds, node


def pred_sig(p):
    return p[0] + "(" + ",".join(c[0] for c in p[1]) + ")"


def test_optimize_splits_ands_and_orders_by_cost(host):
    """FILTER(AND(AND(a, b), c)) becomes a chain of three filters ordered by
    filter_cost (opt.c:1258-1459): the col-col i32 LT (cost 8) outermost,
    the i64 GT against a constant (5), the U8 EQ against a constant (0)
    innermost, i.e. evaluated first."""
    e = host
    t = e.table([("qty", np.arange(8, dtype=np.int64)), ("flag", np.ones(8, np.uint8)),
                 ("a", np.arange(8, dtype=np.int32)), ("b", np.full(8, 4, np.int32))])
    g = e.graph(t)
    p = g.and_(g.and_(g.gt(g.scan("qty"), g.const_i64(2)), g.eq(g.scan("flag"), g.const_i64(1))),
               g.lt(g.scan("a"), g.scan("b")))
    root = g.optimize(g.filter(g.const_table(t), p))
    preds, below = filter_chain(plan_tree(e, g, root))
    assert [pred_sig(x) for x in preds] == ["LT(SCAN a,SCAN b)", "GT(SCAN qty,CONST)",
                                             "EQ(SCAN flag,CONST)"]
    assert below[0] == "CONST"
    g.free()
    e.release(t)


def test_optimize_splits_filters_below_the_root(host):
    """A FILTER(AND) feeding a JOIN is split too (every FILTER node, not only
    the root, opt.c:1360-1380), and its chain is cost-ordered."""
    e = host
    f = e.table([("fk", np.arange(8, dtype=np.int64)), ("qty", np.arange(8, dtype=np.int64)),
                 ("d", np.arange(8, dtype=np.int32))])
    d = e.table([("pk", np.arange(8, dtype=np.int64)), ("c", np.arange(8, dtype=np.int32))])
    g = e.graph(f)
    p = g.and_(g.ge(g.scan("qty"), g.const_i64(3)), g.ne(g.scan("d"), g.const_i64(5)))
    flt = g.filter(g.const_table(f), p)
    root = g.join(flt, [g.scan("fk")], g.const_table(d), [g.scan("pk")], 0)
    tree = plan_tree(e, g, g.optimize(root))
    assert tree[0] == "JOIN"
    preds, below = filter_chain(tree[1][0])
    # GE on i64 vs const: 0 + 3 + 2 = 5; NE on i32 vs const: 0 + 2 + 0 = 2
    assert [pred_sig(x) for x in preds] == ["GE(SCAN qty,CONST)", "NE(SCAN d,CONST)"]
    g.free()
    e.release(f)
    e.release(d)


@pytest.mark.parametrize("keep", [True, False])
def test_optimize_folds_constant_filters(host, keep):
    """FILTER(x, const) -> MATERIALIZE(x) when the constant is true, HEAD(x, 0)
    when false (fold_filter_const_predicate, opt.c:461-494), also for a
    predicate that folds to a constant first."""
    e = host
    t = e.table([("x", np.arange(5, dtype=np.int64))])
    g = e.graph(t)
    pred = g.gt(g.const_i64(3 if keep else 1), g.const_i64(2))      # folds to a constant
    tree = plan_tree(e, g, g.optimize(g.filter(g.const_table(t), pred)))
    assert tree[0] == ("MATERIALIZE" if keep else "HEAD")
    assert [c[0] for c in tree[1]] == ["CONST"]
    g.free()
    e.release(t)
