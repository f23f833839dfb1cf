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
