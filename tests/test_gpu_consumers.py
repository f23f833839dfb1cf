"""GPU parity of the consumers of a group result: HAVING (FILTER over GROUP),
SORT and HEAD(SORT) (top-N) and HEAD(GROUP), against goldens written by the
reference (tests/golden/having_60k*, sort_8k*) and against the oracle's
sort on random keys (integers of every width, f64 with NaN / signed zeros /
infinities, symbols by their strings, several keys with ties)."""
import json
import os

import numpy as np
import pytest

import orc
import workload_dags as D
from parity import compare_groups
from tdtio import read_tdt

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))


def having_dag(eng, t, minc, maxs, top=0):
    g = eng.graph(t)
    grp = g.group([g.scan("sym")], [("SUM", g.scan("v_f64")), ("COUNT", g.scan("v_i64")),
                                    ("MAX", g.scan("v_i64"))])
    pred = g.and_(g.ge(g.scan("v_i64_count"), g.const_i64(minc)),
                  g.lt(g.scan("v_f64_sum"), g.const_f64(maxs)))
    root = g.filter(grp, pred)
    if top:
        root = g.head(g.sort(root, [g.scan("v_f64_sum")], [True]), top)
    return g, root


@pytest.mark.parametrize("resident", [False, True])
def test_having_golden(eng, resident):
    argv = MANIFEST["having_60k"]["argv"]
    t, _ = D.hc_table(eng, int(argv[1]), int(argv[2]), int(argv[3]), int(argv[4]))
    if resident:
        t2 = eng.put(t)
        eng.release(t)
        t = t2
    minc, maxs, top = int(argv[5]), float(argv[6]), int(argv[7])
    g, root = having_dag(eng, t, minc, maxs)
    out = eng.read_table(g.execute(root))
    g.free()
    ref = read_tdt(os.path.join(GOLD, "having_60k.tdt"))
    assert [c[0] for c in out] == [c[0] for c in ref]
    assert [c[1] for c in out] == [c[1] for c in ref]
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]],
                   [ref[0][3].astype(np.int64)], [c[3] for c in ref[1:]], [True, False, False],
                   "having")
    np.testing.assert_array_equal(out[0][3], ref[0][3])       # ascending group order
    g, root = having_dag(eng, t, minc, maxs, top)
    out = eng.read_table(g.execute(root))
    g.free()
    reft = read_tdt(os.path.join(GOLD, "having_60k_top.tdt"))
    np.testing.assert_array_equal(out[0][3], reft[0][3])      # the same top groups, in order
    eng.release(t)


@pytest.mark.parametrize("resident", [False, True])
def test_sort_golden(eng, resident):
    argv = MANIFEST["sort_8k"]["argv"]
    t, _ = D.hc_table(eng, int(argv[1]), int(argv[2]), int(argv[3]), int(argv[4]))
    if resident:
        t2 = eng.put(t)
        eng.release(t)
        t = t2
    g = eng.graph(t)
    root = g.sort(g.const_table(t), [g.scan("key"), g.scan("v_i64")], [False, True])
    out = eng.read_table(g.execute(root))
    g.free()
    ref = read_tdt(os.path.join(GOLD, "sort_8k.tdt"))
    assert [c[0] for c in out] == [c[0] for c in ref]
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(np.asarray(a[3]).view(np.uint8),
                                      np.asarray(b[3]).view(np.uint8), err_msg=a[0])
    g = eng.graph(t)
    root = g.head(g.sort(g.const_table(t), [g.scan("v_f64")], [True]), int(argv[5]))
    out = eng.read_table(g.execute(root))
    g.free()
    reft = read_tdt(os.path.join(GOLD, "sort_8k_top.tdt"))
    for a, b in zip(out, reft):
        np.testing.assert_array_equal(a[3], b[3], err_msg=a[0])
    eng.release(t)


def _mixed_f64(rng, n):
    x = np.round(rng.normal(size=n) * 4, 1)                   # ties
    special = np.array([np.nan, -0.0, 0.0, np.inf, -np.inf, np.nan])
    idx = rng.integers(0, n, 60)
    x[idx] = special[rng.integers(0, len(special), 60)]
    return x


KEYS = {
    "i64": lambda r, n: r.integers(-50, 50, n),
    "i32": lambda r, n: r.integers(-2 ** 31, 2 ** 31 - 1, n).astype(np.int32),
    "i16": lambda r, n: r.integers(-300, 300, n).astype(np.int16),
    "u8": lambda r, n: r.integers(0, 7, n).astype(np.uint8),
    "f64": _mixed_f64,
}


@pytest.mark.parametrize("combo", [("i64",), ("f64",), ("i32", "f64"), ("u8", "i16", "i64"),
                                   ("f64", "i64")])
@pytest.mark.parametrize("dirs", [0, 1, 2])
@pytest.mark.parametrize("limit", [None, 17])
def test_sort_vs_oracle(eng, combo, dirs, limit):
    rng = np.random.default_rng(len(combo) * 100 + dirs * 10 + (limit or 0))
    n = 50003
    cols = [KEYS[k](rng, n) for k in combo]
    pay = np.arange(n, dtype=np.int64)
    descs = [bool((dirs >> i) & 1) if dirs < 2 else bool(i % 2 == 0) for i in range(len(combo))]
    nfs = [not d for d in descs] if dirs == 1 else descs
    t = eng.table([(f"c{i}", c) for i, c in enumerate(cols)] + [("row", pay)])
    g = eng.graph(t)
    root = g.sort(g.const_table(t), [g.scan(f"c{i}") for i in range(len(cols))], descs, nfs)
    if limit:
        root = g.head(root, limit)
    out = eng.read_table(g.execute(root))
    g.free()
    eng.release(t)
    p = orc.sort_perm(cols, descs, nfs)
    if limit:
        p = p[:limit]
    np.testing.assert_array_equal(out[-1][3], pay[p], err_msg=f"{combo} {descs} {nfs}")
    for i, c in enumerate(cols):
        np.testing.assert_array_equal(np.asarray(out[i][3]).view(np.uint8),
                                      np.asarray(c[p]).view(np.uint8))


def test_sort_symbols_by_string(eng):
    rng = np.random.default_rng(3)
    words = [f"w{rng.integers(0, 10 ** 6):06d}".encode() + b"x" * int(rng.integers(0, 3))
             for _ in range(300)] + [b"", b"a", b"ab", b"abc", b"b"]
    ids = np.array([eng.sym(w.decode()) for w in words], np.int64)
    n = 40000
    pick = rng.integers(0, len(words), n)
    sym = ids[pick].astype(np.uint32)
    val = rng.integers(0, 5, n)
    t = eng.table([("s", sym, 20, 2), ("v", val)])
    g = eng.graph(t)
    root = g.sort(g.const_table(t), [g.scan("s"), g.scan("v")], [True, False])
    out = eng.read_table(g.execute(root))
    g.free()
    eng.release(t)
    n_ids = int(sym.max()) + 1
    strings = []
    for i in range(n_ids):
        s = eng.sym_name(i)
        strings.append(s.encode() if s is not None else b"")
    rank = orc.sym_rank(strings)
    p = orc.sort_perm([sym, val], [True, False], ranks=[rank, None])
    np.testing.assert_array_equal(out[0][3], sym[p])
    np.testing.assert_array_equal(out[1][3], val[p])


def test_head_of_group(eng):
    rng = np.random.default_rng(8)
    k = rng.integers(0, 5000, 300000)
    v = rng.normal(size=300000)
    t = eng.table([("k", k), ("v", v)])
    g = eng.graph(t)
    root = g.head(g.group([g.scan("k")], [("SUM", g.scan("v")), ("COUNT", g.scan("v"))]), 7)
    out = eng.read_table(g.execute(root))
    g.free()
    eng.release(t)
    ref = orc.group([k], ["SUM", "COUNT"], [v, v])
    np.testing.assert_array_equal(out[0][3], ref["keys"][0][:7])    # ascending keys
    np.testing.assert_array_equal(out[2][3], ref["vals"][1][:7])
    np.testing.assert_allclose(out[1][3], ref["vals"][0][:7], rtol=1e-12, atol=1e-12)
