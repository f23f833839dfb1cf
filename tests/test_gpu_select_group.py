"""GPU parity: selection bitmaps, filters and group-by against the oracle and
the reference's golden tables.  Every call goes through the td.h C ABI of
libteide_amd.so."""
import ctypes as C
import os
import re
import zlib

import numpy as np
import pytest

import orc
import workload_dags as D
from parity import assert_f64_close, assert_var_close, compare_groups, key_order
from tdtio import read_tdt
from teide_amd import workloads as W
from teide_amd.td import TD_SEL

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sel_payload(e, s):
    """Host bytes of a TD_SEL (device-resident or not)."""
    h = e.hdr(s)
    owned = None
    if h.mmod == 3:
        owned = s = e.check(e.L.td_dev_get(s))
    n = e.hdr(s).val
    nbytes = orc.lib().orc_sel_payload_bytes(n)
    out = np.frombuffer(C.string_at(s + 32, nbytes), np.uint8).copy()
    if owned:
        e.release(owned)
    return out


# ------------------------------------------------------------- selection --

def test_sel_kats(eng):
    e = eng
    b = e.vec(np.array([1, 0, 1, 0, 1], np.uint8), 1)
    s = e.check(e.L.td_sel_from_pred(b))
    assert e.hdr(s).type == TD_SEL and orc.sel_parts(sel_payload(e, s), 5)[0] == 3
    a = e.check(e.L.td_sel_from_pred(e.vec(np.array([1, 1, 0, 0, 1], np.uint8), 1)))
    ab = e.check(e.L.td_sel_and(a, s))
    assert orc.sel_parts(sel_payload(e, ab), 5)[0] == 2


@pytest.mark.parametrize("n", [1, 64, 1000, 1024, 1025, 65537, 3000001])
def test_sel_from_pred_bit_exact(eng, n):
    rng = np.random.default_rng(n)
    pred = (rng.random(n) < 0.4).astype(np.uint8)
    if n > 5000:
        pred[1024:2048] = 1
        pred[2048:3072] = 0
    s = eng.check(eng.L.td_sel_from_pred(eng.vec(pred, 1)))
    np.testing.assert_array_equal(sel_payload(eng, s), orc.sel_from_bool(pred))
    # AND with a second selection, and recompute after a round trip
    p2 = (rng.random(n) < 0.7).astype(np.uint8)
    s2 = eng.check(eng.L.td_sel_from_pred(eng.vec(p2, 1)))
    s12 = eng.check(eng.L.td_sel_and(s, s2))
    np.testing.assert_array_equal(sel_payload(eng, s12),
                                  orc.sel_and(orc.sel_from_bool(pred), orc.sel_from_bool(p2), n))
    for p in (s, s2, s12):
        eng.release(p)


def test_filter_integration_kat(eng):
    # test_sel.c /filter_integration: SUM(FILTER(x, x > 25)) = 120
    t = eng.table([("x", np.array([10, 20, 30, 40, 50], np.int64))])
    g = eng.graph(t)
    x = g.scan("x")
    r = g.execute(g.sum(g.filter(x, g.gt(x, g.const_i64(25)))))
    assert eng.atom_value(r) == 120
    g.free()


@pytest.mark.parametrize("op,c", [("eq", 3), ("ne", 3), ("lt", 3), ("le", 3), ("gt", 3),
                                  ("ge", 3), ("gt", 2.5), ("eq", 3.0)])
def test_filter_table_compaction(eng, op, c):
    rng = np.random.default_rng(11)
    n = 70001
    a = rng.integers(0, 7, n)
    f = rng.normal(size=n)
    t = eng.table([("a", a), ("f", f), ("s", a.astype(np.int32))])
    g = eng.graph(t)
    const = g.const_f64(c) if isinstance(c, float) else g.const_i64(c)
    pred = getattr(g, op)(g.scan("a"), const)
    r = g.execute(g.filter(g.const_table(t), pred))
    got = eng.read_table(r)
    m = orc.cmp_const(a, op.upper(), c).astype(bool)
    assert [x[0] for x in got] == ["a", "f", "s"]
    np.testing.assert_array_equal(got[0][3], a[m])
    np.testing.assert_array_equal(got[1][3], f[m])
    np.testing.assert_array_equal(got[2][3], a[m].astype(np.int32))
    g.free()


def test_filter_and_or_not(eng):
    rng = np.random.default_rng(12)
    n = 100000
    a, b = rng.integers(0, 100, n), rng.normal(size=n)
    t = eng.table([("a", a), ("b", b)])
    g = eng.graph(t)
    pred = g.or_(g.and_(g.gt(g.scan("a"), g.const_i64(40)), g.lt(g.scan("b"), g.const_f64(0.5))),
                 g.not_(g.ne(g.scan("a"), g.const_i64(7))))
    r = g.execute(g.optimize(g.filter(g.const_table(t), pred)))
    m = ((a > 40) & (b < 0.5)) | (a == 7)
    np.testing.assert_array_equal(eng.read_table(r)[0][3], a[m])
    g.free()


# ---------------------------------------------------------- filter+group --

def check_fg(eng, out, fixture):
    got = eng.read_table(out)
    ref = read_tdt(os.path.join(GOLD, fixture))
    assert [c[0] for c in got] == [c[0] for c in ref]
    assert [c[1] for c in got] == [c[1] for c in ref]
    # direct-array path: ascending keys, like the reference's DA path
    np.testing.assert_array_equal(got[0][3], ref[0][3])
    np.testing.assert_array_equal(got[2][3], ref[2][3])
    assert_f64_close(got[1][3], ref[1][3], what="amount_sum")


@pytest.mark.parametrize("resident", [False, True])
def test_fg_reference_form_golden(eng, resident):
    t, _ = D.fg_table(eng, 1000000, 42)
    if resident:
        t = eng.put(t)
    out = D.fg_reference_form(eng, t)
    assert bool(eng.L.td_dev_is_resident(out)) == resident
    check_fg(eng, out, "fg_1m.tdt")


def test_fg_readme_and_selection_forms(eng):
    t, d = D.fg_table(eng, 100000, 7)
    check_fg(eng, D.fg_readme_form(eng, t), "fg_100k_s7.tdt")
    sel = eng.check(eng.L.td_sel_from_pred(eng.vec((d["flag"] == 0).astype(np.uint8), 1)))
    check_fg(eng, D.run_group_fg(eng, t, sel), "fg_100k_s7.tdt")


# ------------------------------------------------------------- group-by --

def key64(a):
    """Key column as the oracle reports it: f64 raw bits, integers widened."""
    return a.view(np.int64) if a.dtype == np.float64 else a.astype(np.int64)


VAR_OPS = {"VAR": (False, False), "VAR_POP": (False, True), "STDDEV": (True, False),
           "STDDEV_POP": (True, True)}       # op -> (compared as square, population)


def var_scales(keys, x, mask, ref_keys, pop):
    """Per group (in ref_keys order) sum(x^2) / (n or n-1): the magnitude the
    variance formula cancels at."""
    K = np.stack([key64(k) for k in keys], 1)
    x = np.asarray(x, np.float64)
    if mask is not None:
        K, x = K[mask], x[mask]
    uk, inv = np.unique(K, axis=0, return_inverse=True)
    inv = inv.ravel()
    sq = np.bincount(inv, weights=x * x, minlength=len(uk))
    cnt = np.bincount(inv, minlength=len(uk)).astype(np.float64)
    denom = cnt if pop else np.maximum(cnt - 1, 1)
    pos = {r: i for i, r in enumerate(map(tuple, uk.tolist()))}
    idx = np.array([pos[r] for r in map(tuple, np.stack(ref_keys, 1).tolist())], np.int64)
    return sq[idx] / denom[idx]


AGG_SETS = [
    (["SUM", "COUNT", "AVG"], ["vi", "vi", "vf"]),
    (["MIN", "MAX", "MIN", "MAX"], ["vi", "vi", "vf", "vf"]),
    (["SUM", "FIRST", "LAST", "VAR", "STDDEV_POP"], ["vf", "vi", "vf", "vf", "vi"]),
]

KEYSETS = {
    "i64_small": lambda r, n: [r.integers(-3, 9, n)],                  # LDS direct array
    "i64_medium": lambda r, n: [r.integers(0, 200000, n)],             # global direct array
    "i64_scattered": lambda r, n: [r.integers(-2 ** 62, 2 ** 62, n) // 1000 * 1000
                                   + r.integers(0, 3, n)],            # hash
    "i32": lambda r, n: [r.integers(-50, 50, n).astype(np.int32)],
    "f64": lambda r, n: [np.round(r.normal(size=n), 1)],                 # hash on raw bits
    "two_keys": lambda r, n: [r.integers(0, 30, n), r.integers(-4, 4, n).astype(np.int16)],
    "sym_w32": lambda r, n: [r.integers(0, 5000, n).astype(np.uint32)],
    "sym_w8": lambda r, n: [r.integers(0, 256, n).astype(np.uint8)],
    "sym_w16": lambda r, n: [r.integers(0, 65536, n).astype(np.uint16)],
    "sym_w64": lambda r, n: [r.integers(0, 3000, n) * 7919 + 2 ** 40],
    # runs of one key (sorted or partitioned data): whole waves fold into one slot
    "i64_sorted": lambda r, n: [np.sort(r.integers(-20, 40, n))],
    # key tuples wider than one 64-bit word: numbered through a hash of their
    # raw key words, then grouped by that number (-0.0 and NaN are raw bits)
    "f64_i64": lambda r, n: [np.where(np.arange(n) % 997 == 0, np.nan,
                                      np.round(r.normal(size=n), 1)), r.integers(-3, 3, n)],
    "two_scattered": lambda r, n: [r.integers(-2 ** 62, 2 ** 62, 5000)[r.integers(0, 5000, n)],
                                   r.integers(-2 ** 61, 2 ** 61, 7)[r.integers(0, 7, n)]],
    "i32_f64_i16": lambda r, n: [r.integers(0, 40, n).astype(np.int32),
                                 r.integers(-2, 3, n) * 0.5,
                                 r.integers(-5, 5, n).astype(np.int16)],
}
TUPLE_KEYSETS = ("f64_i64", "two_scattered", "i32_f64_i16")


# group path 3: the radix-partitioned direct array, forced with small leaf
# tables so that even these sizes take two or three scatter passes; group
# path 4: the hash-partitioned hash group-by with small LDS tables (keys
# that do not fit a part's table fold into the global table)
PDA_KEYSETS = ("i64_small", "i64_medium", "i32", "two_keys", "sym_w32", "sym_w8", "sym_w16",
               "zipf")
# SYM key columns: (td.h type tag, index width in attrs bits 0-1)
SYM_KEYS = {"sym_w8": (20, 0), "sym_w16": (20, 1), "sym_w32": (20, 2), "sym_w64": (20, 3)}


@pytest.fixture
def group_path(eng):
    def setp(p):
        eng.force_paths(0, p, 0)
    yield setp
    eng.force_paths(0, 0, 0)


@pytest.mark.parametrize("kset", list(KEYSETS) + ["zipf"])
@pytest.mark.parametrize("aggi", range(len(AGG_SETS)))
@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("gpath", [0, 3, 4, 5, 6])
def test_group_vs_oracle(eng, group_path, kset, aggi, masked, gpath):
    rng = np.random.default_rng(zlib.crc32(f"{kset}/{aggi}/{masked}".encode()))
    n = 200003
    if kset == "zipf":
        keys = [np.minimum(rng.zipf(1.2, n), 150000).astype(np.int64)]   # heavy skew
    else:
        keys = KEYSETS[kset](rng, n)
    group_path(gpath)
    vi = rng.integers(-10 ** 9, 10 ** 9, n)
    vf = rng.normal(size=n) * 100
    ktag = SYM_KEYS.get(kset)
    cols = [(f"k{i}", k) + (ktag or ()) for i, k in enumerate(keys)] + [("vi", vi), ("vf", vf)]
    t = eng.table(cols)
    g = eng.graph(t)
    ops, ins = AGG_SETS[aggi]
    mask = mb = None
    if masked:
        mb = rng.random(n) < 0.3
        sel = eng.check(eng.L.td_sel_from_pred(eng.vec(mb.astype(np.uint8), 1)))
        g.set_selection(sel)
        eng.release(sel)
        bits = np.packbits(mb, bitorder="little")
        mask = np.pad(bits, (0, (-len(bits)) % 8)).view(np.uint64)
    root = g.group([g.scan(f"k{i}") for i in range(len(keys))],
                   [(o, g.scan(c)) for o, c in zip(ops, ins)])
    out = eng.read_table(g.execute(root))
    plan = eng.last_plan()
    g.free()
    if kset in TUPLE_KEYSETS:
        assert re.search(r"path=2\d", plan), plan   # tuples numbered, then grouped
    if gpath == 3 and kset in PDA_KEYSETS:
        assert "path=4" in plan, plan
    if gpath in (4, 6):
        assert re.search(r"path=2?5", plan), plan  # hash-partitioned hash
    if gpath == 6 and int(re.search(r"groups=(\d+)", plan).group(1)) > 4096:
        # the first try's 1,024-entry table and list overflowed and regrew
        assert int(re.search(r"tries=(\d+)", plan).group(1)) > 1, plan
    if gpath == 5 and kset == "zipf":
        assert "path=6" in plan, plan              # hot keys folded first
    src = {"vi": vi, "vf": vf}
    kt = [ktag] if ktag else None
    ref = orc.group(keys, ops, [src[c] for c in ins], mask_bits=mask, key_types=kt)
    nk = len(keys)
    f64 = ["skip" if ops[a] in VAR_OPS else
           ref["out_types"][a] == 7 and ops[a] in ("SUM", "AVG") for a in range(len(ops))]
    assert [c[1] for c in out[nk:]] == ref["out_types"]
    gk = [key64(c[3]) for c in out[:nk]]
    compare_groups(gk, [c[3] for c in out[nk:]], ref["keys"], ref["vals"], f64,
                   f"{kset}/{aggi}/{masked}", exact=ref["comp"], sum_abs=ref["sum_abs"])
    go, ro = key_order(gk), key_order(ref["keys"])
    for a, op in enumerate(ops):
        if op in VAR_OPS:
            is_sqrt, pop = VAR_OPS[op]
            scale = var_scales(keys, src[ins[a]], mb if masked else None, ref["keys"], pop)
            assert_var_close(out[nk + a][3][go], ref["vals"][a][ro], scale[ro], is_sqrt,
                             f"{kset}/{aggi}/{masked}: {op}")
    if ref["order"] == 1 and "f64" not in kset and gpath < 4:
        # direct-array eligible in the reference: same ascending key order
        np.testing.assert_array_equal(key64(out[0][3]), ref["keys"][0])


def test_group_empty_and_all_filtered(eng):
    t = eng.table([("k", np.zeros(0, np.int64)), ("v", np.zeros(0, np.float64))])
    g = eng.graph(t)
    out = eng.read_table(g.execute(g.group([g.scan("k")], [("SUM", g.scan("v"))])))
    assert len(out[0][3]) == 0 and len(out) == 2
    g.free()
    k = np.arange(5000) % 7
    t = eng.table([("k", k), ("v", k * 1.0)])
    g = eng.graph(t)
    flt = g.filter(g.const_table(t), g.gt(g.scan("k"), g.const_i64(100)))
    r = eng.read_table(g.execute(flt))
    assert len(r[0][3]) == 0
    g.free()


def test_group_kats(eng):
    t = eng.table([("id1", np.array([1, 1, 2, 2, 3, 3, 1, 2, 3, 1], np.int64)),
                   ("v1", np.array([10, 20, 30, 40, 50, 60, 70, 80, 90, 100], np.int64))])
    g = eng.graph(t)
    out = eng.read_table(g.execute(g.group([g.scan("id1")], [("SUM", g.scan("v1"))])))
    assert dict(zip(out[0][3].tolist(), out[1][3].tolist())) == {1: 200, 2: 150, 3: 200}
    g.free()
    ids = np.arange(100000, dtype=np.int64) % 4
    t = eng.table([("id", ids), ("v", ids + 1)])
    g = eng.graph(t)
    out = eng.read_table(g.execute(g.group([g.scan("id")], [("SUM", g.scan("v"))])))
    assert out[1][3].tolist() == [25000, 50000, 75000, 100000]
    g.free()
    x = np.arange(100000, dtype=np.float64) - 50000
    t = eng.table([("x", x)])
    g = eng.graph(t)
    assert eng.atom_value(g.execute(g.min_(g.scan("x")))) == -50000.0
    g.free()
    g = eng.graph(t)
    assert eng.atom_value(g.execute(g.max_(g.scan("x")))) == 49999.0
    g.free()


@pytest.mark.parametrize("by", ["key", "sym"])
def test_hc_golden(eng, by):
    t, _ = D.hc_table(eng, 60000, 4096, 2048)
    out = eng.read_table(D.hc_group(eng, t, by))
    ref = read_tdt(os.path.join(GOLD, f"hc_60k_{by}.tdt"))
    assert [c[0] for c in out] == [c[0] for c in ref]
    assert [c[1] for c in out] == [c[1] for c in ref]
    f64 = [False, False, False, True, False, False, False]
    compare_groups([out[0][3]], [c[3] for c in out[1:]], [ref[0][3]], [c[3] for c in ref[1:]],
                   f64, by)
