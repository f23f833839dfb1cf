"""GPU parity: equi-joins (direct-map and radix-partitioned paths) and the
fused filter -> join -> group pipeline, against the oracle and the
reference's golden tables.  Every call goes through the td.h C ABI."""
import json
import os

import numpy as np
import pytest

import orc
import workload_dags as D
from parity import compare_groups, row_multiset_equal
from tdtio import canonical_digest, read_tdt
from teide_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))

# (join path, radix bits): 0 = planner's choice, 1 = direct map, 2 = radix
PATHS = [(0, 0), (1, 0), (2, 0), (2, 3), (2, 9)]


@pytest.fixture
def forced(eng):
    def setp(path, bits=0, group=0):
        eng.force_paths(path, group, bits)
    yield setp
    eng.force_paths(0, 0, 0)


def run_join(eng, lcols, rcols, lkeys, rkeys, jt, resident=False):
    lt, rt = eng.table(lcols), eng.table(rcols)
    if resident:
        lt2, rt2 = eng.put(lt), eng.put(rt)
        eng.release(lt)
        eng.release(rt)
        lt, rt = lt2, rt2
    g = eng.graph(lt)
    root = g.join(g.const_table(lt), [g.scan(k) for k in lkeys], g.const_table(rt),
                  [g.scan(k) for k in rkeys], jt)
    out = g.execute(g.optimize(root))
    assert bool(eng.L.td_dev_is_resident(out)) == resident
    res = eng.read_table(out)
    g.free()
    eng.release(out)
    eng.release(lt)
    eng.release(rt)
    return res


def join_case(argv):
    nl, nr, krange, jt, seed = (int(a) for a in argv[1:6])
    lk = W.i64_uniform(nl, 0, krange - 1, seed, 601)
    lv = W.i64_uniform(nl, -1000000, 1000000, seed, 602)
    rk = W.i64_uniform(nr, 0, krange - 1, seed, 603)
    rv = W.f64_uniform(nr, -1.0, 1.0, seed, 604)
    return lk, lv, rk, rv, jt


@pytest.mark.parametrize("case", ["join_inner_small", "join_left_small", "join_full_small",
                                  "join_inner_radix", "join_full_radix"])
@pytest.mark.parametrize("path", PATHS)
def test_join_golden(eng, forced, case, path):
    lk, lv, rk, rv, jt = join_case(MANIFEST[case]["argv"])
    forced(*path)
    got = run_join(eng, [("k", lk), ("lv", lv)], [("k", rk), ("rv", rv)], ["k"], ["k"], jt)
    assert [c[0] for c in got] == ["k", "lv", "rv"]
    arrays = [c[3] for c in got]
    ent = MANIFEST[case]
    if ent["outputs"]:
        ref = read_tdt(os.path.join(GOLD, ent["outputs"][0]))
        row_multiset_equal(arrays, [c[3] for c in ref], case)
    else:
        assert len(arrays[0]) == ent["rows_out"][0]
        assert canonical_digest(arrays) == ent["digest"][0]


def test_join_kats(eng, forced):
    # test_exec.c /join (691-744): 4 rows
    for path in PATHS:
        forced(*path)
        got = run_join(eng, [("id", np.array([1, 2, 3]))], [("id", np.array([1, 2, 2, 3])),
                       ("s", np.array([10, 20, 21, 30]))], ["id"], ["id"], 0)
        assert len(got[0][3]) == 4
        assert sorted(got[1][3].tolist()) == [10, 20, 21, 30]
    # /join_large (747-831): 200,000 rows, score sum 99,999,000,000
    i = np.arange(100000)
    for path in PATHS:
        forced(*path)
        got = run_join(eng, [("id", i % 50000)], [("id", i % 50000), ("score", i * 10)],
                       ["id"], ["id"], 0)
        assert len(got[0][3]) == 200000 and int(got[1][3].sum()) == 99999000000
    # /join_empty (917-1025): 0 / 0 / 3 / 3 / 3
    a, e = np.array([1, 2, 3]), np.zeros(0, np.int64)
    for path in PATHS:
        forced(*path)
        n = [len(run_join(eng, [("id", x)], [("id", y), ("v", y * 2)], ["id"], ["id"], t)[0][3])
             for x, y, t in [(a, e, 0), (e, a, 0), (a, e, 1), (a, e, 2), (e, a, 2)]]
        assert n == [0, 0, 3, 3, 3]
    # /join_left_large (1028-1089) and /join_full_large (1092-1151)
    for path in PATHS:
        forced(*path)
        got = run_join(eng, [("id", i)], [("id", i * 2), ("v", i)], ["id"], ["id"], 1)
        assert len(got[0][3]) == 100000
        j = np.arange(80000)
        got = run_join(eng, [("id", j)], [("id", j + 40000), ("v", j)], ["id"], ["id"], 2)
        assert len(got[0][3]) == 120000
    # /join_skewed (1154-1228): every score 294
    forced(0)
    got = run_join(eng, [("id", np.full(100000, 42))], [("id", i), ("score", i * 7)],
                   ["id"], ["id"], 0)
    assert len(got[0][3]) == 100000 and set(got[1][3].tolist()) == {294}
    # /join_multikey (1299-1369): two keys, scores {1000, 2000}
    got = run_join(eng, [("a", np.array([1, 1, 2, 2, 3])), ("b", np.array([10., 20., 10., 20., 10.]))],
                   [("a", np.array([1, 2, 3])), ("b", np.array([10., 20., 30.])),
                    ("score", np.array([1000, 2000, 3000]))], ["a", "b"], ["a", "b"], 0)
    assert sorted(got[2][3].tolist()) == [1000, 2000]


def oracle_join_table(lcols, rcols, lkeys, rkeys, jt):
    li, ri = orc.join([dict(lcols)[k] for k in lkeys], [dict(rcols)[k] for k in rkeys], jt)
    out = [orc.gather(a, li, jt == 2) for _, a in lcols]
    out += [orc.gather(a, ri, jt >= 1) for n, a in rcols if n not in rkeys]
    return out


KEYGEN = {
    "i64_dense": lambda r, nl, nr: (r.integers(0, 3000, nl), r.integers(0, 3000, nr)),
    "i64_sparse": lambda r, nl, nr: (r.integers(-2 ** 40, 2 ** 40, nl) // 997 * 997,
                                     r.integers(-2 ** 40, 2 ** 40, nr) // 997 * 997),
    "i32": lambda r, nl, nr: (r.integers(-5000, 5000, nl).astype(np.int32),
                              r.integers(-5000, 5000, nr).astype(np.int32)),
    "f64": lambda r, nl, nr: (np.round(r.normal(size=nl), 2), np.round(r.normal(size=nr), 2)),
}


@pytest.mark.parametrize("kind", list(KEYGEN))
@pytest.mark.parametrize("jt", [0, 1, 2])
@pytest.mark.parametrize("path", [(0, 0), (2, 0), (2, 5)])
def test_join_vs_oracle(eng, forced, kind, jt, path):
    rng = np.random.default_rng(1000 + 10 * jt + list(KEYGEN).index(kind))
    nl, nr = 120001, 40003
    lk, rk = KEYGEN[kind](rng, nl, nr)
    lcols = [("k", lk), ("lv", rng.integers(-10 ** 6, 10 ** 6, nl)),
             ("lf", rng.normal(size=nl).astype(np.float64))]
    rcols = [("k", rk), ("rv", rng.integers(0, 100, nr).astype(np.int32)),
             ("rf", rng.random(nr))]
    forced(*path)
    got = run_join(eng, lcols, rcols, ["k"], ["k"], jt)
    ref = oracle_join_table(lcols, rcols, ["k"], ["k"], jt)
    assert [c[0] for c in got] == ["k", "lv", "lf", "rv", "rf"]
    row_multiset_equal([c[3] for c in got], ref, f"{kind}/{jt}/{path}")


# SYM key columns of each index width (td.h:121-133): dtype, attrs
SYM_W = {8: (np.uint8, 0), 16: (np.uint16, 1), 32: (np.uint32, 2), 64: (np.int64, 3)}


@pytest.mark.parametrize("widths", [(8, 64), (16, 32), (32, 16), (64, 8), (8, 8), (64, "i64")])
@pytest.mark.parametrize("path", [(0, 0), (1, 0), (2, 0), (2, 4)])
def test_join_sym_widths(eng, forced, widths, path):
    """SYM keys join by their widened index value (read_col_i64, exec.c:90-109,
    compared in join_keys_eq, exec.c:7754-7768), whatever the two widths."""
    rng = np.random.default_rng(500 + 8 * widths[0] + (widths[1] if widths[1] != "i64" else 1))
    nl, nr = 60001, 9001
    hi = min(256, 1 << min(w if w != "i64" else 64 for w in widths))
    lk = rng.integers(0, hi, nl)
    rk = rng.permutation(np.arange(nr) % hi)[:nr] if nr > hi else rng.integers(0, hi, nr)
    if widths == (64, "i64"):
        lk, rk = lk * 7919 + 2 ** 40, rk * 7919 + 2 ** 40

    def col(name, k, w):
        if w == "i64":
            return (name, k.astype(np.int64))
        dt, attrs = SYM_W[w]
        return (name, k.astype(dt), 20, attrs)

    lv = rng.integers(-10 ** 6, 10 ** 6, nl)
    rv = rng.random(nr)
    lcols = [col("k", lk, widths[0]), ("lv", lv)]
    rcols = [col("k", rk, widths[1]), ("rv", rv)]
    forced(*path)
    got = run_join(eng, lcols, rcols, ["k"], ["k"], 0)
    assert [c[0] for c in got] == ["k", "lv", "rv"]
    if widths[0] != "i64":
        assert got[0][1] == 20 and got[0][2] == SYM_W[widths[0]][1]    # left key keeps its width
    li, ri = orc.join([lk.astype(np.int64)], [rk.astype(np.int64)], 0)
    ref = [lk.astype(np.int64)[li], lv[li], rv[ri]]
    row_multiset_equal([got[0][3].astype(np.int64), got[1][3], got[2][3]], ref,
                       f"sym {widths} {path}")


def test_join_two_keys_and_resident(eng, forced):
    rng = np.random.default_rng(77)
    nl, nr = 50000, 20000
    lcols = [("a", rng.integers(0, 50, nl)), ("b", rng.integers(0, 40, nl).astype(np.int32)),
             ("v", rng.normal(size=nl))]
    rcols = [("a", rng.integers(0, 50, nr)), ("b", rng.integers(0, 40, nr).astype(np.int32)),
             ("w", rng.integers(0, 9, nr))]
    ref = oracle_join_table(lcols, rcols, ["a", "b"], ["a", "b"], 0)
    for path, resident in [((0, 0), False), ((2, 4), False), ((0, 0), True)]:
        forced(*path)
        got = run_join(eng, lcols, rcols, ["a", "b"], ["a", "b"], 0, resident)
        row_multiset_equal([c[3] for c in got], ref, f"two keys {path} {resident}")


# ------------------------------------------------------- join -> group --

def check_jg(out, case):
    ref = read_tdt(os.path.join(GOLD, case + ".tdt"))
    assert [c[0] for c in out] == [c[0] for c in ref]
    assert [c[1] for c in out] == [c[1] for c in ref]
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]],
                   [ref[0][3].astype(np.int64)], [c[3] for c in ref[1:]],
                   [True, False, False], case)


@pytest.mark.parametrize("case", ["jg_2m_250k", "jg_100k_8k_c1m"])
@pytest.mark.parametrize("form", ["two_call", "fused"])
@pytest.mark.parametrize("path", [(0, 0), (2, 0), (3, 0), (4, 0)])
@pytest.mark.parametrize("resident", [False, True])
def test_jg_golden(eng, forced, case, form, path, resident):
    nf, nd, nc, seed = (int(a) for a in MANIFEST[case]["argv"][1:5])
    tf, td, _, _ = D.jg_tables(eng, nf, nd, nc, seed)
    if resident:
        tf2, td2 = eng.put(tf), eng.put(td)
        eng.release(tf)
        eng.release(td)
        tf, td = tf2, td2
    forced(*path)
    run = D.jg_fused if form == "fused" else D.jg_two_call
    out = run(eng, tf, td)
    assert bool(eng.L.td_dev_is_resident(out)) == resident
    check_jg(eng.read_table(out), case)
    eng.release(out)
    eng.release(tf)
    eng.release(td)


@pytest.mark.parametrize("variant", ["dup_dim", "agg_on_dim", "key_on_fact", "left_join",
                                     "zipf_fk", "scattered", "scattered_dim_aggs",
                                     "scattered_dup"])
@pytest.mark.parametrize("jpath", [0, 3, 4])
def test_join_group_general(eng, forced, variant, jpath):
    """Fused JOIN -> GROUP shapes off the one-index fast path: duplicate
    build keys, aggregates over build columns, keys on the probe side, a
    left join (unmatched rows carry zero-filled build columns), and build
    keys scattered over 64-bit values (the key range is no direct map: the
    join runs radix partitioned into only the grouped columns, or through
    row pairs when build keys repeat)."""
    rng = np.random.default_rng(5)
    nf, nd = 300000, (200000 if variant == "zipf_fk" else 20000)
    fk = rng.integers(0, nd + 500, nf)
    if variant == "zipf_fk":
        # heavy key skew: the partitioned probes' regions overflow and the
        # direct-map probe takes over
        fk = np.minimum(rng.zipf(1.3, nf) - 1, nd + 499).astype(np.int64)
    qty = rng.integers(1, 101, nf)
    price = rng.lognormal(3, 1, nf)
    pk = rng.permutation(nd)
    if variant == "dup_dim":
        pk = pk % (nd // 2)
    if variant.startswith("scattered"):
        spread = lambda v: v * 7_000_000_019 - 2 ** 62         # unique, 64-bit range
        if variant == "scattered_dup":
            pk = pk % (nd // 2)
        fk, pk = spread(fk), spread(pk)
    cat = rng.integers(0, 77, nd).astype(np.int32)
    wgt = rng.normal(size=nd)
    tf = eng.table([("fk", fk), ("qty", qty), ("price", price)])
    td = eng.table([("pk", pk), ("cat", cat), ("wgt", wgt)])
    jt = 1 if variant == "left_join" else 0
    if variant == "key_on_fact":
        keys, aggs = ["qty"], [("SUM", "price"), ("MAX", "cat"), ("COUNT", "price")]
    elif variant in ("agg_on_dim", "scattered_dim_aggs"):
        keys, aggs = ["cat"], [("SUM", "wgt"), ("MIN", "qty"), ("AVG", "price")]
    else:
        keys, aggs = ["cat"], [("SUM", "price"), ("SUM", "qty"), ("COUNT", "price")]
    forced(jpath)
    g = eng.graph(tf)
    flt = g.filter(g.const_table(tf), g.gt(g.scan("qty"), g.const_i64(50)))
    g.join(flt, [g.scan("fk")], g.const_table(td), [g.scan("pk")], jt)
    root = g.group([g.scan(k) for k in keys], [(o, g.scan(c)) for o, c in aggs])
    out = eng.read_table(g.execute(root))
    plan = eng.last_plan()
    g.free()
    if variant in ("scattered", "scattered_dim_aggs") and jpath == 0:
        assert "join_group path=1" in plan and "path=1 " not in plan, plan   # 11..15
    # oracle composition
    m = qty > 50
    f = {"fk": fk[m], "qty": qty[m], "price": price[m]}
    li, ri = orc.join([f["fk"]], [pk], jt)
    cols = {c: orc.gather(a, li) for c, a in f.items()}
    cols.update({"cat": orc.gather(cat, ri, True), "wgt": orc.gather(wgt, ri, True)})
    ref = orc.group([cols[k] for k in keys], [o for o, _ in aggs], [cols[c] for _, c in aggs])
    f64 = [ref["out_types"][i] == 7 and o in ("SUM", "AVG") for i, (o, _) in enumerate(aggs)]
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]], ref["keys"],
                   ref["vals"], f64, variant, exact=ref["comp"], sum_abs=ref["sum_abs"])


TILED_SHAPES = {
    # name: (predicate builder, numpy mask, aggregates)
    "two_filters": (lambda g: [g.gt(g.scan("qty"), g.const_i64(30)),
                               g.lt(g.scan("price"), g.const_f64(50.0))],
                    lambda f: (f["qty"] > 30) & (f["price"] < 50.0),
                    [("SUM", "price"), ("SUM", "qty"), ("COUNT", "price")]),
    "or_not": (lambda g: [g.or_(g.gt(g.scan("qty"), g.const_i64(80)),
                                g.not_(g.eq(g.scan("disc"), g.const_i64(3))))],
               lambda f: (f["qty"] > 80) | ~(f["disc"] == 3),
               [("SUM", "price"), ("MIN", "qty"), ("MAX", "price"), ("AVG", "price")]),
    "cmp_col": (lambda g: [g.gt(g.scan("qty"), g.scan("disc"))],
                lambda f: f["qty"] > f["disc"],
                [("SUM", "disc"), ("MAX", "disc"), ("COUNT", "qty")]),
    "i32_key": (lambda g: [g.ge(g.scan("qty"), g.const_i64(50))],
                lambda f: f["qty"] >= 50,
                [("SUM", "price"), ("MIN", "price"), ("SUM", "disc")]),
}


@pytest.mark.parametrize("shape", list(TILED_SHAPES))
@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("jpath", [0, 3, 4])
def test_jg_tiled_shapes(eng, forced, shape, resident, jpath):
    """Fused JOIN -> GROUP through the tiled slot-map probe: stacked
    filters (a selection), OR / NOT and column-column predicates, 4-byte key
    and aggregate columns, probe keys missing from the build side, and a row
    count that leaves a partial last tile."""
    rng = np.random.default_rng(list(TILED_SHAPES).index(shape))
    nf, nd = 500003, 40000
    f = {"fk": rng.integers(0, nd + 300, nf), "qty": rng.integers(1, 101, nf),
         "price": rng.lognormal(3, 1, nf), "disc": rng.integers(0, 10, nf).astype(np.int32)}
    if shape == "i32_key":
        f["fk"] = f["fk"].astype(np.int32)
    pk = rng.permutation(nd)
    cat = rng.integers(0, 300, nd).astype(np.int32)
    tf = eng.table(list(f.items()))
    td = eng.table([("pk", pk), ("cat", cat)])
    if resident:
        tf2, td2 = eng.put(tf), eng.put(td)
        eng.release(tf)
        eng.release(td)
        tf, td = tf2, td2
    preds, mask_fn, aggs = TILED_SHAPES[shape]
    forced(jpath)
    g = eng.graph(tf)
    node = g.const_table(tf)
    for p in preds(g):
        node = g.filter(node, p)
    g.join(node, [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
    root = g.group([g.scan("cat")], [(o, g.scan(c)) for o, c in aggs])
    out = eng.read_table(g.execute(root))
    plan = eng.last_plan()
    g.free()
    eng.release(tf)
    eng.release(td)
    assert plan.startswith("join_group"), plan
    if jpath in (3, 4) and shape in ("two_filters", "i32_key"):
        assert f"path={jpath}" in plan, plan   # the partitioned probe ran
    # OR / NOT and column-column predicates are not range checks: the
    # partitioned probe declines them and the direct-map probe runs
    m = mask_fn(f)
    fm = {c: a[m] for c, a in f.items()}
    li, ri = orc.join([fm["fk"]], [pk], 0)
    cols = {c: orc.gather(a, li) for c, a in fm.items()}
    cols["cat"] = orc.gather(cat, ri)
    ref = orc.group([cols["cat"]], [o for o, _ in aggs], [cols[c] for _, c in aggs])
    f64 = [ref["out_types"][i] == 7 and o in ("SUM", "AVG") for i, (o, _) in enumerate(aggs)]
    assert [c[1] for c in out[1:]] == ref["out_types"]
    compare_groups([out[0][3].astype(np.int64)], [c[3] for c in out[1:]], ref["keys"],
                   ref["vals"], f64, shape, exact=ref["comp"], sum_abs=ref["sum_abs"])


# ------------------------------------------- materialising radix join --

@pytest.mark.parametrize("jt", [0, 1])
@pytest.mark.parametrize("case", ["i64", "i32_key", "two_keys", "f64_key", "filtered",
                                  "oversized", "sym16_key"])
def test_join_materialised_vs_oracle(eng, forced, jt, case):
    """INNER / LEFT joins on unique build keys through the materialising radix
    join (plan path 3): output columns of every type written straight from
    the partition passes, probe rows without a partner compacted away (inner)
    or zero-filled (left), against the oracle's pairs gathered column by
    column.  `oversized` forces 2 partitions of a 30,000-row build side, more
    than one LDS table holds: the pair-producing join answers instead."""
    rng = np.random.default_rng(900 + 10 * jt + len(case))
    nl, nr = 300_007, 30_011
    pool = rng.choice(np.arange(-2 ** 50, 2 ** 50, 2 ** 31 + 7), 2 * nr, replace=False)
    rk = pool[:nr]
    lk = pool[rng.integers(0, 2 * nr, nl)]                # about half the probes miss
    lcols = [("k", lk), ("a", rng.integers(-9, 9, nl).astype(np.int32)),
             ("b", rng.normal(size=nl)), ("c", rng.integers(0, 255, nl).astype(np.uint8)),
             ("d", rng.integers(-300, 300, nl).astype(np.int16))]
    rcols = [("k", rk), ("w", rng.integers(0, 10 ** 9, nr)), ("x", rng.random(nr)),
             ("y", rng.integers(0, 60000, nr).astype(np.uint16), 20, 1)]
    lkeys = rkeys = ["k"]
    if case == "i32_key":
        u = rng.permutation(2 ** 20)[:2 * nr].astype(np.int32) * 997
        lcols[0] = ("k", u[rng.integers(0, 2 * nr, nl)])
        rcols[0] = ("k", u[:nr])
    elif case == "sym16_key":
        u = rng.permutation(65536)[:2 * 20000].astype(np.uint16)
        nr2 = 20000
        rcols = [("k", u[:nr2], 20, 1), ("w", rng.integers(0, 10 ** 9, nr2))]
        lcols[0] = ("k", u[rng.integers(0, 2 * nr2, nl)], 20, 1)
    elif case == "two_keys":
        lcols.append(("k2", (lk % 7).astype(np.int32)))
        rcols.append(("k2", (rk % 7).astype(np.int32)))
        lkeys = rkeys = ["k", "k2"]
    elif case == "f64_key":
        lcols[0] = ("k", lk.astype(np.float64) / 4.0)
        rcols[0] = ("k", rk.astype(np.float64) / 4.0)
    forced(2, 1 if case == "oversized" else 0)
    lt, rt = eng.table(lcols), eng.table(rcols)
    g = eng.graph(lt)
    left = g.const_table(lt)
    if case == "filtered":
        left = g.filter(left, g.gt(g.scan("a"), g.const_i64(0)))
    root = g.join(left, [g.scan(k) for k in lkeys], g.const_table(rt),
                  [g.scan(k) for k in rkeys], jt)
    out = eng.read_table(g.execute(root))
    plan = eng.last_plan()
    g.free()
    eng.release(lt)
    eng.release(rt)
    assert ("path=3" in plan) == (case != "oversized"), plan
    L = {c[0]: c[1] for c in lcols}
    Rr = {c[0]: c[1] for c in rcols}
    m = L["a"] > 0 if case == "filtered" else np.ones(nl, bool)
    li, ri = orc.join([L[k][m] for k in lkeys], [Rr[k] for k in rkeys], jt)
    ref = [orc.gather(c[1][m], li) for c in lcols]
    ref += [orc.gather(c[1], ri, jt == 1) for c in rcols if c[0] not in rkeys]
    assert [c[0] for c in out] == [c[0] for c in lcols] + [c[0] for c in rcols
                                                          if c[0] not in rkeys]
    row_multiset_equal([c[3] for c in out], ref, f"mat {case} {jt}")


@pytest.mark.parametrize("jt", [0, 1])
@pytest.mark.parametrize("case", ["perm", "gaps", "i32_key", "sym16_key", "filtered",
                                  "far_probes", "empty_build", "dup"])
def test_join_map_vs_oracle(eng, jt, case):
    """INNER / LEFT joins whose single integer build key is dense in its range
    take the materialising map (plan path 4) at the planner's choice: the
    selected probe rows in row order, left columns compacted, right columns
    read through the map (zeros without a partner).  `far_probes` puts probe
    keys at the int64 extremes, outside the map; `dup` repeats build keys,
    which the map declines for the pair-producing join."""
    rng = np.random.default_rng(950 + 10 * jt + len(case))
    nl, nr = 400_009, 50_021
    rk = rng.permutation(nr) - 20_000
    lk = rng.integers(-25_000, nr - 15_000, nl)          # some probes outside the range
    if case == "gaps":
        rk = np.sort(rng.choice(3 * nr, nr, replace=False)) - 7
        rng.shuffle(rk)
        lk = rng.integers(-10, 3 * nr, nl)
    elif case == "far_probes":
        lk[rng.integers(0, nl, 5000)] = np.iinfo(np.int64).max
        lk[rng.integers(0, nl, 5000)] = np.iinfo(np.int64).min
    elif case == "empty_build":
        rk = rk[:0]
    elif case == "dup":
        rk = rng.integers(-20_000, 30_000, nr)
    nr = len(rk)
    lcols = [("k", lk), ("a", rng.integers(-9, 9, nl).astype(np.int32)),
             ("b", rng.normal(size=nl)), ("c", rng.integers(0, 255, nl).astype(np.uint8))]
    rcols = [("k", rk), ("w", rng.integers(0, 10 ** 9, nr)), ("x", rng.random(nr)),
             ("y", rng.integers(0, 60000, nr).astype(np.uint16), 20, 1),
             ("z", rng.integers(-100, 100, nr).astype(np.int16))]
    if case == "i32_key":
        lcols[0] = ("k", lk.astype(np.int32))
        rcols[0] = ("k", rk.astype(np.int32))
    elif case == "sym16_key":
        u = rng.permutation(40000).astype(np.uint16) + 2000
        rcols = [("k", u[:30000], 20, 1)] + [(c[0], c[1][:30000], *c[2:]) for c in rcols[1:]]
        lcols[0] = ("k", rng.integers(0, 45000, nl).astype(np.uint16), 20, 1)
    lt, rt = eng.table(lcols), eng.table(rcols)
    g = eng.graph(lt)
    left = g.const_table(lt)
    if case == "filtered":
        left = g.filter(left, g.gt(g.scan("a"), g.const_i64(0)))
    root = g.join(left, [g.scan("k")], g.const_table(rt), [g.scan("k")], jt)
    out = eng.read_table(g.execute(root))
    plan = eng.last_plan()
    g.free()
    eng.release(lt)
    eng.release(rt)
    assert ("join path=4" in plan) == (case != "dup"), plan
    L = {c[0]: c[1] for c in lcols}
    Rr = {c[0]: c[1] for c in rcols}
    m = L["a"] > 0 if case == "filtered" else np.ones(nl, bool)
    li, ri = orc.join([L["k"][m]], [Rr["k"]], jt)
    ref = [orc.gather(c[1][m], li) for c in lcols]
    ref += [orc.gather(c[1], ri, jt == 1) for c in rcols if c[0] != "k"]
    assert [c[0] for c in out] == [c[0] for c in lcols] + [c[0] for c in rcols[1:]]
    row_multiset_equal([c[3] for c in out], ref, f"map {case} {jt}")
    if case != "dup":
        # the map keeps the probe side's row order
        keep = np.sort(li) if jt == 0 else np.arange(int(m.sum()))
        np.testing.assert_array_equal(out[0][3], L["k"][m][keep])
