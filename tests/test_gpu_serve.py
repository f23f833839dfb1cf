"""GPU parity of the request server (td_dev_serve_*): every request's result
in a mix of concurrent requests equals the same DAG run alone through
td_execute and the same request run alone through the server -- keys,
group sets, COUNT and integer sums bit-exact, f64 sums within the f64
tolerance -- with requests joining the shared scan at different steps,
under request-slot and device-memory admission limits, and with requests
outside the served family answered by td_execute."""
import numpy as np
import pytest

import orc
from parity import compare_groups
from test_gpu_batch import check_same

pytestmark = pytest.mark.gpu

CMPS = ["gt", "ge", "lt", "le", "eq"]
CMP_NP = {"gt": np.greater, "ge": np.greater_equal, "lt": np.less, "le": np.less_equal,
          "eq": np.equal}
# host copies of the served tables' columns, for the oracle
HOST = {}


@pytest.fixture(scope="module")
def tables(eng):
    rng = np.random.default_rng(31)
    nf, nd = 300_000, 40_000
    fk = rng.integers(0, nd + nd // 10, nf)            # a tenth of the probes miss
    qty = rng.integers(1, 101, nf)
    price = rng.lognormal(3, 1, nf)
    pk = rng.permutation(nd) + 5
    cat = rng.integers(0, 50, nd).astype(np.int32)
    k64 = rng.integers(0, 3000, nd)
    flag = rng.integers(0, 4, nf)
    wide = rng.integers(0, 2_000_000, nd)              # 2 M-group key range
    mid = rng.integers(0, 300_000, nd)                 # 300 K groups x 100 qty values
    c7 = (cat % 7).astype(np.int64)
    pz = np.round(price, 0) - 20.0                     # ties, zeros of both signs, NaN
    pz[rng.random(nf) < 0.01] = np.nan
    pz[pz == 0.0] = np.where(rng.random(int((pz == 0.0).sum())) < 0.5, -0.0, 0.0)
    HOST.update(fk=fk, qty=qty, price=price, pk=pk, cat=cat, k64=k64, flag=flag, wide=wide,
                mid=mid, c7=c7, pz=pz)
    tf = eng.table([("fk", fk), ("qty", qty), ("price", price), ("flag", flag), ("pz", pz)])
    td = eng.table([("pk", pk), ("cat", cat), ("k64", k64), ("wide", wide), ("mid", mid),
                    ("c7", c7)])
    rf, rd = eng.put(tf), eng.put(td)
    eng.release(tf)
    eng.release(td)
    yield rf, rd
    eng.release(rf)
    eng.release(rd)


def request_dag(eng, rf, rd, key, cmp, t, aggs, filtered=True):
    g = eng.graph(rf)
    src = g.const_table(rf)
    if filtered:
        src = g.filter(src, getattr(g, cmp)(g.scan("qty"), g.const_i64(t)))
    g.join(src, [g.scan("fk")], g.const_table(rd), [g.scan("pk")], 0)
    root = g.group([g.scan(key)], [(o, g.scan(c)) for o, c in aggs])
    return g, root


AGG_SETS = [
    [("SUM", "price"), ("SUM", "qty"), ("COUNT", "price")],
    [("AVG", "price"), ("AVG", "qty"), ("COUNT", "qty")],
    [("COUNT", "price")],
    [("SUM", "qty"), ("SUM", "price"), ("AVG", "price")],
]


def random_requests(rng, n):
    reqs = []
    for i in range(n):
        cmp = CMPS[int(rng.integers(0, len(CMPS)))] if i % 3 else "gt"
        t = int(rng.integers(1, 101))
        if i % 11 == 5:
            t = int(rng.choice([-5, 0, 100, 250]))       # empty / everything
        reqs.append(dict(key="cat" if rng.random() < 0.5 else "k64", cmp=cmp, t=t,
                         aggs=AGG_SETS[int(rng.integers(0, len(AGG_SETS)))]))
    return reqs


def open_server(eng, rf, rd, **kw):
    return eng.serve_open(rf, rd, "fk", "pk", "qty", ["price"], ["cat", "k64"],
                          chunk_rows=kw.pop("chunk_rows", 40960), **kw)


def run_to_end(eng, h, results):
    for _ in range(10000):
        info = eng.serve_info(h)
        if info["queued"] == 0 and info["active"] == 0:
            break
        eng.serve_step(h)
        for rid, res in eng.serve_poll(h):
            results[rid] = res
    for rid, res in eng.serve_poll(h):
        results[rid] = res


def alone(eng, rf, rd, r):
    g, root = request_dag(eng, rf, rd, r["key"], r["cmp"], r["t"], r["aggs"],
                          r.get("filtered", True))
    out = g.execute(root)
    g.free()
    return out


def check_oracle(eng, got, r, what):
    """GROUP(key; aggs) over JOIN(FILTER(fact, qty CMP t), fk = pk) by the
    oracle's join, gathers and group-by on the host copies."""
    h = HOST
    m = CMP_NP[r["cmp"]](h["qty"], r["t"]) if r.get("filtered", True) else \
        np.ones(len(h["qty"]), bool)
    li, ri = orc.join([h["fk"][m]], [h["pk"]], 0)
    src = {c: orc.gather(h[c][m], li) for c in ("price", "qty", "flag")}
    ops = [o for o, _ in r["aggs"]]
    ref = orc.group([orc.gather(h[r["key"]], ri)], ops, [src[c] for _, c in r["aggs"]])
    a = eng.read_table(got)
    assert [c[1] for c in a[1:]] == ref["out_types"], what
    f64 = [t == 7 and o in ("SUM", "AVG") for t, o in zip(ref["out_types"], ops)]
    compare_groups([a[0][3].astype(np.int64)], [c[3] for c in a[1:]], ref["keys"], ref["vals"],
                   f64, what, exact=ref["comp"], sum_abs=ref["sum_abs"])
    if ref["order"] == 1:
        np.testing.assert_array_equal(a[0][3].astype(np.int64), ref["keys"][0], err_msg=what)


def test_serve_mix_matches_alone(eng, tables):
    rf, rd = tables
    rng = np.random.default_rng(32)
    reqs = random_requests(rng, 36)
    reqs.append(dict(key="k64", cmp="gt", t=0, aggs=AGG_SETS[0], filtered=False))
    # outside the served family: answered by td_execute at admission
    reqs.append(dict(key="cat", cmp="gt", t=40, aggs=[("MIN", "price"), ("MAX", "qty")]))
    h = open_server(eng, rf, rd)
    assert eng.serve_info(h)["chunks"] == 8
    dags, ids, results = {}, {}, {}
    # requests arrive in waves while earlier ones are mid-scan
    waves = [reqs[0:10], reqs[10:11], reqs[11:25], reqs[25:]]
    for w in waves:
        for r in w:
            g, root = request_dag(eng, rf, rd, r["key"], r["cmp"], r["t"], r["aggs"],
                                  r.get("filtered", True))
            rid = eng.serve_submit(h, g, root)
            dags[rid] = g
            ids[rid] = r
        for _ in range(3):
            eng.serve_step(h)
            for rid, res in eng.serve_poll(h):
                results[rid] = res
    run_to_end(eng, h, results)
    info = eng.serve_info(h)
    assert info["alone"] == 1 and info["shared"] == len(reqs) - 1
    assert info["held"] == info["base"]
    assert sorted(results) == sorted(ids)
    for rid, r in ids.items():
        ref = alone(eng, rf, rd, r)
        check_same(eng, results[rid], ref, f"request {rid} {r}")
        check_oracle(eng, results[rid], r, f"request {rid} {r} vs oracle")
        eng.release(results[rid])
        eng.release(ref)
    # the same requests, each alone in the server
    for rid in list(ids)[:8]:
        r = ids[rid]
        g, root = request_dag(eng, rf, rd, r["key"], r["cmp"], r["t"], r["aggs"],
                              r.get("filtered", True))
        one = {}
        rid1 = eng.serve_submit(h, g, root)
        run_to_end(eng, h, one)
        ref = alone(eng, rf, rd, r)
        check_same(eng, one[rid1], ref, f"alone in server {r}")
        check_oracle(eng, one[rid1], r, f"alone in server {r} vs oracle")
        eng.release(one[rid1])
        eng.release(ref)
        g.free()
    eng.serve_close(h)
    for g in dags.values():
        g.free()


def test_serve_large_cube_steps(eng, tables):
    """A key whose cube (100 qty values x 300 K groups) has more slices
    than one step folds in LDS: steps where it is active build the cubes
    with global atomics, steps with only the small key by slices; requests
    of both kinds, joining at different steps, give the oracle's groups."""
    rf, rd = tables
    h = eng.serve_open(rf, rd, "fk", "pk", "qty", ["price"], ["cat", "mid"], chunk_rows=40960)
    rng = np.random.default_rng(33)
    first = [dict(key="cat", cmp=c, t=int(rng.integers(1, 101)), aggs=AGG_SETS[i % 4])
             for i, c in enumerate(["gt", "le", "eq"])]
    later = [dict(key=k, cmp=c, t=int(rng.integers(1, 101)), aggs=AGG_SETS[i % 4])
             for i, (k, c) in enumerate([("mid", "ge"), ("cat", "lt"), ("mid", "gt")])]
    dags, ids, results = [], {}, {}
    for wave in (first, later):
        for r in wave:
            g, root = request_dag(eng, rf, rd, r["key"], r["cmp"], r["t"], r["aggs"])
            ids[eng.serve_submit(h, g, root)] = r
            dags.append(g)
        for _ in range(3):
            eng.serve_step(h)
            for rid, res in eng.serve_poll(h):
                results[rid] = res
    run_to_end(eng, h, results)
    info = eng.serve_info(h)
    assert info["alone"] == 0 and info["shared"] == len(ids), info
    for rid, r in ids.items():
        check_oracle(eng, results[rid], r, f"request {rid} {r} vs oracle")
        eng.release(results[rid])
    eng.serve_close(h)
    for g in dags:
        g.free()


def test_serve_three_keys_two_values(eng, tables):
    """Three served keys (the joint map's 4-wide rows) and two value columns
    (two value words per cube cell): cube requests over every key, with
    SUM / AVG of both value columns and every comparison, match the oracle
    and all share the scan."""
    rf, rd = tables
    h = eng.serve_open(rf, rd, "fk", "pk", "qty", ["price", "flag"], ["cat", "k64", "c7"],
                       chunk_rows=40960)
    aggs = [[("SUM", "price"), ("SUM", "flag"), ("COUNT", "price")],
            [("AVG", "flag"), ("SUM", "qty"), ("COUNT", "qty")],
            [("SUM", "flag")],
            [("AVG", "price"), ("AVG", "flag"), ("AVG", "qty")]]
    rng = np.random.default_rng(34)
    reqs = [dict(key=k, cmp=CMPS[i % len(CMPS)], t=int(rng.integers(1, 101)),
                 aggs=aggs[i % len(aggs)])
            for i, k in enumerate(["cat", "k64", "c7"] * 4)]
    dags, ids, results = [], {}, {}
    for i, r in enumerate(reqs):
        g, root = request_dag(eng, rf, rd, r["key"], r["cmp"], r["t"], r["aggs"])
        ids[eng.serve_submit(h, g, root)] = r
        dags.append(g)
        if i % 4 == 3:                                 # later requests join mid-scan
            eng.serve_step(h)
            for rid, res in eng.serve_poll(h):
                results[rid] = res
    run_to_end(eng, h, results)
    info = eng.serve_info(h)
    assert info["alone"] == 0 and info["shared"] == len(reqs), info
    assert sorted(results) == sorted(ids)
    for rid, r in ids.items():
        check_oracle(eng, results[rid], r, f"request {rid} {r} vs oracle")
        eng.release(results[rid])
    eng.serve_close(h)
    for g in dags:
        g.free()


def test_serve_slots_and_budget(eng, tables):
    rf, rd = tables
    req = dict(key="k64", cmp="ge", t=30, aggs=AGG_SETS[0])

    def submit(h, n):
        out = []
        for _ in range(n):
            g, root = request_dag(eng, rf, rd, req["key"], req["cmp"], req["t"], req["aggs"])
            out.append((eng.serve_submit(h, g, root), g))
        return out

    ref = alone(eng, rf, rd, req)
    # bytes one request holds while active
    h = open_server(eng, rf, rd)
    sub = submit(h, 1)
    eng.serve_step(h)
    info = eng.serve_info(h)
    per = info["held"] - info["base"]
    assert per > 0 and info["active"] == 1
    res = {}
    run_to_end(eng, h, res)
    eng.release(res[sub[0][0]])
    sub[0][1].free()
    base = info["base"]
    eng.serve_close(h)

    for kw, cap in ((dict(max_active=3), 3), (dict(budget=base + 2 * per + per // 2), 2)):
        h = open_server(eng, rf, rd, **kw)
        sub = submit(h, 7)
        res = {}
        peak = 0
        while True:
            info = eng.serve_info(h)
            if info["queued"] == 0 and info["active"] == 0:
                break
            eng.serve_step(h)
            peak = max(peak, eng.serve_info(h)["active"])
            if kw.get("budget"):
                assert eng.serve_info(h)["held"] <= kw["budget"]
            for rid, r in eng.serve_poll(h):
                res[rid] = r
        assert peak == cap
        for rid, g in sub:
            check_same(eng, res[rid], ref, f"admission {kw} request {rid}")
            check_oracle(eng, res[rid], req, f"admission {kw} request {rid} vs oracle")
            eng.release(res[rid])
            g.free()
        eng.serve_close(h)

    # a request that can never fit the budget finishes with an error
    h = open_server(eng, rf, rd, budget=base + per // 2)
    sub = submit(h, 1)
    eng.serve_step(h)
    done = eng.serve_poll(h)
    assert len(done) == 1 and done[0][0] == sub[0][0] and done[0][1] < 32
    sub[0][1].free()
    eng.serve_close(h)
    eng.release(ref)


# filters the cube cannot answer: f64 and wide-domain columns, several
# comparisons -- each request evaluates its own predicate on the shared chunks
WIDE = [
    ("price > t", lambda g, t: g.gt(g.scan("price"), g.const_f64(t)),
     lambda h, t: h["price"] > t, [7.5, 20.0, 55.25, 1e9, -1.0]),
    ("price <= t", lambda g, t: g.le(g.scan("price"), g.const_f64(t)),
     lambda h, t: h["price"] <= t, [3.0, 30.0]),
    ("fk < t", lambda g, t: g.lt(g.scan("fk"), g.const_i64(int(t))),
     lambda h, t: h["fk"] < int(t), [5000, 30000]),
    ("qty > 20 and price < t",
     lambda g, t: g.and_(g.gt(g.scan("qty"), g.const_i64(20)), g.lt(g.scan("price"), g.const_f64(t))),
     lambda h, t: (h["qty"] > 20) & (h["price"] < t), [15.0, 40.0]),
    ("qty < 10 or price > t",
     lambda g, t: g.or_(g.lt(g.scan("qty"), g.const_i64(10)), g.gt(g.scan("price"), g.const_f64(t))),
     lambda h, t: (h["qty"] < 10) | (h["price"] > t), [60.0]),
]


def wide_dag(eng, rf, rd, r):
    g = eng.graph(rf)
    src = g.filter(g.const_table(rf), WIDE[r["w"]][1](g, r["t"]))
    g.join(src, [g.scan("fk")], g.const_table(rd), [g.scan("pk")], 0)
    root = g.group([g.scan(r["key"])], [(o, g.scan(c)) for o, c in r["aggs"]])
    return g, root


def check_wide_oracle(eng, got, r, what):
    h = HOST
    m = WIDE[r["w"]][2](h, r["t"])
    li, ri = orc.join([h["fk"][m]], [h["pk"]], 0)
    src = {c: orc.gather(h[c][m], li) for c in ("price", "qty", "flag")}
    ops = [o for o, _ in r["aggs"]]
    ref = orc.group([orc.gather(h[r["key"]], ri)], ops, [src[c] for _, c in r["aggs"]])
    a = eng.read_table(got)
    assert [c[1] for c in a[1:]] == ref["out_types"], what
    f64 = [t == 7 and o in ("SUM", "AVG") for t, o in zip(ref["out_types"], ops)]
    compare_groups([a[0][3].astype(np.int64)], [c[3] for c in a[1:]], ref["keys"], ref["vals"],
                   f64, what, exact=ref["comp"], sum_abs=ref["sum_abs"])


def test_serve_wide_filters_share_the_scan(eng, tables):
    """Requests filtering on price (f64), on fk (a wide integer domain) or
    on several comparisons run in the shared scan beside cube requests:
    none runs alone, each matches the oracle and its own td_execute."""
    rf, rd = tables
    rng = np.random.default_rng(33)
    wide = [dict(w=w, t=t, key=k, aggs=AGG_SETS[(w + i) % len(AGG_SETS)])
            for w, (_, _, _, ts) in enumerate(WIDE) for i, t in enumerate(ts)
            for k in ("cat", "k64")]
    cube = random_requests(rng, 12)
    h = open_server(eng, rf, rd)
    dags, ids, results = [], {}, {}
    order = wide[:len(wide) // 2] + cube[:6] + wide[len(wide) // 2:] + cube[6:]
    for i, r in enumerate(order):
        g, root = wide_dag(eng, rf, rd, r) if "w" in r else \
            request_dag(eng, rf, rd, r["key"], r["cmp"], r["t"], r["aggs"])
        ids[eng.serve_submit(h, g, root)] = r
        dags.append(g)
        if i % 5 == 4:                                     # arrivals mid-scan
            eng.serve_step(h)
            for rid, res in eng.serve_poll(h):
                results[rid] = res
    run_to_end(eng, h, results)
    info = eng.serve_info(h)
    assert info["alone"] == 0 and info["shared"] == len(order), info
    assert info["held"] == info["base"]
    assert sorted(results) == sorted(ids)
    for rid, r in ids.items():
        what = f"request {rid} {r}"
        if "w" in r:
            g, root = wide_dag(eng, rf, rd, r)
            ref = g.execute(root)
            g.free()
            check_same(eng, results[rid], ref, what)
            check_wide_oracle(eng, results[rid], r, what + " vs oracle")
        else:
            ref = alone(eng, rf, rd, r)
            check_same(eng, results[rid], ref, what)
            check_oracle(eng, results[rid], r, what + " vs oracle")
        eng.release(results[rid])
        eng.release(ref)
    eng.serve_close(h)
    for g in dags:
        g.free()


def test_serve_threshold_buckets(eng, tables):
    """Requests `pz CMP t` (< <= > >=) fold through threshold buckets: more
    requests than one group holds, repeated thresholds, thresholds equal to
    column values, +0.0 / -0.0 and NaN in the column; each matches the
    oracle (NumPy's comparisons, false for NaN)."""
    rf, rd = tables
    cmps = ["gt", "ge", "lt", "le"]
    ts = [-20.0, -5.0, -0.0, 0.0, 0.0, 3.0, 7.0, 7.0, 15.5, 40.0, 1e9, -1e9]
    reqs = [dict(cmp=cmps[i % 4], t=t, key="cat" if i % 3 else "k64",
                 aggs=AGG_SETS[0] if i % 5 else AGG_SETS[2])
            for i, t in enumerate(ts * 3)]
    # one family of more requests than a group's thresholds (63): two groups
    reqs += [dict(cmp="gt", t=float(t), key="cat", aggs=AGG_SETS[0])
             for t in np.arange(-30.0, 40.0, 1.0)]
    h = open_server(eng, rf, rd)
    dags, ids, results = [], {}, {}
    for r in reqs:
        g = eng.graph(rf)
        pred = getattr(g, r["cmp"])(g.scan("pz"), g.const_f64(r["t"]))
        src = g.filter(g.const_table(rf), pred)
        g.join(src, [g.scan("fk")], g.const_table(rd), [g.scan("pk")], 0)
        root = g.group([g.scan(r["key"])], [(o, g.scan(c)) for o, c in r["aggs"]])
        ids[eng.serve_submit(h, g, root)] = r
        dags.append(g)
    run_to_end(eng, h, results)
    info = eng.serve_info(h)
    assert info["alone"] == 0 and info["shared"] == len(reqs), info
    hh = HOST
    for rid, r in ids.items():
        m = CMP_NP[r["cmp"]](hh["pz"], r["t"])
        li, ri = orc.join([hh["fk"][m]], [hh["pk"]], 0)
        src = {c: orc.gather(hh[c][m], li) for c in ("price", "qty", "flag")}
        ops = [o for o, _ in r["aggs"]]
        ref = orc.group([orc.gather(hh[r["key"]], ri)], ops, [src[c] for _, c in r["aggs"]])
        a = eng.read_table(results[rid])
        what = f"request {rid} {r}"
        assert [c[1] for c in a[1:]] == ref["out_types"], what
        f64 = [t == 7 and o in ("SUM", "AVG") for t, o in zip(ref["out_types"], ops)]
        compare_groups([a[0][3].astype(np.int64)], [c[3] for c in a[1:]], ref["keys"],
                       ref["vals"], f64, what, exact=ref["comp"], sum_abs=ref["sum_abs"])
        eng.release(results[rid])
    eng.serve_close(h)
    for g in dags:
        g.free()


def test_serve_wide_filters_many_groups(eng, tables):
    """price > t requests grouped by a key of 2 M groups: more table slices
    than the bucket plan takes, so the shared scan folds them with global
    atomics; still shared, still the oracle's groups."""
    rf, rd = tables
    h = eng.serve_open(rf, rd, "fk", "pk", "flag", ["price", "qty"], ["wide"],
                       chunk_rows=65536)
    reqs = [dict(w=0, t=t, key="wide", aggs=AGG_SETS[i % len(AGG_SETS)])
            for i, t in enumerate([10.0, 25.0, 80.0])]
    dags, ids, results = [], {}, {}
    for r in reqs:
        g, root = wide_dag(eng, rf, rd, r)
        ids[eng.serve_submit(h, g, root)] = r
        dags.append(g)
    run_to_end(eng, h, results)
    info = eng.serve_info(h)
    assert info["alone"] == 0 and info["shared"] == len(reqs), info
    for rid, r in ids.items():
        check_wide_oracle(eng, results[rid], r, f"request {rid} {r} vs oracle")
        eng.release(results[rid])
    eng.serve_close(h)
    for g in dags:
        g.free()
