"""GPU parity of batched execution (td_dev_execute_batch / td_dev_batch_*):
every query's result equals the oracle's join + group-by of the same
columns and its own td_execute result -- keys, COUNT and integer sums
bit-exact, f64 sums within the f64 tolerance -- including queries the batch
hands back to td_execute (host-resident inputs, duplicate build keys, other
shapes), captured-graph replays interleaved with other device work, and a
sample of the full 4,096-query seed-42 batch."""
import numpy as np
import pytest

import orc
from parity import compare_groups
from teide_amd import workloads as W

pytestmark = pytest.mark.gpu

# host copies of each query's columns, by fact-table handle, for the oracle
HOST = {}


def make_query(eng, rng, nf, nd, ncat, thr, resident=True, dup=False):
    fk = rng.integers(0, nd + max(1, nd // 20), nf)          # some probe keys miss
    qty = rng.integers(1, 101, nf)
    price = rng.lognormal(3, 1, nf)
    pk = rng.permutation(nd)
    if dup:
        pk = pk % max(1, nd // 2)
    cat = rng.integers(0, ncat, nd).astype(np.int32)
    tf = eng.table([("fk", fk), ("qty", qty), ("price", price)])
    td = eng.table([("pk", pk), ("cat", cat)])
    if resident:
        tf2, td2 = eng.put(tf), eng.put(td)
        eng.release(tf)
        eng.release(td)
        tf, td = tf2, td2
    HOST[tf] = dict(fk=fk, qty=qty, price=price, pk=pk, cat=cat)
    return tf, td, thr


def oracle_query(cols, thr, aggs=None):
    """GROUP(cat; aggs) over JOIN(FILTER(fact, qty > thr), fk = pk): the
    oracle's inner join, gathers and group-by."""
    aggs = aggs or [("SUM", "price"), ("SUM", "qty"), ("COUNT", "price")]
    m = cols["qty"] > thr
    li, ri = orc.join([cols["fk"][m]], [cols["pk"]], 0)
    src = {"price": orc.gather(cols["price"][m], li), "qty": orc.gather(cols["qty"][m], li)}
    return orc.group([orc.gather(cols["cat"], ri)], [o for o, _ in aggs],
                     [src[c] for _, c in aggs])


def check_oracle(eng, got, cols, thr, what, aggs=None):
    ref = oracle_query(cols, thr, aggs)
    a = eng.read_table(got)
    assert [c[1] for c in a[1:]] == ref["out_types"], what
    f64 = [t == 7 and o in ("SUM", "AVG") for t, (o, _) in
           zip(ref["out_types"], aggs or [("SUM", 0), ("SUM", 0), ("COUNT", 0)])]
    compare_groups([a[0][3].astype(np.int64)], [c[3] for c in a[1:]], ref["keys"], ref["vals"],
                   f64, what, exact=ref["comp"], sum_abs=ref["sum_abs"])
    if ref["order"] == 1:                            # direct-array order: ascending keys
        np.testing.assert_array_equal(a[0][3].astype(np.int64), ref["keys"][0], err_msg=what)


def build_dag(eng, tf, td, thr, aggs=None):
    g = eng.graph(tf)
    flt = g.filter(g.const_table(tf), g.gt(g.scan("qty"), g.const_i64(thr)))
    g.join(flt, [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
    aggs = aggs or [("SUM", "price"), ("SUM", "qty"), ("COUNT", "price")]
    root = g.group([g.scan("cat")], [(o, g.scan(c)) for o, c in aggs])
    return g, root


def check_same(eng, got, ref, what):
    a, b = eng.read_table(got), eng.read_table(ref)
    assert [c[0] for c in a] == [c[0] for c in b], what
    assert [c[1] for c in a] == [c[1] for c in b], what
    f64 = [c[1] == 7 for c in b[1:]]
    compare_groups([a[0][3].astype(np.int64)], [c[3] for c in a[1:]],
                   [b[0][3].astype(np.int64)], [c[3] for c in b[1:]], f64, what)
    np.testing.assert_array_equal(a[0][3], b[0][3], err_msg=what)   # ascending keys


def run_singles(eng, qs, aggs=None):
    out = []
    for tf, td, thr in qs:
        g, root = build_dag(eng, tf, td, thr, aggs)
        out.append(g.execute(root))
        g.free()
    return out


def test_batch_matches_single_queries(eng):
    rng = np.random.default_rng(21)
    qs = []
    for i in range(40):
        nf = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        nd = max(1, nf // 8)
        ncat = int(np.exp(rng.uniform(np.log(4), np.log(65536))))
        qs.append(make_query(eng, rng, nf, nd, ncat, int(rng.integers(1, 100))))
    # queries the batch does not take: host inputs, duplicate keys
    qs.append(make_query(eng, rng, 5000, 600, 10, 30, resident=False))
    qs.append(make_query(eng, rng, 5000, 600, 10, 30, dup=True))
    ref = run_singles(eng, qs)
    dags = [build_dag(eng, *q) for q in qs]
    rc, got = eng.execute_batch([g for g, _ in dags], [r for _, r in dags])
    assert rc == 0
    for i, (a, b) in enumerate(zip(got, ref)):
        check_same(eng, a, b, f"query {i}")
        check_oracle(eng, a, HOST[qs[i][0]], qs[i][2], f"query {i} vs oracle")
        eng.release(a)
        eng.release(b)
    for g, _ in dags:
        g.free()


def test_batch_graph_replay(eng):
    rng = np.random.default_rng(22)
    qs = [make_query(eng, rng, int(rng.integers(2000, 60000)), 500 + 37 * i, 3 + 41 * i,
                     int(rng.integers(1, 100))) for i in range(16)]
    ref = run_singles(eng, qs)
    dags = [build_dag(eng, *q) for q in qs]
    h = eng.batch_prepare([g for g, _ in dags], [r for _, r in dags])
    nb, tile_rows, block = eng.batch_info(h)
    assert nb == len(qs) and tile_rows > 0 and block in (256, 512)
    for rep in range(4):
        rc, got = eng.batch_run(h, len(qs))
        assert rc == 0
        for i, (a, b) in enumerate(zip(got, ref)):
            check_same(eng, a, b, f"replay {rep} query {i}")
            if rep in (0, 3):
                check_oracle(eng, a, HOST[qs[i][0]], qs[i][2], f"replay {rep} query {i} oracle")
            eng.release(a)
        # unrelated eager device work between replays
        extra = run_singles(eng, qs[:2])
        for x in extra:
            eng.release(x)
        eng.L.td_dev_sync()
    eng.L.td_dev_batch_free(h)
    for r in ref:
        eng.release(r)
    for g, _ in dags:
        g.free()


def test_batch_mixed_shapes_and_preds(eng):
    """Different aggregate lists cannot share one output shape: the batch
    takes the first shape and hands the rest to td_execute."""
    rng = np.random.default_rng(23)
    qs = [make_query(eng, rng, 30000, 3000, 50, 20 + i) for i in range(6)]
    shapes = [None, None, [("MIN", "price"), ("MAX", "qty"), ("AVG", "price")], None,
              [("SUM", "qty")], None]
    ref = [run_singles(eng, [q], s)[0] for q, s in zip(qs, shapes)]
    dags = [build_dag(eng, *q, s) for q, s in zip(qs, shapes)]
    rc, got = eng.execute_batch([g for g, _ in dags], [r for _, r in dags])
    assert rc == 0
    for i, (a, b) in enumerate(zip(got, ref)):
        check_same(eng, a, b, f"query {i}")
        check_oracle(eng, a, HOST[qs[i][0]], qs[i][2], f"query {i} vs oracle",
                     [(o, c) for o, c in (shapes[i] or
                      [("SUM", "price"), ("SUM", "qty"), ("COUNT", "price")])])
        eng.release(a)
        eng.release(b)
    for g, _ in dags:
        g.free()


def test_batch_full_size_sample(eng):
    """The full batched workload -- 4,096 queries of seed 42 (fact rows
    log-uniform in [1K, 2M], 1.08e9 in all, dim = fact / 8, groups
    log-uniform in [4, 65536]) -- prepared, autotuned and run as one batch
    with graph replay; 64 of its queries, spread over the size range, are
    checked against the oracle on the host copies of their columns."""
    n = 4096
    nf, ncat, thr = W.batch_spec(n, 42)
    tables, dags = [], []
    for q in range(n):
        nd = max(1, int(nf[q]) // 8)
        fact = W.join_fact(int(nf[q]), nd, 42 + q)
        dim = W.join_dim(nd, int(ncat[q]), 42 + q)
        hf = eng.table([("fk", fact["fk"]), ("qty", fact["qty"]), ("price", fact["price"])])
        hd = eng.table([("pk", dim["pk"]), ("cat", dim["cat"])])
        tf, td = eng.put(hf), eng.put(hd)
        eng.release(hf)
        eng.release(hd)
        tables += [tf, td]
        dags.append(build_dag(eng, tf, td, int(thr[q])))
    h = eng.batch_prepare([g for g, _ in dags], [r for _, r in dags])
    nb, _, _ = eng.batch_info(h)
    assert nb == n                                   # every query batched
    order = np.argsort(nf)
    sample = sorted(set(order[np.linspace(0, n - 1, 64).astype(int)].tolist()) |
                    {int(np.argmax(ncat)), int(np.argmin(ncat))})
    assert len(sample) >= 64
    for rep in range(2):                             # capture, then replay
        rc, got = eng.batch_run(h, n)
        assert rc == 0
        for q in (sample if rep else sample[:8]):
            nd = max(1, int(nf[q]) // 8)
            cols = dict(W.join_fact(int(nf[q]), nd, 42 + q))
            cols.update(W.join_dim(nd, int(ncat[q]), 42 + q))
            check_oracle(eng, got[q], cols, int(thr[q]), f"C4 query {q} run {rep}")
        for r in got:
            eng.release(r)
    eng.L.td_dev_batch_free(h)
    for g, _ in dags:
        g.free()
    for t in tables:
        eng.release(t)


def make_hash_query(eng, rng, nf, nd, ncat, thr, kind):
    """A query whose build keys are not a dense unique range: `scattered`
    unique 64-bit keys, `dup` keys repeated about twice, `hot` a handful of
    keys shared by every dim row (build partitions beyond one LDS table)."""
    if kind == "scattered":
        pk = (rng.permutation(nd).astype(np.int64) * 7_000_000_019) - (1 << 62)
    elif kind == "dup":
        pk = rng.integers(0, max(1, nd // 2), nd) * 1_000_003 + 11
    else:
        pk = rng.integers(0, 3, nd) * (1 << 40) - 5
    fk = np.concatenate([pk, rng.integers(-(1 << 62), 1 << 62, max(1, nd // 10))])
    fk = fk[rng.integers(0, len(fk), nf)]                     # most probe keys hit, some miss
    qty = rng.integers(1, 101, nf)
    price = rng.lognormal(3, 1, nf)
    cat = rng.integers(0, ncat, nd).astype(np.int32)
    hf = eng.table([("fk", fk), ("qty", qty), ("price", price)])
    hd = eng.table([("pk", pk), ("cat", cat)])
    tf, td = eng.put(hf), eng.put(hd)
    eng.release(hf)
    eng.release(hd)
    HOST[tf] = dict(fk=fk, qty=qty, price=price, pk=pk, cat=cat)
    return tf, td, thr


def test_batch_hash_queries(eng):
    """Queries with scattered, repeated and hot build keys join in the batch
    through its radix-partitioned hash path (no query runs alone), mixed
    with slot-map queries, over a capture and two replays; every query is
    checked against the oracle, MIN / MAX / AVG included."""
    rng = np.random.default_rng(31)
    qs = []
    for i in range(24):
        kind = ("scattered", "dup", "hot", "dense")[i % 4]
        nf = int(np.exp(rng.uniform(np.log(500), np.log(300000))))
        nd = max(1, nf // 8) if kind != "hot" else int(rng.integers(3000, 9000))
        ncat = int(np.exp(rng.uniform(np.log(4), np.log(20000))))
        thr = int(rng.integers(1, 100))
        if kind == "dense":
            qs.append(make_query(eng, rng, nf, nd, ncat, thr))
        elif kind == "hot":
            # ~nd/3 partners per probe row: a few thousand rows per group keep
            # the f64 sums inside the 1e-12 relative tolerance
            qs.append(make_hash_query(eng, rng, 3000, nd, max(ncat, 500), thr, kind))
        else:
            qs.append(make_hash_query(eng, rng, nf, nd, ncat, thr, kind))
    aggs = [("SUM", "price"), ("MIN", "price"), ("MAX", "qty"), ("AVG", "qty"),
            ("COUNT", "price")]
    dags = [build_dag(eng, *q, aggs) for q in qs]
    h = eng.batch_prepare([g for g, _ in dags], [r for _, r in dags])
    nb, _, _ = eng.batch_info(h)
    assert nb == len(qs)                             # every query in the batch
    nh, part_rows, max_bits = eng.batch_radix(h)
    assert nh == 18 and part_rows > 0 and max_bits >= 1   # the non-dense ones: radix path
    for rep in range(3):
        rc, got = eng.batch_run(h, len(qs))
        assert rc == 0
        for i, a in enumerate(got):
            check_oracle(eng, a, HOST[qs[i][0]], qs[i][2], f"run {rep} query {i}", aggs)
            eng.release(a)
    eng.L.td_dev_batch_free(h)
    for g, _ in dags:
        g.free()
