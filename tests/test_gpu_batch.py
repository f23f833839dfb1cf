"""GPU parity of batched execution (td_dev_execute_batch / td_dev_batch_*):
every query's result equals its own td_execute result -- integers bit-exact,
f64 sums within the f64 tolerance -- including queries the batch hands back
to td_execute (host-resident inputs, duplicate build keys, other shapes),
and captured-graph replays interleaved with other device work."""
import numpy as np
import pytest

from parity import compare_groups

pytestmark = pytest.mark.gpu


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
    return tf, td, thr


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
        eng.release(a)
        eng.release(b)
    for g, _ in dags:
        g.free()
