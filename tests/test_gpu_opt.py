"""GPU parity of optimized DAGs: constant filters folded into MATERIALIZE /
HEAD 0 (opt.c:461-494) and AND predicates split into cost-ordered filter
chains (opt.c:1258-1459) give the rows the oracle's predicate gives, as
root filters, under a join and under a group-by."""
import numpy as np
import pytest

import orc
from parity import compare_groups, row_multiset_equal

pytestmark = pytest.mark.gpu


def table(eng, rng, n):
    cols = dict(qty=rng.integers(1, 101, n), flag=(rng.random(n) < 0.5).astype(np.uint8),
                a=rng.integers(-50, 50, n).astype(np.int32),
                b=rng.integers(-50, 50, n).astype(np.int32), price=rng.lognormal(3, 1, n))
    return cols, eng.table(list(cols.items()))


@pytest.mark.parametrize("keep", [True, False])
def test_constant_filter_folds(eng, keep):
    rng = np.random.default_rng(5)
    cols, t = table(eng, rng, 70001)
    g = eng.graph(t)
    pred = g.gt(g.const_i64(3 if keep else 1), g.const_i64(2))
    out = eng.read_table(g.execute(g.optimize(g.filter(g.const_table(t), pred))))
    assert [c[0] for c in out] == list(cols)
    for (name, _, _, arr) in out:
        np.testing.assert_array_equal(arr, cols[name] if keep else cols[name][:0])
    g.free()
    eng.release(t)


def test_split_and_chain_vs_oracle(eng):
    rng = np.random.default_rng(6)
    n = 200003
    cols, t = table(eng, rng, n)
    g = eng.graph(t)
    p = g.and_(g.and_(g.gt(g.scan("qty"), g.const_i64(20)), g.eq(g.scan("flag"), g.const_i64(1))),
               g.lt(g.scan("a"), g.scan("b")))
    out = eng.read_table(g.execute(g.optimize(g.filter(g.const_table(t), p))))
    m = orc.bool_op("AND", orc.bool_op("AND", orc.cmp_const(cols["qty"], "GT", 20),
                                       orc.cmp_const(cols["flag"], "EQ", 1)),
                    orc.cmp_cols(cols["a"], cols["b"], "LT")).astype(bool)
    for (name, _, _, arr) in out:
        np.testing.assert_array_equal(arr, cols[name][m], err_msg=name)
    g.free()
    eng.release(t)


def test_split_under_join_and_group(eng):
    """GROUP(c; SUM price, COUNT) over JOIN(FILTER(fact, qty >= 30 AND a != 5),
    fk = pk), optimized: the split chain under the join, fused or not, gives
    the oracle's groups."""
    rng = np.random.default_rng(7)
    nf, nd = 300007, 20011
    fk = rng.integers(0, nd, nf)
    qty = rng.integers(1, 101, nf)
    a = rng.integers(0, 10, nf).astype(np.int32)
    price = rng.lognormal(3, 1, nf)
    pk = rng.permutation(nd)
    c = rng.integers(0, 300, nd).astype(np.int32)
    tf = eng.table([("fk", fk), ("qty", qty), ("a", a), ("price", price)])
    td = eng.table([("pk", pk), ("c", c)])
    g = eng.graph(tf)
    p = g.and_(g.ge(g.scan("qty"), g.const_i64(30)), g.ne(g.scan("a"), g.const_i64(5)))
    g.join(g.filter(g.const_table(tf), p), [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
    root = g.group([g.scan("c")], [("SUM", g.scan("price")), ("COUNT", g.scan("price"))])
    out = eng.read_table(g.execute(g.optimize(root)))
    m = (qty >= 30) & (a != 5)
    li, ri = orc.join([fk[m]], [pk], 0)
    ref = orc.group([c[ri]], ["SUM", "COUNT"], [price[m][li], None])
    compare_groups([out[0][3].astype(np.int64)], [x[3] for x in out[1:]], ref["keys"],
                   ref["vals"], [True, False], "split under join", exact=ref["comp"],
                   sum_abs=ref["sum_abs"])
    g.free()
    # the join alone, optimized: the filtered fact rows joined
    g = eng.graph(tf)
    root = g.join(g.filter(g.const_table(tf), g.and_(g.ge(g.scan("qty"), g.const_i64(30)),
                                                     g.ne(g.scan("a"), g.const_i64(5)))),
                  [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
    got = eng.read_table(g.execute(g.optimize(root)))
    ref = [fk[m][li], qty[m][li], a[m][li], price[m][li], c[ri]]
    row_multiset_equal([x[3] for x in got], ref, "split join rows")
    g.free()
    eng.release(tf)
    eng.release(td)
