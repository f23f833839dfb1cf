"""GPU parity of the scalar reductions over a column (exec_reduction,
exec.c:1729-1832, and exec_count_distinct, exec.c:1658-1727): the
reference's known answers (test_exec.c /reductions 526-598 and
/count_distinct 1518-1580), wrapping i64 products against numpy's uint64
product, f64 products against the sequential product within the rounding
bound of two n-term products, distinct counts against numpy's unique on
canonicalised f64 bits, and the empty-input values of the reference's
accumulator."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64_EPS = 2.0 ** -53


def reduce_one(eng, cols, op, col):
    t = eng.table(cols)
    g = eng.graph(t)
    r = g.execute(getattr(g, op)(g.scan(col)))
    typ = eng.hdr(r).type
    v = eng.atom_value(r)
    eng.release(r)
    g.free()
    eng.release(t)
    return typ, v


EXEC_TABLE = [("id1", np.array([1, 1, 2, 2, 3, 3, 1, 2, 3, 1], np.int64)),
              ("v1", np.arange(10, 101, 10, dtype=np.int64)),
              ("v3", np.arange(10, dtype=np.float64) + 1.5)]


def test_reduction_kats(eng):
    # test_exec.c /reductions (526-598)
    assert reduce_one(eng, EXEC_TABLE, "prod", "id1") == (-6, 216)
    assert reduce_one(eng, EXEC_TABLE, "min_", "v1")[1] == 10
    assert reduce_one(eng, EXEC_TABLE, "max_", "v1")[1] == 100
    assert reduce_one(eng, EXEC_TABLE, "avg", "v1") == (-7, 55.0)
    assert reduce_one(eng, EXEC_TABLE, "first", "v1")[1] == 10
    assert reduce_one(eng, EXEC_TABLE, "last", "v1")[1] == 100
    # test_exec.c /count_distinct (1518-1580)
    assert reduce_one(eng, EXEC_TABLE, "count_distinct", "id1") == (-6, 3)
    assert reduce_one(eng, EXEC_TABLE, "count_distinct", "v1")[1] == 10
    f = [("f", np.array([1.5, 2.5, 1.5, 2.5, 3.5]))]
    assert reduce_one(eng, f, "count_distinct", "f")[1] == 3
    assert reduce_one(eng, [("x", np.ones(5, np.int64))], "count_distinct", "x")[1] == 1


@pytest.mark.parametrize("n", [1, 63, 1000, 65537, 3_000_001])
@pytest.mark.parametrize("dtype", ["i64", "i32", "i16", "u8", "sym16"])
def test_prod_integer_wraps(eng, n, dtype):
    rng = np.random.default_rng(n + len(dtype))
    if dtype == "i64":
        a = rng.integers(-2 ** 62, 2 ** 62, n) | 1          # odd: the wrapped product stays non-zero
        col = ("x", a)
    elif dtype == "i32":
        a = (rng.integers(-40000, 40000, n) | 1).astype(np.int32)
        col = ("x", a)
    elif dtype == "i16":
        a = (rng.integers(-300, 300, n) | 1).astype(np.int16)
        col = ("x", a)
    elif dtype == "u8":
        a = (rng.integers(0, 256, n) | 1).astype(np.uint8)
        col = ("x", a)
    else:
        a = (rng.integers(0, 65536, n) | 1).astype(np.uint16)
        col = ("x", a, 20, 1)
    want = int(np.multiply.reduce(a.astype(np.int64).astype(np.uint64)).astype(np.int64))
    typ, got = reduce_one(eng, [col], "prod", "x")
    assert typ == -6 and got == want


@pytest.mark.parametrize("n", [1, 5, 4097, 1_000_003])
def test_prod_f64(eng, n):
    rng = np.random.default_rng(77 + n)
    a = np.exp(rng.normal(0, 2.0 / math.sqrt(n) if n > 1 else 1.0, n))
    a[rng.random(n) < 0.3] *= -1
    seq = 1.0
    for v in a.tolist():            # the reference's order on one worker
        seq *= v
    typ, got = reduce_one(eng, [("x", a)], "prod", "x")
    assert typ == -7
    # each side is within (n - 1) roundings of the exact product
    assert abs(got - seq) <= 2.0 * n * F64_EPS * abs(seq), (got, seq)
    assert math.copysign(1.0, got) == (-1.0) ** int((a < 0).sum())


def test_prod_f64_special_values(eng):
    base = np.full(1000, 1.5)
    for planted, check in ((0.0, lambda v: v == 0.0), (np.inf, math.isinf),
                           (np.nan, math.isnan), (-np.inf, lambda v: v == -np.inf)):
        a = base.copy()
        a[517] = planted
        a[:2] = [1.0, 1.0]
        assert check(reduce_one(eng, [("x", a)], "prod", "x")[1]), planted


def test_reductions_of_no_rows(eng):
    # accumulator defaults of exec_reduction (exec.c:1593-1600, 1808-1832)
    ei, ef = [("x", np.zeros(0, np.int64))], [("x", np.zeros(0, np.float64))]
    assert reduce_one(eng, ei, "prod", "x") == (-6, 1)
    assert reduce_one(eng, ef, "prod", "x") == (-7, 1.0)
    for op in ("sum", "min_", "max_", "first", "last"):
        assert reduce_one(eng, ei, op, "x") == (-6, 0), op
        assert reduce_one(eng, ef, op, "x") == (-7, 0.0), op
    assert reduce_one(eng, ei, "avg", "x") == (-7, 0.0)
    assert reduce_one(eng, ef, "avg", "x") == (-7, 0.0)
    for op in ("var", "var_pop", "stddev", "stddev_pop"):
        assert math.isnan(reduce_one(eng, ef, op, "x")[1]), op
    assert reduce_one(eng, ei, "count_distinct", "x") == (-6, 0)
    assert reduce_one(eng, ei, "count", "x") == (-6, 0)


def canon_bits(a):
    """f64 bits with all NaNs one value and -0.0 equal to 0.0."""
    b = a.view(np.int64).copy()
    b[np.isnan(a)] = 0x7ff8000000000000
    b[a == 0.0] = 0
    return b


@pytest.mark.parametrize("kind", ["i64_scattered", "i64_dense", "i32", "i16", "u8", "bool",
                                  "sym8", "sym64", "f64_special"])
def test_count_distinct_vs_unique(eng, kind):
    rng = np.random.default_rng(len(kind) * 31)
    n = 1_000_003
    if kind == "i64_scattered":
        a = rng.integers(-2 ** 62, 2 ** 62, 300_000)[rng.integers(0, 300_000, n)]
        col, want = ("x", a), len(np.unique(a))
    elif kind == "i64_dense":
        a = rng.integers(-500, 70_000, n)
        col, want = ("x", a), len(np.unique(a))
    elif kind in ("i32", "i16", "u8"):
        dt = {"i32": np.int32, "i16": np.int16, "u8": np.uint8}[kind]
        a = rng.integers(np.iinfo(dt).min, int(np.iinfo(dt).max) + 1, n).astype(dt)
        col, want = ("x", a), len(np.unique(a))
    elif kind == "bool":
        a = rng.random(n) < 0.5
        col, want = ("x", a), 2
    elif kind == "sym8":
        a = rng.integers(0, 200, n).astype(np.uint8)
        col, want = ("x", a, 20, 0), len(np.unique(a))
    elif kind == "sym64":
        a = rng.integers(0, 5000, n) * 104729 + 7
        col, want = ("x", a, 20, 3), len(np.unique(a))
    else:
        a = np.round(rng.normal(size=n) * 100, 1)
        a[::7] = -0.0
        a[1::7] = 0.0
        nan_bits = np.array([0x7ff8000000000000, 0x7ff8000000000001, -0x0008000000000000,
                             0x7ff0000000000123], np.int64).view(np.float64)
        a[2::11] = nan_bits[rng.integers(0, 4, len(a[2::11]))]
        col, want = ("x", a), len(np.unique(canon_bits(a)))
    assert reduce_one(eng, [col], "count_distinct", "x") == (-6, want)
