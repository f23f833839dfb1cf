"""Failure paths on the GPU, forced from the host so every kernel still
exits normally: cancellation (td_cancel, test_pool.c:272-318) at every
cancellation point of multi-pass queries, and queries over the arena
budget, which return TD_ERR_OOM (or fall back to a path that fits) with
their device scratch freed and the engine usable afterwards."""
import numpy as np
import pytest

from parity import compare_groups, row_multiset_equal
from teide_amd.td import TdError

pytestmark = pytest.mark.gpu

TD_ERR_OOM, TD_ERR_CANCEL = 1, 11


def test_cancel_then_next_query_works(eng):
    """The reference's /cancel: a td_cancel() before td_execute is consumed by
    it (td_execute clears the flag when it starts), and the next query runs
    normally: SUM of 1..100000 = n(n+1)/2."""
    n = 100000
    t = eng.table([("val", np.arange(1, n + 1, dtype=np.int64))])
    eng.L.td_cancel()
    g = eng.graph(t)
    try:
        r = g.execute(g.sum(g.scan("val")))
        eng.release(r)
    except TdError as e:
        assert e.code == TD_ERR_CANCEL
    g.free()
    g = eng.graph(t)
    assert eng.atom_value(g.execute(g.sum(g.scan("val")))) == n * (n + 1) // 2
    g.free()
    eng.release(t)


def hash_group(eng, t):
    g = eng.graph(t)
    root = g.group([g.scan("k")], [("SUM", g.scan("v")), ("MAX", g.scan("v")),
                                   ("COUNT", g.scan("v"))])
    return g, root


def radix_join(eng, t, d):
    g = eng.graph(t)
    root = g.join(g.filter(g.const_table(t), g.gt(g.scan("v"), g.const_i64(0))),
                  [g.scan("k")], g.const_table(d), [g.scan("k")], 0)
    return g, root


def join_group(eng, t, d):
    g = eng.graph(t)
    g.join(g.filter(g.const_table(t), g.gt(g.scan("v"), g.const_i64(0))), [g.scan("k")],
           g.const_table(d), [g.scan("k")], 0)
    root = g.group([g.scan("w")], [("SUM", g.scan("v")), ("COUNT", g.scan("v"))])
    return g, root


def same(eng, a, b, what):
    x, y = eng.read_table(a), eng.read_table(b)
    assert [c[0] for c in x] == [c[0] for c in y], what
    row_multiset_equal([c[3] for c in x], [c[3] for c in y], what)


@pytest.fixture(scope="module")
def tables(eng):
    rng = np.random.default_rng(41)
    n, nd = 400_003, 60_013
    keys = rng.integers(-2 ** 60, 2 ** 60, nd)
    t = eng.put(eng.table([("k", keys[rng.integers(0, nd, n)]),
                           ("v", rng.integers(-10 ** 6, 10 ** 6, n))]))
    d = eng.put(eng.table([("k", keys), ("w", rng.integers(0, 700, nd))]))
    yield t, d
    eng.release(t)
    eng.release(d)


@pytest.mark.parametrize("shape,paths", [("hash_group", (0, 4, 0)), ("hash_group", (0, 0, 0)),
                                         ("radix_join", (2, 0, 9)), ("join_group", (0, 0, 0))])
def test_cancel_at_every_point(eng, tables, shape, paths):
    t, d = tables
    build = {"hash_group": lambda: hash_group(eng, t), "radix_join": lambda: radix_join(eng, t, d),
             "join_group": lambda: join_group(eng, t, d)}[shape]
    eng.force_paths(*paths)
    try:
        g, root = build()
        p0 = eng.L.td_dev_cancel_polls()
        ref = g.execute(root)
        npolls = eng.L.td_dev_cancel_polls() - p0
        assert npolls >= 2, npolls
        eng.L.td_dev_sync()
        base = eng.mem_stats()["in_use"]
        for k in range(1, npolls + 1):
            eng.L.td_dev_cancel_after(k)
            with pytest.raises(TdError) as ei:
                eng.release(g.execute(root))
            assert ei.value.code == TD_ERR_CANCEL, (shape, k)
            eng.L.td_dev_cancel_after(0)
            eng.L.td_dev_sync()
            assert eng.mem_stats()["in_use"] == base, (shape, k)   # scratch freed
        got = g.execute(root)                            # usable afterwards
        same(eng, got, ref, f"{shape} after cancels")
        eng.release(got)
        eng.release(ref)
        g.free()
    finally:
        eng.L.td_dev_cancel_after(0)
        eng.force_paths(0, 0, 0)


def test_over_budget_returns_oom_and_frees(eng):
    """A hash-partitioned group-by of 60 M rows needs record buffers of
    several hundred MB each.  With the arena's budget pinned at what it
    holds, a query either runs to the same result or returns TD_ERR_OOM;
    either way the bytes in use return to where they were, and restoring
    the budget lets the query run."""
    rng = np.random.default_rng(42)
    n = 60_000_000
    keys = rng.integers(-2 ** 62, 2 ** 62, 3_000_000)
    t = eng.put(eng.table([("k", keys[rng.integers(0, len(keys), n)]),
                           ("v", rng.integers(-10 ** 6, 10 ** 6, n))]))
    g, root = hash_group(eng, t)
    eng.force_paths(0, 4, 0)
    try:
        ref = g.execute(root)
        ref_groups = eng.read_table(ref)
        eng.release(ref)
        outcomes = []
        for extra_gib in (0, 1, 2, 4, 8, 16):
            eng.L.td_dev_sync()
            eng.L.td_heap_gc()                       # return cached blocks
            m = eng.mem_stats()
            base = m["in_use"]
            eng.L.td_dev_set_budget(m["reserved"] + extra_gib * (1 << 30))
            try:
                got = g.execute(root)
                out = eng.read_table(got)
                eng.release(got)
                compare_groups([out[0][3]], [c[3] for c in out[1:]], [ref_groups[0][3]],
                               [c[3] for c in ref_groups[1:]], [False, False, False],
                               f"budget +{extra_gib} GiB")
                outcomes.append("ok")
            except TdError as e:
                assert e.code == TD_ERR_OOM, e
                outcomes.append("oom")
            eng.L.td_dev_sync()
            assert eng.mem_stats()["in_use"] == base, extra_gib
            assert eng.mem_stats()["reserved"] <= m["reserved"] + extra_gib * (1 << 30)
        assert outcomes[0] == "oom" and outcomes[-1] == "ok", outcomes
    finally:
        eng.L.td_dev_set_budget(0)
        eng.force_paths(0, 0, 0)
    got = g.execute(root)
    eng.release(got)
    g.free()
    eng.release(t)


@pytest.mark.parametrize("shape", ["hash_group", "radix_join", "join_group"])
def test_out_of_memory_falls_back(eng, shape):
    """The first large request of the partitioned plan fails (the injected
    out-of-memory of td_dev_fail_allocs): the group-by falls back to the
    global hash table, the materialising join to the pair-producing join,
    as the reference falls back to its sequential group table and chained
    join (exec.c:6486-6498, 8219-8231); same rows, nothing leaked.  With a
    forced plan the query returns TD_ERR_OOM instead."""
    rng = np.random.default_rng(43)
    n, nd = 20_000_000, 1_000_003
    keys = rng.integers(-2 ** 62, 2 ** 62, nd)
    t = eng.put(eng.table([("k", keys[rng.integers(0, nd, n)]),
                           ("v", rng.integers(-10 ** 6, 10 ** 6, n))]))
    d = eng.put(eng.table([("k", keys), ("w", rng.integers(0, 700, nd))]))
    g, root = {"hash_group": lambda: hash_group(eng, t), "radix_join": lambda: radix_join(eng, t, d),
               "join_group": lambda: join_group(eng, t, d)}[shape]()
    try:
        ref = g.execute(root)
        plan_ref = eng.last_plan()
        eng.L.td_dev_sync()
        base = eng.mem_stats()["in_use"]
        # the smallest request size that only the partitioned plan makes
        # (join_group partitions only the filtered rows' grouped columns)
        big = (64 if shape == "join_group" else 128) << 20
        eng.L.td_dev_fail_allocs(big, 1)
        got = g.execute(root)
        plan = eng.last_plan()
        eng.L.td_dev_fail_allocs(0, 0)
        assert plan != plan_ref, (plan_ref, plan)             # another plan answered
        if shape == "hash_group":
            assert "path=5" in plan_ref and "group path=3" in plan, (plan_ref, plan)
        same(eng, got, ref, f"{shape} after out-of-memory")
        eng.release(got)
        eng.L.td_dev_sync()
        assert eng.mem_stats()["in_use"] == base, shape
        # forced plans report the failure
        eng.force_paths(2, 4, 0)
        eng.L.td_dev_fail_allocs(big, 1)
        with pytest.raises(TdError) as ei:
            eng.release(g.execute(root))
        assert ei.value.code == TD_ERR_OOM
        eng.L.td_dev_sync()
        assert eng.mem_stats()["in_use"] == base, shape
        eng.release(ref)
    finally:
        eng.L.td_dev_fail_allocs(0, 0)
        eng.force_paths(0, 0, 0)
    g.free()
    eng.release(t)
    eng.release(d)
