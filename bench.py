#!/usr/bin/env python3
"""Benchmark: filter -> hash join -> group-by on one MI355X.

Workload (the reference's radix join + group-by configuration):
  fact 200,000,000 rows: fk i64 uniform [0,16M), qty i64 uniform [1,100],
                         price f64 lognormal(3,1)
  dim   16,000,000 rows: pk = permutation of [0,16M), cat i32 uniform [0,1000)
  GROUP(cat; SUM price, SUM qty, COUNT price) over
  JOIN(FILTER(fact, qty > 50), fk = pk, dim)    -- one td.h DAG, seed 42

A step is one td_execute of that DAG with both tables resident in HBM
(td_dev_put) and the result left resident.  `value` is fact rows/s over
the K timed steps.  `roofline` prices the dominant kernel's algorithmic
bytes against its average HIP-event duration over the same timed steps.
`cpu_baseline` times the reference's own multithreaded CPU path (built from
source into oracle/_ref) on a smaller sample of the same workload.

The same line then carries the two serving forms on their own workloads
(skip them with --no-forms): "batched" -- 4096 small filter+join+group-by
queries run as one batch, beside one td_execute per query -- and "served"
-- a Poisson request stream (2000/s) through the request server over
resident fact 100M x dim 8M tables, with p50/p99 latency, the closed-loop
capacity and the plain per-request rate.  `--workload batch|served|hc` runs
one of those workloads (or the high-cardinality group-by) alone.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0                      # MI355X HBM3E, GB/s
N_FACT, N_DIM, N_CAT, SEED = 200_000_000, 16_000_000, 1000, 42
CPU_SAMPLE = (200_000_000, 16_000_000)     # reference CPU path sample: the full C2 size


def kernel_bytes(name, nf, nd, ncat):
    """Algorithmic (compulsory) HBM bytes of one launch of a kernel of the
    fused pipeline: every referenced column read once, outputs written once."""
    if name.startswith("jg_probe"):
        # fact fk, qty, price (24 B/row) + the dim-row -> group-slot map read
        # once (4 B per dim row)
        return nf * 24 + nd * 4
    if name == "jg_slotmap":
        return nd * (8 + 4) + nd * 4       # pk, cat in; slot map out
    if name == "jg_build":
        return nd * 8 + nd * 4             # pk in; head array out
    return None


def cpu_threads():
    """Threads for the reference's CPU path: the reference's own default
    (td_pool_init(0): nproc - 1 workers + the caller) applied to the host
    cores this process may use, capped at the 16-core share a GPU box
    allots per GPU."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(2, min(16, n))


def host_info():
    """nproc, the cores this process may use and the CPU model."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": aff, "cpu_model": model}


def cpu_baseline_reference(nf, nd, ncat, reps=5):
    exe = os.path.join(REPO, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    thr = cpu_threads()
    env = dict(os.environ, TD_REF_THREADS=str(thr))
    out = subprocess.run([exe, "bench", "jg", str(nf), str(nd), str(ncat), str(reps)],
                         check=True, capture_output=True, text=True, env=env, timeout=600)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return dict({"value": nf / (r["median_ms"] / 1e3), "unit": "rows/s", "cores": r["threads"],
                 "kind": "reference", "median_ms": r["median_ms"],
                 "sample": f"fact {nf:,} x dim {nd:,} rows, {ncat} cats, same distributions and "
                           f"DAG (join graph + group graph), median of {reps} after 1 warm-up, "
                           f"td_pool_init({r['threads'] - 1}) workers + caller"}, **host_info())


def cpu_baseline_port(nf, nd, ncat):
    """Single-threaded oracle restatement, used when oracle/_ref is absent."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    from teide_amd import workloads as W
    f, d = W.join_fact(nf, nd, SEED), W.join_dim(nd, ncat, SEED)
    t0 = time.perf_counter()
    idx = orc.sel_to_idx(orc.sel_from_bool(orc.cmp_const(f["qty"], "GT", 50)), nf)
    fk, qty, price = (orc.gather(f[c], idx) for c in ("fk", "qty", "price"))
    li, ri = orc.join([fk], [d["pk"]], 0)
    cat = orc.gather(d["cat"], ri)
    orc.group([cat], ["SUM", "SUM", "COUNT"], [orc.gather(price, li), orc.gather(qty, li), None])
    dt = time.perf_counter() - t0
    return {"value": nf / dt, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"fact {nf:,} x dim {nd:,} rows, one run of the scalar oracle"}


def load_traffic(workload, name):
    """HBM bytes per launch of the workload's kernel measured with rocprofv3
    --pmc (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), when recorded."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh)
    ent = rec.get(f"{workload}/{name}")
    return ent.get("bytes_per_launch") if ent else None


def calibrate_copy(e, nbytes):
    """One device-to-device copy of nbytes through the library's copy kernel:
    reads nbytes and writes nbytes of HBM, the reference point for scaling
    FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C
    L = e.L
    L.tdd_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.tdd_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.tdd_fill.argtypes = [C.c_void_p, C.c_uint64, C.c_size_t]
    L.tdd_free.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    if L.tdd_malloc(C.byref(src), nbytes) or L.tdd_malloc(C.byref(dst), nbytes):
        raise SystemExit("calibrate-copy: device allocation failed")
    L.tdd_fill(src, 0x0123456789abcdef, nbytes)
    L.tdd_fill(dst, 0, nbytes)
    L.td_dev_sync()
    L.tdd_d2d(dst, src, nbytes)
    L.td_dev_sync()
    L.tdd_free(src)
    L.tdd_free(dst)
    print(f"# calibrate-copy: {nbytes} bytes", file=sys.stderr, flush=True)


class JoinGroup:
    """filter qty > 50 -> join fact x dim on fk = pk -> group by cat."""
    name = "jg"

    def __init__(self, e, a):
        from teide_amd import workloads as W
        self.e, self.a = e, a
        t0 = time.perf_counter()
        fact = W.join_fact(a.fact, a.dim, SEED)
        dim = W.join_dim(a.dim, a.cats, SEED)
        hf = e.table([("fk", fact["fk"]), ("qty", fact["qty"]), ("price", fact["price"])])
        del fact
        hd = e.table([("pk", dim["pk"]), ("cat", dim["cat"])])
        del dim
        t1 = time.perf_counter()
        self.tf, self.td = e.put(hf), e.put(hd)
        e.L.td_dev_sync()
        t2 = time.perf_counter()
        e.release(hf)
        e.release(hd)
        gbs = (a.fact * 24 + a.dim * 12) / (t2 - t1) / 1e9
        print(f"# data: gen {t1 - t0:.1f}s, h2d {t2 - t1:.2f}s ({gbs:.1f} GB/s)",
              file=sys.stderr, flush=True)
        self.rows = a.fact
        self.query_bytes = a.fact * 24 + a.dim * 12 + a.cats * 28
        self.config = {"workload": f"filter qty>50 -> join fact {a.fact:,} x dim {a.dim:,} on "
                                   f"fk=pk -> group by cat ({a.cats} cats): SUM price, SUM qty, "
                                   f"COUNT"}

    def step(self):
        e, tf, td = self.e, self.tf, self.td
        g = e.graph(tf)
        flt = g.filter(g.const_table(tf), g.gt(g.scan("qty"), g.const_i64(50)))
        g.join(flt, [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
        root = g.group([g.scan("cat")], [("SUM", g.scan("price")), ("SUM", g.scan("qty")),
                                         ("COUNT", g.scan("price"))])
        out = g.execute(root)
        g.free()
        e.release(out)

    def kernel_bytes(self, name):
        return kernel_bytes(name, self.a.fact, self.a.dim, self.a.cats)

    def cpu_baseline(self):
        nf, nd = min(CPU_SAMPLE[0], self.a.fact), min(CPU_SAMPLE[1], self.a.dim)
        return cpu_baseline_reference(nf, nd, self.a.cats) or cpu_baseline_port(
            min(nf, 5_000_000), min(nd, 400_000), self.a.cats)

    def release(self):
        self.e.release(self.tf)
        self.e.release(self.td)


class JoinGroupTwoCall(JoinGroup):
    """The join + group-by in the reference's own form: graph 1 runs
    JOIN(FILTER(fact, qty > 50), fk = pk) to a joined table in HBM, graph 2
    runs GROUP(cat; SUM price, SUM qty, COUNT price) over it."""
    name = "jg2"

    def __init__(self, e, a):
        super().__init__(e, a)
        self.config = dict(self.config, form="two graphs: join to a resident joined table, "
                                             "then group-by over it")

    def step(self):
        e = self.e
        g = e.graph(self.tf)
        flt = g.filter(g.const_table(self.tf), g.gt(g.scan("qty"), g.const_i64(50)))
        j = g.execute(g.join(flt, [g.scan("fk")], g.const_table(self.td), [g.scan("pk")], 0))
        g.free()
        g2 = e.graph(j)
        r = g2.group([g2.scan("cat")], [("SUM", g2.scan("price")), ("SUM", g2.scan("qty")),
                                         ("COUNT", g2.scan("price"))])
        e.release(g2.execute(r))
        g2.free()
        e.release(j)

    def kernel_bytes(self, name):
        a = self.a
        if name == "compact":
            # the passing fact rows' fk, qty, price read (every line of the
            # three columns at 50 % selectivity) and written, the selection
            # words read, cat read at the row's key and written
            return a.fact * 24 + a.fact // 2 * (24 + 4 + 4) + a.fact // 8
        if name == "rjoin_map_sel":
            return a.fact * 8 + a.fact // 8 * 2
        if name == "rjoin_map":
            return a.dim * (8 + 4 + 4)
        return None


class JoinGroupScattered(JoinGroup):
    """The join + group-by workload with its join keys spread over 64-bit
    values (fk, pk -> key * 7,000,000,019 - 2^62: still unique, no longer a
    dense range), so the join runs radix partitioned: the materialising
    join writes the filtered rows' grouped columns (price, qty, cat), which
    are then grouped."""
    name = "jgs"

    def __init__(self, e, a):
        from teide_amd import workloads as W
        self.e, self.a = e, a
        t0 = time.perf_counter()
        fact = W.join_fact(a.fact, a.dim, SEED)
        dim = W.join_dim(a.dim, a.cats, SEED)
        spread = lambda v: v * 7_000_000_019 - (1 << 62)
        hf = e.table([("fk", spread(fact["fk"])), ("qty", fact["qty"]), ("price", fact["price"])])
        self.sel = int((fact["qty"] > 50).sum())
        del fact
        hd = e.table([("pk", spread(dim["pk"])), ("cat", dim["cat"])])
        del dim
        self.tf, self.td = e.put(hf), e.put(hd)
        e.L.td_dev_sync()
        e.release(hf)
        e.release(hd)
        print(f"# data: gen + h2d {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        self.rows = a.fact
        self.query_bytes = a.fact * 24 + a.dim * 12 + a.cats * 28
        self.config = {"workload": f"filter qty>50 -> join fact {a.fact:,} x dim {a.dim:,} on "
                                   f"fk=pk with keys scattered over 64-bit values -> group by "
                                   f"cat ({a.cats} cats): SUM price, SUM qty, COUNT"}

    def kernel_bytes(self, name):
        n, m, s = self.a.fact, self.a.dim, self.sel
        if name == "rjoin_scatter":
            # per launch on average over 2 passes x 2 sides: fact pass 1 reads
            # fk of every row and qty, price of the passing rows and writes the
            # key word, qty, price of those; pass 2 reads and writes the 24 B;
            # dim passes read pk + cat, write key word + cat, then 12 B each way
            return (n * 8 + s * 16 + s * 24 + s * 48 + m * 24 + m * 24) // 4
        if name == "rjoin_build_probe":
            return m * 12 + s * (8 + 4 + 1)    # build key words + cat; probe key, cat out, flag
        if name == "rjoin_hist":
            return (n * 8 + m * 8) // 2
        return None

    def cpu_baseline(self):
        return None


HC_AGGS = [("SUM", "v_i64"), ("MIN", "v_i64"), ("MAX", "v_i64"), ("SUM", "v_f64"),
           ("MIN", "v_f64"), ("MAX", "v_f64"), ("COUNT", "v_i64")]


class HighCard:
    """group by a Zipf(1.1) key over 64 M values and by a u32 symbol column
    over 1 M symbols: SUM/MIN/MAX of an i64 and an f64 column, and COUNT."""
    name = "hc"

    def __init__(self, e, a):
        from teide_amd import workloads as W
        self.e, self.a = e, a
        t0 = time.perf_counter()
        d = W.high_card(a.hc_rows, a.hc_values, a.hc_syms, 1.1, SEED)
        h = e.table([("key", d["key"]), ("sym", d["sym"], 20, 2), ("v_i64", d["v_i64"]),
                     ("v_f64", d["v_f64"])])
        del d
        t1 = time.perf_counter()
        self.t = e.put(h)
        e.L.td_dev_sync()
        e.release(h)
        print(f"# data: gen {t1 - t0:.1f}s", file=sys.stderr, flush=True)
        self.rows = 2 * a.hc_rows                       # two group-bys per step
        out = self._group("key")
        self.g_key = self.e.L.td_table_nrows(out)
        # rows the key group-by's partition passes carry: the rows of its
        # hot keys are folded before them
        m = re.search(r"part_rows=(\d+)", e.last_plan())
        self.key_part_rows = int(m.group(1)) if m and int(m.group(1)) > 0 else a.hc_rows
        self.e.release(out)
        self.query_bytes = (a.hc_rows * 24 + self.g_key * 64) + (a.hc_rows * 20 + a.hc_syms * 60)
        self.config = {"workload": f"group {a.hc_rows:,} rows by key (zipf 1.1 over "
                                   f"{a.hc_values:,} values, {self.g_key:,} groups) and by sym "
                                   f"({a.hc_syms:,} symbols): SUM/MIN/MAX v_i64, v_f64, COUNT"}

    def _group(self, by):
        e = self.e
        g = e.graph(self.t)
        root = g.group([g.scan(by)], [(op, g.scan(c)) for op, c in HC_AGGS])
        out = g.execute(g.optimize(root))
        g.free()
        return out

    def step(self):
        self.e.release(self._group("key"))
        self.e.release(self._group("sym"))

    def kernel_bytes(self, name):
        n, c = self.a.hc_rows, self.key_part_rows
        if name == "group_pg_pass":
            # per launch on average over the step's four scatter passes: the
            # key group-by's two passes read and write 24-byte records of the
            # rows its hot-key fold left (key, v_i64, v_f64), the symbol
            # group-by's first reads its rows (sym u32, v_i64, v_f64) and
            # writes 20-byte records, its second reads and writes records
            return (c * (48 + 48) + n * (40 + 40)) // 4
        if name == "group_pg_hagg":
            return c * 24                      # one read of the key records
        if name == "group_pg_hotfold":
            return n * 24 + c * 24             # rows read, the cold rows' records written
        return None

    def cpu_baseline(self):
        exe = os.path.join(REPO, "oracle", "_ref", "ref_harness")
        if not os.path.exists(exe):
            return None
        n = min(self.a.hc_rows, 30_000_000)
        nv = max(1, self.a.hc_values * n // self.a.hc_rows)
        ns = max(1, self.a.hc_syms * n // self.a.hc_rows)
        env = dict(os.environ, TD_REF_THREADS=str(cpu_threads()))
        out = subprocess.run([exe, "bench", "hc", str(n), str(nv), str(ns), "3"], check=True,
                             capture_output=True, text=True, env=env, timeout=900)
        r = json.loads(out.stdout.strip().splitlines()[-1])
        return dict({"value": 2 * n / (r["median_ms"] / 1e3), "unit": "rows/s",
                     "cores": r["threads"], "kind": "reference",
                     "sample": f"{n:,} rows, {nv:,} zipf values, {ns:,} symbols, both "
                               f"group-bys, median of 3 after 1 warm-up, "
                               f"td_pool_init({r['threads'] - 1}) workers + caller"},
                    **host_info())

    def release(self):
        self.e.release(self.t)


class Charge:
    """The pricing expression of a line-item scan,
    price * (1 - disc) * (1 + tax), over three resident f64 columns into a
    new f64 column: one fused expression kernel, 24 bytes read and 8
    written per row."""
    name = "charge"
    metric = "rows/s + % HBM peak of a fused expression (price * (1 - disc) * (1 + tax))"

    def __init__(self, e, a):
        from teide_amd import workloads as W
        self.e, self.a = e, a
        n = a.expr_rows
        cols = [("price", W.f64_uniform(n, 900.0, 105000.0, SEED, 701)),
                ("disc", W.f64_uniform(n, 0.0, 0.1, SEED, 702)),
                ("tax", W.f64_uniform(n, 0.0, 0.08, SEED, 703))]
        h = e.table(cols)
        del cols
        self.t = e.put(h)
        e.L.td_dev_sync()
        e.release(h)
        self.rows = n
        self.query_bytes = n * 32
        self.config = {"workload": f"expression price * (1 - disc) * (1 + tax) over {n:,} rows "
                                   "of three f64 columns into an f64 column"}

    def step(self):
        e = self.e
        g = e.graph(self.t)
        root = g.mul(g.mul(g.scan("price"), g.sub(g.const_f64(1.0), g.scan("disc"))),
                     g.add(g.const_f64(1.0), g.scan("tax")))
        v = g.execute(g.optimize(root))
        g.free()
        e.release(v)

    def kernel_bytes(self, name):
        return self.rows * 32 if name == "expr" else None

    def cpu_baseline(self):
        exe = os.path.join(REPO, "oracle", "_ref", "ref_harness")
        if not os.path.exists(exe):
            return None
        n = min(self.a.expr_rows, 100_000_000)
        env = dict(os.environ, TD_REF_THREADS=str(cpu_threads()))
        out = subprocess.run([exe, "bench", "charge", str(n), "5"], check=True,
                             capture_output=True, text=True, env=env, timeout=900)
        r = json.loads(out.stdout.strip().splitlines()[-1])
        return {"value": n / (r["median_ms"] / 1e3), "unit": "rows/s", "cores": r["threads"],
                "kind": "reference",
                "sample": f"{n:,} rows, median of 5 after 1 warm-up"}

    def release(self):
        self.e.release(self.t)


class Join:
    """A plain inner join (the radix path): fact rows with a scattered 64-bit
    key and an f64 payload against a dim table of unique scattered keys and
    an i64 payload; every fact row matches one dim row.  Output: the joined
    table (fk, x, w) in HBM."""
    name = "join"
    metric = "fact rows/s of a radix hash join into a resident joined table"

    def __init__(self, e, a):
        from teide_amd import workloads as W
        import numpy as np
        self.e, self.a = e, a
        n, m = a.join_fact, a.join_dim
        t0 = time.perf_counter()
        pk = W.permutation(m, 0, SEED, 901) * 1_000_003 + 7     # unique, range far beyond m
        idx = W.i64_uniform(n, 0, m - 1, SEED, 902)
        fk = pk[idx]
        del idx
        hf = e.table([("fk", fk), ("x", W.f64_uniform(n, 0.0, 1.0, SEED, 903))])
        del fk
        hd = e.table([("pk", pk), ("w", W.i64_uniform(m, -10 ** 9, 10 ** 9, SEED, 904))])
        del pk
        self.tf, self.td = e.put(hf), e.put(hd)
        e.L.td_dev_sync()
        e.release(hf)
        e.release(hd)
        print(f"# data: gen + h2d {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        self.rows = n
        # fact fk + x read, dim pk + w read, joined rows (fk, x, w) written
        self.query_bytes = n * 16 + m * 16 + n * 24
        self.config = {"workload": f"inner join fact {n:,} (fk, x) x dim {m:,} (pk, w) on "
                                   "fk = pk, scattered 64-bit keys, every fact row matching "
                                   "one dim row -> joined table (fk, x, w)"}

    def step(self):
        e = self.e
        g = e.graph(self.tf)
        root = g.join(g.const_table(self.tf), [g.scan("fk")], g.const_table(self.td),
                      [g.scan("pk")], 0)
        out = g.execute(root)
        g.free()
        e.release(out)

    def kernel_bytes(self, name):
        n, m = self.a.join_fact, self.a.join_dim
        # materialising radix join: the histogram runs once per side, the
        # scatter twice per side (4,096 partitions = two 64-way passes);
        # bytes are per launch on average
        if name == "rjoin_scatter":
            # fact: fk + x read and written per pass (the second pass writes
            # the output's left columns); dim: pk + w read, w + key word
            # written, then both read and written again
            return (n * 64 + m * 64) // 4
        if name == "rjoin_hist":
            return (n * 8 + m * 8) // 2        # the key column read
        if name == "rjoin_build_probe":
            # dim key words + w read once; fact keys read, w written, and the
            # inner join's per-row match flag written
            return m * 16 + n * (8 + 8 + 1)
        if name == "join_build_probe":
            # both sides' partition entries (64-bit key + 32-bit row) read by
            # each of the two launches (count, fill); the fill launch writes
            # the (left, right) row pairs: half of that per launch on average
            return (n + m) * 12 + n * 16 // 2
        if name == "gather":
            return n * (8 + 8 + 8)             # row index and element read, element written
        return None

    def cpu_baseline(self):
        return None

    def release(self):
        self.e.release(self.tf)
        self.e.release(self.td)


class Parted:
    """A date-partitioned table: P daily partitions of the high-cardinality
    recipe's rows (key, sym, v_i64, v_f64) saved in the reference's on-disk
    layout (page cache), loaded straight into HBM (td_dev_parted_load, each
    column's partitions back to back in one block), then per step
    GROUP(date; SUM v_i64, SUM v_f64, MIN v_f64, MAX v_i64, COUNT) and
    GROUP(sym; SUM v_i64, SUM v_f64, COUNT), the date key repeated over
    each partition's rows on the device."""
    name = "parted"
    metric = "rows/s of group-bys over a date-partitioned table in HBM"

    def __init__(self, e, a):
        import shutil
        import tempfile
        from teide_amd import workloads as W
        self.e, self.a = e, a
        n, P = a.parted_rows, a.parted_parts
        nsym = 100_000
        for i in range(nsym):                     # the sym ids name real symbols
            e.sym(f"s{i:06d}")
        d = W.high_card(n, 4_000_000, nsym, 1.1, SEED)
        self.root = tempfile.mkdtemp(prefix="teide_parted_", dir="/tmp")
        t0 = time.perf_counter()
        per = n // P
        for p in range(P):
            lo, hi = p * per, (n if p == P - 1 else (p + 1) * per)
            t = e.table([("key", d["key"][lo:hi]), ("sym", d["sym"][lo:hi], 20, 2),
                         ("v_i64", d["v_i64"][lo:hi]), ("v_f64", d["v_f64"][lo:hi])])
            part = os.path.join(self.root, f"2024.{1 + p // 28:02d}.{1 + p % 28:02d}")
            os.makedirs(part)
            if e.L.td_splay_save(t, os.path.join(part, "t").encode(), None) != 0:
                raise RuntimeError("td_splay_save failed")
            e.release(t)
        del d
        if e.L.td_sym_save(os.path.join(self.root, "sym").encode()) != 0:
            raise RuntimeError("td_sym_save failed")
        t1 = time.perf_counter()
        self.t = e.L.td_dev_parted_load(self.root.encode(), b"t")
        e.L.td_dev_sync()
        t2 = time.perf_counter()
        if not self.t or self.t < 32:
            raise RuntimeError(f"td_dev_parted_load failed: {self.t}")
        self.load_gbs = n * 28 / (t2 - t1) / 1e9
        print(f"# data: save {t1 - t0:.1f}s, HBM load {t2 - t1:.2f}s", file=sys.stderr, flush=True)
        self.rows = 2 * n
        self.query_bytes = n * (4 + 8 + 8) + n * (4 + 8 + 8)
        self.config = {"workload": f"{n:,} rows in {P} daily partitions (key, sym u32 over "
                                   f"{nsym:,} symbols, v_i64, v_f64), loaded into HBM by "
                                   "td_dev_parted_load; GROUP(date; 5 aggregates) and "
                                   "GROUP(sym; 3 aggregates)",
                       "hbm_load_gbs": round(self.load_gbs, 2)}
        self._shutil = shutil

    def _group(self, w):
        e = self.e
        g = e.graph(self.t)
        vi, vf = g.scan("v_i64"), g.scan("v_f64")
        if w == 0:
            root = g.group([g.scan("date")], [("SUM", vi), ("SUM", vf), ("MIN", vf),
                                              ("MAX", vi), ("COUNT", vi)])
        else:
            root = g.group([g.scan("sym")], [("SUM", vi), ("SUM", vf), ("COUNT", vi)])
        out = g.execute(g.optimize(root))
        g.free()
        return out

    def step(self):
        self.e.release(self._group(0))
        self.e.release(self._group(1))

    def kernel_bytes(self, name):
        return None

    def cpu_baseline(self):
        exe = os.path.join(REPO, "oracle", "_ref", "ref_harness")
        if not os.path.exists(exe):
            return None
        env = dict(os.environ, TD_REF_THREADS=str(cpu_threads()))
        out = subprocess.run([exe, "bench", "partgroup", self.root, "3"], check=True,
                             capture_output=True, text=True, env=env, timeout=900)
        r = json.loads(out.stdout.strip().splitlines()[-1])
        return {"value": 2 * r["rows"] / (r["median_ms"] / 1e3), "unit": "rows/s",
                "cores": r["threads"], "kind": "reference",
                "sample": f"the same database ({r['rows']:,} rows) read by td_read_parted, "
                          "both group-bys, median of 3 after 1 warm-up"}

    def release(self):
        self.e.release(self.t)
        self._shutil.rmtree(self.root, ignore_errors=True)


class BatchQueries:
    """Many independent filter -> join -> group-by queries: fact rows
    log-uniform in [1K, 2M], dim = fact / 8 with unique keys, group
    cardinality log-uniform in [4, 65536], payloads as in the join workload.
    The step runs all of them as one batch (prepared once, replayed as a
    captured graph); the plain form (one td_execute per query) is timed
    beside it."""
    name = "batch"

    def __init__(self, e, a, scattered=False, n=None):
        from teide_amd import workloads as W
        import numpy as np
        self.e, self.a = e, a
        n = n or a.batch_queries
        nf, ncat, thr = W.batch_spec(n, SEED)
        t0 = time.perf_counter()
        self.tables, self.dags = [], []
        qbytes = 0
        for q in range(n):
            nd = max(1, int(nf[q]) // 8)
            fact = W.join_fact(int(nf[q]), nd, SEED + q)
            dim = W.join_dim(nd, int(ncat[q]), SEED + q)
            if scattered:            # the same matches on keys spread over 64-bit values
                fact["fk"] = fact["fk"] * 7_000_000_019 - (1 << 62)
                dim["pk"] = dim["pk"] * 7_000_000_019 - (1 << 62)
            hf = e.table([("fk", fact["fk"]), ("qty", fact["qty"]), ("price", fact["price"])])
            hd = e.table([("pk", dim["pk"]), ("cat", dim["cat"])])
            tf, td = e.put(hf), e.put(hd)
            e.release(hf)
            e.release(hd)
            self.tables += [tf, td]
            g = e.graph(tf)
            flt = g.filter(g.const_table(tf), g.gt(g.scan("qty"), g.const_i64(int(thr[q]))))
            g.join(flt, [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
            root = g.group([g.scan("cat")], [("SUM", g.scan("price")), ("SUM", g.scan("qty")),
                                             ("COUNT", g.scan("price"))])
            self.dags.append((g, root))
            qbytes += int(nf[q]) * 24 + nd * 12 + min(nd, int(ncat[q])) * 28
        print(f"# data: {n} queries, {int(nf.sum()):,} fact rows, {time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
        self.n = n
        self.rows = int(nf.sum())
        self.query_bytes = qbytes
        t0 = time.perf_counter()
        self.h = e.batch_prepare([g for g, _ in self.dags], [r for _, r in self.dags])
        self.prepare_s = time.perf_counter() - t0
        nb, tr, bl = e.batch_info(self.h)
        nh, pr, mb = e.batch_radix(self.h)
        keys = ("unique keys spread over 64-bit values (radix hash path)" if scattered
                else "unique keys")
        self.config = {"workload": f"{n} filter+join+group-by queries, fact rows log-uniform "
                                   f"[1K,2M] (sum {self.rows:,}), dim = fact/8 {keys}, "
                                   f"groups log-uniform [4,65536]; one batch, graph replay",
                       "batched": nb, "tile_rows": tr, "block": bl, "hash_queries": nh,
                       "part_rows": pr, "max_radix_bits": mb,
                       "prepare_s": round(self.prepare_s, 3)}

    def step(self):
        rc, res = self.e.batch_run(self.h, self.n)
        if rc:
            raise SystemExit(f"batch run failed: {rc}")
        for r in res:
            self.e.release(r)

    def plain(self, reps):
        e = self.e
        e.L.td_dev_sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            for g, root in self.dags:
                e.release(g.execute(root))
        e.L.td_dev_sync()
        return (time.perf_counter() - t0) / reps

    def kernel_bytes(self, name):
        return self.query_bytes if name == "batch_probe" else None

    def cpu_baseline(self):
        return None

    def release(self):
        self.e.L.td_dev_batch_free(self.h)
        for g, _ in self.dags:
            g.free()
        for t in self.tables:
            self.e.release(t)


class Served:
    """Request stream over resident tables: fact 100M rows and dim 8M rows
    with the join workload's distributions, plus k64 on dim (i64 uniform over
    65,536 values).  Each request is GROUP(key; SUM price, SUM qty, COUNT)
    over JOIN(FILTER(fact, qty > t), fk = pk) with t uniform in [1,100] and
    key cat or k64 at random, submitted to the request server (shared scan,
    continuous batching, admission against a 200 GB device-memory budget).
    Reported: the open-loop Poisson stream at the offered rate (achieved
    qps, p50 / p99 latency from arrival to result), the closed-loop capacity
    (a fixed number of outstanding requests), and the plain form (one
    td_execute per request) on the same tables."""
    name = "served"
    BUDGET = 200 * 10**9

    def __init__(self, e, a):
        from teide_amd import workloads as W
        import numpy as np
        self.e, self.a, self.np = e, a, np
        nf, nd = a.serve_fact, a.serve_dim
        t0 = time.perf_counter()
        fact = W.join_fact(nf, nd, SEED)
        dim = W.join_dim(nd, N_CAT, SEED)
        k64 = W.i64_uniform(nd, 0, 65535, SEED, 301)
        hf = e.table([("fk", fact["fk"]), ("qty", fact["qty"]), ("price", fact["price"])])
        del fact
        hd = e.table([("pk", dim["pk"]), ("cat", dim["cat"]), ("k64", k64)])
        del dim, k64
        self.tf, self.td = e.put(hf), e.put(hd)
        e.L.td_dev_sync()
        e.release(hf)
        e.release(hd)
        print(f"# served data: {nf:,} x {nd:,}, {time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
        # the clients' request pool: DAGs built ahead, submitted many times
        rng = np.random.default_rng(SEED + 7)
        self.pool = []
        for _ in range(a.serve_pool):
            t = int(rng.integers(1, 101))
            key = "cat" if rng.random() < 0.5 else "k64"
            self.pool.append(self._dag(key, t) + (key, t))
        # requests the cube cannot answer: price > t (f64), t from the
        # lognormal(3, 1) price distribution's quantiles 5 %..95 %
        self.wide_pool = []
        for _ in range(min(a.serve_pool, 256)):
            t = float(np.exp(3.0 + 1.6449 * rng.uniform(-1.0, 1.0)))
            key = "cat" if rng.random() < 0.5 else "k64"
            self.wide_pool.append(self._dag(key, t, wide=True) + (key, t))
        self.req_bytes = nf * 24 + nd * 12 + 2000 * 28      # one request scanned alone

    def _dag(self, key, t, wide=False):
        e, tf, td = self.e, self.tf, self.td
        g = e.graph(tf)
        pred = g.gt(g.scan("price"), g.const_f64(t)) if wide else \
            g.gt(g.scan("qty"), g.const_i64(t))
        flt = g.filter(g.const_table(tf), pred)
        g.join(flt, [g.scan("fk")], g.const_table(td), [g.scan("pk")], 0)
        root = g.group([g.scan(key)], [("SUM", g.scan("price")), ("SUM", g.scan("qty")),
                                        ("COUNT", g.scan("price"))])
        return g, root

    def _open(self):
        return self.e.serve_open(self.tf, self.td, "fk", "pk", "qty", ["price"], ["cat", "k64"],
                                 budget=self.BUDGET, chunk_rows=self.a.serve_chunk,
                                 max_active=self.a.serve_slots)

    def plain(self, n, pool=None):
        """Sequential td_execute of n pool requests: seconds per request."""
        e = self.e
        pool = pool or self.pool
        e.release(pool[0][0].execute(pool[0][1]))
        e.L.td_dev_sync()
        t0 = time.perf_counter()
        for i in range(n):
            g, root = pool[i % len(pool)][:2]
            e.release(g.execute(root))
        e.L.td_dev_sync()
        return (time.perf_counter() - t0) / n

    def stream(self, rate, seconds, seed):
        """Open loop: Poisson arrivals at `rate` for `seconds`."""
        np, e = self.np, self.e
        rng = np.random.default_rng(seed)
        n = max(1, int(rate * seconds))
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        pick = rng.integers(0, len(self.pool), n)
        lat = np.zeros(n)
        h = self._open()
        owner, done, i, steps = {}, 0, 0, 0
        t0 = time.perf_counter()
        while done < n:
            now = time.perf_counter() - t0
            while i < n and arr[i] <= now:
                g, root = self.pool[pick[i]][:2]
                owner[e.serve_submit(h, g, root)] = i
                i += 1
            if i == done:                                  # idle until the next arrival
                time.sleep(max(0.0, arr[i] - now))
                continue
            e.serve_step(h)
            steps += 1
            t = time.perf_counter() - t0
            for rid, res in e.serve_poll(h):
                k = owner.pop(rid)
                lat[k] = t - arr[k]
                if res < 32:
                    raise SystemExit(f"served request failed: {res}")
                e.release(res)
                done += 1
        end = time.perf_counter() - t0
        info = e.serve_info(h)
        e.serve_close(h)
        return {"offered_qps": rate, "requests": n, "qps": round(n / (end - arr[0]), 1),
                "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                "max_ms": round(float(lat.max()) * 1e3, 2), "steps": steps,
                "shared": info["shared"], "alone": info["alone"]}

    def closed(self, clients, seconds, pool=None):
        """Closed loop: `clients` outstanding requests, each replaced on
        completion, for about `seconds`."""
        np, e = self.np, self.e
        pool = pool or self.pool
        h = self._open()
        owner, lat, nxt = {}, [], 0
        t0 = time.perf_counter()
        for _ in range(clients):
            g, root = pool[nxt % len(pool)][:2]
            owner[e.serve_submit(h, g, root)] = 0.0
            nxt += 1
        done = 0
        while True:
            e.serve_step(h)
            t = time.perf_counter() - t0
            for rid, res in e.serve_poll(h):
                lat.append(t - owner.pop(rid))
                if res < 32:
                    raise SystemExit(f"served request failed: {res}")
                e.release(res)
                done += 1
                if t < seconds:
                    g, root = pool[nxt % len(pool)][:2]
                    owner[e.serve_submit(h, g, root)] = t
                    nxt += 1
            if not owner:
                break
        end = time.perf_counter() - t0
        info = e.serve_info(h)
        e.serve_close(h)
        lat = np.array(lat)
        return {"clients": clients, "requests": done, "qps": round(done / end, 1),
                "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                "chunks": info["chunks"], "held_gb_fixed": round(info["base"] / 1e9, 2),
                "shared": info["shared"], "alone": info["alone"]}

    def run(self):
        e = self.e
        e.prof_reset()
        e.prof(True)
        st = self.stream(self.a.serve_rate, self.a.serve_seconds, SEED + 11)
        prof = e.prof_read()
        e.prof(False)
        cap = self.closed(self.a.serve_clients, self.a.serve_cap_seconds)
        plain_s = self.plain(self.a.serve_plain)
        plain_qps = 1.0 / plain_s
        e.prof_reset()
        e.prof(True)
        wide = self.closed(self.a.serve_wide_clients, self.a.serve_cap_seconds, self.wide_pool)
        wprof = e.prof_read()
        e.prof(False)
        wide_plain = 1.0 / self.plain(max(4, self.a.serve_plain // 4), self.wide_pool)
        wide.update(workload="price > t (f64; t between the price's 5 % and 95 % quantiles), "
                             "group by cat or k64, same aggregates: each request's predicate "
                             "evaluated on the shared chunks",
                    plain_qps=round(wide_plain, 1),
                    served_vs_plain=round(wide["qps"] / wide_plain, 2),
                    kernels={k: {"launches": n, "avg_ms": round(tot / max(n, 1), 4)}
                             for k, (n, tot) in wprof.items()})
        kern = {k: {"launches": n, "avg_ms": round(tot / max(n, 1), 4)}
                for k, (n, tot) in prof.items()}
        return {"workload": f"Poisson {self.a.serve_rate}/s for {self.a.serve_seconds}s over "
                            f"resident fact {self.a.serve_fact:,} x dim {self.a.serve_dim:,}; "
                            f"qty > t (t uniform 1..100), group by cat or k64 (65,536 values); "
                            f"SUM price, SUM qty, COUNT; budget 200 GB",
                "stream": st, "capacity": cap, "plain_qps": round(plain_qps, 1),
                "served_vs_plain": round(cap["qps"] / plain_qps, 2),
                "stream_vs_plain": round(st["qps"] / plain_qps, 2),
                "effective_gbs": round(cap["qps"] * self.req_bytes / 1e9, 1),
                "kernels": kern, "wide": wide}

    def release(self):
        for g, _, _, _ in self.pool + self.wide_pool:
            g.free()
        self.e.release(self.tf)
        self.e.release(self.td)


def load_summary(e, a):
    """Splayed-table load into HBM: the join workload's fact columns saved
    in the reference's on-disk format, then loaded (file in the page cache)
    straight into the arena (td_dev_splay_load), beside the host load +
    td_dev_put path."""
    import shutil
    import tempfile
    import ctypes as C
    from teide_amd import workloads as W
    n = a.load_rows
    fact = W.join_fact(n, max(1, n // 12), SEED)
    t = e.table([("fk", fact["fk"]), ("qty", fact["qty"]), ("price", fact["price"])])
    del fact
    tmp = tempfile.mkdtemp(prefix="teide_load_", dir="/tmp")
    try:
        d, sym = os.path.join(tmp, "fact").encode(), os.path.join(tmp, "fact.sym").encode()
        if e.L.td_splay_save(t, d, sym) != 0:
            raise RuntimeError("td_splay_save failed")
        e.release(t)
        nbytes = n * 24
        dev, host = [], []
        for _ in range(a.load_reps):
            t0 = time.perf_counter()
            r = e.L.td_dev_splay_load(d, None)
            e.L.td_dev_sync()
            dev.append(time.perf_counter() - t0)
            if not r or r < 32:
                raise RuntimeError(f"td_dev_splay_load failed: {r}")
            e.release(r)
            t0 = time.perf_counter()
            h = e.L.td_splay_load(d)
            r = e.L.td_dev_put(h)
            e.L.td_dev_sync()
            host.append(time.perf_counter() - t0)
            e.release(r)
            e.release(h)
        dv, hv = sorted(dev)[len(dev) // 2], sorted(host)[len(host) // 2]
        return {"workload": f"splayed fact {n:,} rows x 3 columns (fk, qty, price), "
                            f"{nbytes / 1e9:.2f} GB, file in the page cache",
                "dev_load_gbs": round(nbytes / dv / 1e9, 2), "dev_load_s": round(dv, 4),
                "host_load_then_put_gbs": round(nbytes / hv / 1e9, 2),
                "host_load_then_put_s": round(hv, 4), "reps": a.load_reps}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def measure(e, wl, steps, warmup, prof=True):
    """Warm up, then time `steps` steps with per-kernel HIP events; the
    roofline prices the dominant kernel's algorithmic bytes (the
    workload's kernel_bytes) against its average event duration."""
    for _ in range(warmup):
        wl.step()
    plan = e.last_plan()
    e.L.td_dev_sync()
    e.prof_reset()
    e.prof(prof)
    e.L.td_dev_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        wl.step()
    e.L.td_dev_sync()
    t1 = time.perf_counter()
    pr = e.prof_read()
    e.prof(False)
    ms = (t1 - t0) * 1e3 / steps
    roofline = None
    kernels = {k: {"launches": n, "avg_ms": round(tot / max(n, 1), 5)} for k, (n, tot) in pr.items()}
    if pr:
        dom = max(pr, key=lambda k: pr[k][1])
        n, tot = pr[dom]
        avg_ms = tot / n
        nbytes = wl.kernel_bytes(dom)
        if nbytes:
            ach = nbytes / (avg_ms / 1e3) / 1e9
            roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": load_traffic(getattr(wl, "name", ""), dom), "kernel": dom,
                        "algorithmic_bytes": nbytes, "avg_ms": round(avg_ms, 5)}
    return {"ms": ms, "plan": plan, "kernels": kernels, "roofline": roofline}


def two_call_summary(e, wl, steps=5, warmup=2):
    """The reference's own form of the join + group-by, timed on the GPU
    beside the fused DAG: graph 1 runs JOIN(FILTER(fact, qty > 50), fk =
    pk) to a joined table in HBM, graph 2 runs GROUP(cat; SUM price,
    SUM qty, COUNT price) over it (bench_queries.c's shape; the CPU
    baseline times exactly this)."""
    def step():
        g = e.graph(wl.tf)
        flt = g.filter(g.const_table(wl.tf), g.gt(g.scan("qty"), g.const_i64(50)))
        root = g.join(flt, [g.scan("fk")], g.const_table(wl.td), [g.scan("pk")], 0)
        j = g.execute(root)
        g.free()
        g2 = e.graph(j)
        r = g2.group([g2.scan("cat")], [("SUM", g2.scan("price")), ("SUM", g2.scan("qty")),
                                         ("COUNT", g2.scan("price"))])
        e.release(g2.execute(r))
        g2.free()
        e.release(j)

    class TwoCall:
        name = JoinGroupTwoCall.name

        def step(self):
            step()

        def kernel_bytes(self, name):
            return JoinGroupTwoCall.kernel_bytes(wl, name)

    m = measure(e, TwoCall(), steps, warmup)
    return {"workload": "the same query as the reference runs it: join graph, then group "
                        "graph over the joined table (both on the GPU, resident)",
            "ms_per_query": round(m["ms"], 4), "rows_per_s": wl.rows / (m["ms"] / 1e3),
            "fused_vs_two_call": None, "roofline": m["roofline"], "kernels": m["kernels"]}


def scattered_summary(e, a, steps=5, warmup=2):
    """The join + group-by with its join keys spread over 64-bit values (no
    dense slot map): the radix-partitioned path, with its own roofline."""
    wl = JoinGroupScattered(e, a)
    m = measure(e, wl, steps, warmup)
    out = {"workload": wl.config["workload"], "ms_per_query": round(m["ms"], 4),
           "rows_per_s": wl.rows / (m["ms"] / 1e3),
           "query_gbs": round(wl.query_bytes / (m["ms"] / 1e3) / 1e9, 1),
           "query_frac_hbm": round(wl.query_bytes / (m["ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "roofline": m["roofline"], "plan": m["plan"], "kernels": m["kernels"]}
    wl.release()
    return out


def hc_summary(e, a, steps=5, warmup=2, cpu=True):
    """The high-cardinality group-by (300 M rows, by key and by sym) with
    its own roofline and CPU baseline."""
    wl = HighCard(e, a)
    m = measure(e, wl, steps, warmup)
    out = {"workload": wl.config["workload"], "ms_per_step": round(m["ms"], 4),
           "rows_per_s": wl.rows / (m["ms"] / 1e3),
           "query_gbs": round(wl.query_bytes / (m["ms"] / 1e3) / 1e9, 1),
           "query_frac_hbm": round(wl.query_bytes / (m["ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "roofline": m["roofline"], "plan": m["plan"], "kernels": m["kernels"]}
    wl.release()
    out["cpu_baseline"] = run_form(wl.cpu_baseline) if cpu else None
    return out


def run_form(fn):
    try:
        return fn()
    except Exception as ex:                       # reported in the line, not fatal
        return {"error": f"{type(ex).__name__}: {ex}"}


def batched_summary(e, a, steps=5, warmup=2, scattered=False, n=None):
    """The batched form on its own workload (4096 queries), beside the
    plain form (one td_execute per query); `scattered`: the join keys spread
    over 64-bit values, so every query takes the radix hash path."""
    t0 = time.perf_counter()
    wl = BatchQueries(e, a, scattered, n)
    for _ in range(warmup):
        wl.step()
    e.L.td_dev_sync()
    t1 = time.perf_counter()
    for _ in range(steps):
        wl.step()
    e.L.td_dev_sync()
    ms = (time.perf_counter() - t1) * 1e3 / steps
    plain_s = wl.plain(1)
    out = {"workload": wl.config["workload"], "queries": wl.n, "ms_per_batch": round(ms, 3),
           "batched_qps": round(wl.n / (ms / 1e3), 1), "plain_qps": round(wl.n / plain_s, 1),
           "batched_vs_plain": round(plain_s * 1e3 / ms, 2),
           "rows_per_s": wl.rows / (ms / 1e3),
           "tile_rows": wl.config["tile_rows"], "block": wl.config["block"],
           "hash_queries": wl.config["hash_queries"], "part_rows": wl.config["part_rows"],
           "max_radix_bits": wl.config["max_radix_bits"], "setup_s": round(t1 - t0, 1)}
    wl.release()
    return out


def served_summary(e, a):
    wl = Served(e, a)
    r = wl.run()
    wl.release()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fact", type=int, default=N_FACT)
    ap.add_argument("--dim", type=int, default=N_DIM)
    ap.add_argument("--cats", type=int, default=N_CAT)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="join workload only: skip the batched and served forms")
    ap.add_argument("--workload", choices=["jg", "jg2", "jgs", "hc", "batch", "batchs", "served", "load",
                                           "charge", "join", "parted"],
                    default="jg",
                    help="jg: join + group-by (default); jg2: the same query as two graphs "
                         "(join, then group-by over the joined table); jgs: jg with scattered "
                         "64-bit join keys; hc: high-cardinality group-by; "
                         "batch: many small join + group-by queries per batch; "
                         "batchs: the same with scattered keys (radix hash path); "
                         "served: request stream through the request server; "
                         "load: splayed-table load into HBM")
    ap.add_argument("--load-rows", type=int, default=50_000_000)
    ap.add_argument("--expr-rows", type=int, default=250_000_000)
    ap.add_argument("--join-fact", type=int, default=100_000_000)
    ap.add_argument("--join-dim", type=int, default=10_000_000)
    ap.add_argument("--parted-rows", type=int, default=120_000_000)
    ap.add_argument("--parted-parts", type=int, default=30)
    ap.add_argument("--load-reps", type=int, default=3)
    ap.add_argument("--batch-queries", type=int, default=4096)
    ap.add_argument("--batch-hash-queries", type=int, default=1024,
                    help="queries of the batched form's scattered-key (radix hash) run")
    ap.add_argument("--serve-fact", type=int, default=100_000_000)
    ap.add_argument("--serve-dim", type=int, default=8_000_000)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="offered requests/s")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-clients", type=int, default=2048,
                    help="outstanding requests of the closed-loop capacity run")
    ap.add_argument("--serve-cap-seconds", type=float, default=3.0)
    ap.add_argument("--serve-wide-clients", type=int, default=256,
                    help="outstanding requests of the closed-loop run with price > t filters")
    ap.add_argument("--serve-chunk", type=int, default=0, help="fact rows per scan step")
    ap.add_argument("--serve-slots", type=int, default=4096)
    ap.add_argument("--serve-pool", type=int, default=512, help="distinct request DAGs")
    ap.add_argument("--serve-plain", type=int, default=40, help="plain requests timed")
    ap.add_argument("--hc-rows", type=int, default=300_000_000)
    ap.add_argument("--hc-values", type=int, default=64_000_000)
    ap.add_argument("--hc-syms", type=int, default=1_000_000)
    ap.add_argument("--join-path", type=int, default=0,
                    help="0 planner's choice, 1 direct slot map, 3 radix-partitioned")
    ap.add_argument("--no-prof", action="store_true",
                    help="no per-kernel HIP events in the timed steps")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device copy of MIB MiB (the known-size kernel that "
                         "scales rocprof byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    from teide_amd.td import engine
    e = engine()
    if e.L.td_dev_available() != 1:
        raise SystemExit("bench.py: no usable HIP device")

    if a.calibrate_copy:
        calibrate_copy(e, a.calibrate_copy << 20)
    e.force_paths(a.join_path, 0, 0)

    if a.workload == "load":
        r = load_summary(e, a)
        line = {"metric": "splayed table file -> HBM load rate", "value": r["dev_load_gbs"],
                "unit": "GB/s", "n_gpus": 1, "steps": r["reps"], "warmup": 0,
                "ms_per_step": round(r["dev_load_s"] * 1e3, 3), "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "i64/f64",
                "data": "synthetic (seeded generator, saved in the reference's column format)",
                "config": {"workload": r.pop("workload")}, "roofline": None, "cpu_baseline": None}
        line.update(r)
        print(json.dumps(line), flush=True)
        return

    if a.workload == "served":
        wl = Served(e, a)
        r = wl.run()
        wl.release()
        st = r["stream"]
        line = {"metric": "served qps + p50/p99 latency (request stream)",
                "value": st["qps"], "unit": "requests/s", "n_gpus": 1,
                "steps": st["steps"], "warmup": 0,
                "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "i64/f64",
                "data": "synthetic (seeded generator, tables resident in HBM)",
                "config": {"workload": r.pop("workload")}, "roofline": None,
                "cpu_baseline": None}
        line.update(r)
        print(json.dumps(line), flush=True)
        return

    wl = {"hc": HighCard, "batch": BatchQueries, "charge": Charge, "jgs": JoinGroupScattered,
          "batchs": lambda e, a: BatchQueries(e, a, scattered=True),
          "jg2": JoinGroupTwoCall, "parted": Parted, "join": Join}.get(a.workload,
                                                                       JoinGroup)(e, a)

    m = measure(e, wl, a.steps, a.warmup, not a.no_prof)
    ms, plan, kernels, roofline = m["ms"], m["plan"], m["kernels"], m["roofline"]
    cpu = None if a.no_cpu_baseline else wl.cpu_baseline()
    query_gbs = wl.query_bytes / (ms / 1e3) / 1e9
    extra = {}
    if a.workload in ("batch", "batchs"):
        plain_s = wl.plain(1)
        extra = {"batched_qps": round(wl.n / (ms / 1e3), 1),
                 "plain_qps": round(wl.n / plain_s, 1),
                 "batched_vs_plain": round((plain_s * 1e3) / ms, 2)}
    line = {
        "metric": getattr(wl, "metric",
                          "rows/s + % HBM peak for join+group-by; batched/served qps & p99 vs plain"),
        "value": wl.rows / (ms / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i64/f64",
        "data": "synthetic (seeded generator, inputs resident in HBM; integer column "
                "min/max statistics computed once at td_dev_put are reused by every step "
                "where the reference scans key ranges per query)",
        "config": dict(wl.config, plan=plan),
        "query_gbs": round(query_gbs, 1),
        "query_frac_hbm": round(query_gbs / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    line.update(extra)
    if a.workload == "jg" and not a.no_forms:
        tc = run_form(lambda: two_call_summary(e, wl))
        if "ms_per_query" in tc:
            tc["fused_vs_two_call"] = round(tc["ms_per_query"] / ms, 2)
        line["two_call"] = tc
    wl.release()
    if a.workload == "jg" and not a.no_forms:
        line["scattered"] = run_form(lambda: scattered_summary(e, a))
        line["hc"] = run_form(lambda: hc_summary(e, a, cpu=not a.no_cpu_baseline))
        line["batched"] = run_form(lambda: batched_summary(e, a))
        line["batched_hash"] = run_form(lambda: batched_summary(e, a, scattered=True,
                                                              n=a.batch_hash_queries))
        line["served"] = run_form(lambda: served_summary(e, a))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
