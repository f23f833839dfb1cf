"""GPU parity on partitioned tables (store/part.c): the reference's
partitioned database tests/golden/parted_8k_db grouped by its partition
key (the MAPCOMMON date column), by a parted column and by both, against
the reference's own results of the same group-bys over td_read_parted
(tests/golden/parted_8k_*.tdt), for the table read on the host, put on the
device (segments copied back to back into one block) and loaded straight
into HBM (td_dev_parted_load); plus a filter, a head and a sort over the
parted table against the flat rows.  The symbol file is loaded with the
table, so each case runs in a fresh process."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from parity import compare_groups
from tdtio import read_tdt
from test_storage import PDB, parted_data

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")

GROUPS = """
def groups(e, t):
    import workload_dags as D
    out = []
    for w in range(3):
        g = e.graph(t)
        vi, vf = g.scan("v_i64"), g.scan("v_f64")
        if w == 0:
            root = g.group([g.scan("date")], [("SUM", vi), ("SUM", vf), ("MIN", vf), ("MAX", vi),
                                              ("COUNT", vi)])
        elif w == 1:
            root = g.group([g.scan("sym")], [("SUM", vi), ("SUM", vf), ("COUNT", vi)])
        else:
            root = g.group([g.scan("date"), g.scan("sym")], [("COUNT", vi), ("SUM", vi)])
        r = g.execute(g.optimize(root))
        out.append(e.read_table(r))
        e.release(r)
        g.free()
    return out
"""


def run_fresh(body, tmp_path):
    prog = (f"import sys, json, numpy as np\nsys.path[:0] = [{REPO!r}, {HERE!r}]\n"
            "from teide_amd.td import Engine, LIB_PATH\n"
            "e = Engine(LIB_PATH)\nassert e.L.td_dev_available() == 1\n" + GROUPS + body)
    print("parted case in a fresh process", flush=True)
    r = subprocess.run([sys.executable, "-c", prog], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-4000:]
    return r.stdout


def save_groups_code(var):
    return f"""
for w, tab in enumerate(groups(e, {var})):
    np.savez(f"g{{w}}.npz", *[np.asarray(c[3]) for c in tab])
    json.dump([[c[0], c[1]] for c in tab], open(f"g{{w}}.json", "w"))
"""


def check_groups(tmp_path):
    ref = [read_tdt(os.path.join(GOLD, f)) for f in
           ("parted_8k_date.tdt", "parted_8k_sym.tdt", "parted_8k_date_sym.tdt")]
    f64 = [[False, True, False, False, False], [False, True, False], [False, False]]
    for w in range(3):
        names = json.load(open(tmp_path / f"g{w}.json"))
        z = np.load(tmp_path / f"g{w}.npz")
        got = [z[f"arr_{i}"] for i in range(len(z.files))]
        assert [n for n, _ in names] == [c[0] for c in ref[w]], w
        assert [t for _, t in names] == [c[1] for c in ref[w]], w
        nk = 2 if w == 2 else 1
        compare_groups([g.astype(np.int64) for g in got[:nk]], got[nk:],
                       [c[3].astype(np.int64) for c in ref[w][:nk]], [c[3] for c in ref[w][nk:]],
                       f64[w], f"parted group {w}")


def test_parted_host_groups(eng, tmp_path):
    run_fresh(f"""
t = e.L.td_read_parted({PDB.encode()!r}, b"t")
assert t and t > 32, t
assert e.L.td_dev_is_resident(t) == 0
""" + save_groups_code("t"), tmp_path)
    check_groups(tmp_path)


def test_parted_put_groups(eng, tmp_path):
    run_fresh(f"""
h = e.L.td_read_parted({PDB.encode()!r}, b"t")
t = e.put(h)
e.release(h)
assert e.L.td_dev_is_resident(t) == 1
""" + save_groups_code("t"), tmp_path)
    check_groups(tmp_path)


def test_parted_dev_load_groups_and_rows(eng, tmp_path):
    """Loaded straight into HBM: the groups, and a filter, a head and a sort
    of the table (the date column repeated over each partition's rows)."""
    run_fresh(f"""
t = e.L.td_dev_parted_load({PDB.encode()!r}, b"t")
assert t and t > 32, t
assert e.L.td_dev_is_resident(t) == 1
""" + save_groups_code("t") + """
g = e.graph(t)
r = g.execute(g.filter(g.const_table(t), g.gt(g.scan("v_i64"), g.const_i64(0))))
np.savez("flt.npz", *[np.asarray(c[3]) for c in e.read_table(r)])
e.release(r); g.free()
g = e.graph(t)
r = g.execute(g.head(g.const_table(t), 700))
np.savez("head.npz", *[np.asarray(c[3]) for c in e.read_table(r)])
e.release(r); g.free()
g = e.graph(t)
r = g.execute(g.sort(g.const_table(t), [g.scan("v_f64")], [True]))
np.savez("sort.npz", *[np.asarray(c[3]) for c in e.read_table(r)])
e.release(r); g.free()
""", tmp_path)
    check_groups(tmp_path)
    d, sizes = parted_data()
    date = np.repeat(8766 + np.arange(len(sizes)), sizes).astype(np.int32)
    cols = [date, d["key"], d["sym"], d["v_i64"], d["v_f64"]]
    keep = d["v_i64"] > 0
    z = np.load(tmp_path / "flt.npz")
    for i, c in enumerate(cols):
        np.testing.assert_array_equal(z[f"arr_{i}"], c[keep], err_msg=f"filter col {i}")
    z = np.load(tmp_path / "head.npz")
    for i, c in enumerate(cols):
        np.testing.assert_array_equal(z[f"arr_{i}"], c[:700], err_msg=f"head col {i}")
    import orc
    p = orc.sort_perm([d["v_f64"]], [True])
    z = np.load(tmp_path / "sort.npz")
    for i, c in enumerate(cols):
        np.testing.assert_array_equal(z[f"arr_{i}"], c[p], err_msg=f"sort col {i}")


def test_parted_many_sym_partitions(eng, tmp_path):
    """1,100 partitions named 1.0000 .. 1.1099 (not dates: a SYM MAPCOMMON
    column of 8-byte ids, partitions beyond what the fill stages in LDS),
    of 1-7 rows each: the partition column of the flattened table repeats
    each partition's id over its rows, in name order, and a group-by over it
    gives each partition's row count and sum."""
    body = """
import os
rng = np.random.default_rng(11)
sizes = rng.integers(1, 8, 1100)
vals = rng.integers(-10 ** 6, 10 ** 6, int(sizes.sum()))
os.makedirs("db")
off = 0
for p, s in enumerate(sizes):
    os.makedirs(f"db/1.{p:04d}")
    h = e.table([("v", vals[off:off + s])])
    assert e.L.td_splay_save(h, f"db/1.{p:04d}/t".encode(), b"db/sym") == 0
    e.release(h)
    off += s
t = e.L.td_dev_parted_load(b"db", b"t")
assert t and t > 32, t
g = e.graph(t)
r = g.execute(g.head(g.const_table(t), int(sizes.sum())))
np.savez("rows.npz", *[np.asarray(c[3]) for c in e.read_table(r)])
e.release(r); g.free()
g = e.graph(t)
r = g.execute(g.group([g.scan("part")], [("SUM", g.scan("v")), ("COUNT", g.scan("v"))]))
np.savez("grp.npz", *[np.asarray(c[3]) for c in e.read_table(r)])
e.release(r); g.free()
np.savez("src.npz", sizes=sizes, vals=vals)
"""
    prog = (f"import sys, json, numpy as np\nsys.path[:0] = [{REPO!r}, {HERE!r}]\n"
            "from teide_amd.td import Engine, LIB_PATH\n"
            "e = Engine(LIB_PATH)\nassert e.L.td_dev_available() == 1\n" + body)
    print("parted many-partition case in a fresh process", flush=True)
    r = subprocess.run([sys.executable, "-c", prog], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-4000:]
    src = np.load(tmp_path / "src.npz")
    sizes, vals = src["sizes"], src["vals"]
    rows = np.load(tmp_path / "rows.npz")
    part, v = rows["arr_0"].astype(np.int64), rows["arr_1"]
    np.testing.assert_array_equal(v, vals)
    ids = part[np.concatenate([[0], np.cumsum(sizes)[:-1]])]      # first row of each partition
    assert len(np.unique(ids)) == len(sizes)
    assert np.all(np.diff(ids) > 0), "ids follow the name order in which they were interned"
    np.testing.assert_array_equal(part, np.repeat(ids, sizes))
    grp = np.load(tmp_path / "grp.npz")
    gk = grp["arr_0"].astype(np.int64)
    order = np.argsort(gk)
    pos = np.searchsorted(ids, gk[order])
    np.testing.assert_array_equal(gk[order], ids[pos])
    np.testing.assert_array_equal(grp["arr_2"][order], sizes[pos])
    sums = np.add.reduceat(vals, np.concatenate([[0], np.cumsum(sizes)[:-1]]))
    np.testing.assert_array_equal(grp["arr_1"][order], sums[pos])
    assert len(gk) == len(sizes)
