"""GPU side of the loaders: the reference's column files and splayed table
(tests/golden/splay_hc3k) streamed straight into the HBM arena through
pinned staging buffers (td_dev_col_load / td_dev_splay_load) hold the
generator's values, carry the integer column statistics that td_dev_put
computes, and give the same query results as tables put from host memory."""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import workload_dags as D
from parity import compare_groups
from teide_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
SPLAY = os.path.join(GOLD, "splay_hc3k")
SYMS = os.path.join(GOLD, "splay_hc3k.sym")
ARGV = json.load(open(os.path.join(GOLD, "manifest.json")))["splay_hc3k"]["argv"]


def data():
    return W.high_card(int(ARGV[1]), int(ARGV[2]), int(ARGV[3]), 1.1, int(ARGV[4]))


def hbm_stats(v):
    """(lo, hi, ranged) of a resident vector's column statistics."""
    lo = C.c_int64.from_address(v + 32 + 16).value
    hi = C.c_int64.from_address(v + 32 + 24).value
    ranged = C.c_int32.from_address(v + 32 + 32).value
    return lo, hi, ranged


def test_dev_col_load(eng):
    d = data()
    for c in ["key", "sym", "v_i64", "v_f64"]:
        v = eng.L.td_dev_col_load(os.path.join(SPLAY, c).encode())
        assert v and v > 32, (c, v)
        assert eng.hdr(v).mmod == 3                              # resident in HBM
        np.testing.assert_array_equal(eng.read_vec(v), d[c], err_msg=c)
        lo, hi, ranged = hbm_stats(v)
        if c == "v_f64":
            assert ranged == 0
        else:
            assert ranged == 1 and (lo, hi) == (int(d[c].min()), int(d[c].max())), c
        eng.release(v)
    assert eng.L.td_dev_col_load(os.path.join(SPLAY, "nope").encode()) == 8      # TD_ERR_IO


def test_dev_splay_load_queries(eng, tmp_path):
    """In a fresh process (the saved column name ids resolve through the
    saved symbol table): load the splayed table into HBM, group it by key
    and by sym, and write the results next to the host-put ones."""
    prog = f"""
import sys, numpy as np
sys.path[:0] = [{REPO!r}, {HERE!r}]
from teide_amd.td import Engine, LIB_PATH
import workload_dags as D
e = Engine(LIB_PATH)
assert e.L.td_dev_available() == 1
t = e.L.td_dev_splay_load({SPLAY.encode()!r}, {SYMS.encode()!r})
assert t and t > 32, t
assert e.L.td_dev_is_resident(t) == 1
for name, typ, attrs, vals in e.read_table(t):
    np.save(name + ".npy", np.asarray(vals))
for by in ("key", "sym"):
    out = e.read_table(D.hc_group(e, t, by))
    np.savez(by + "_grp.npz", *[np.asarray(c[3]) for c in out])
print("ok", e.last_plan())
"""
    r = subprocess.run([sys.executable, "-c", prog], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    d = data()
    for c in ["key", "sym", "v_i64", "v_f64"]:
        np.testing.assert_array_equal(np.load(tmp_path / (c + ".npy")), d[c], err_msg=c)
    t, _ = D.hc_table(eng, int(ARGV[1]), int(ARGV[2]), int(ARGV[3]), int(ARGV[4]))
    rt = eng.put(t)
    eng.release(t)
    for by in ("key", "sym"):
        got = np.load(tmp_path / (by + "_grp.npz"))
        got = [got[f"arr_{i}"] for i in range(len(got.files))]
        ref = eng.read_table(D.hc_group(eng, rt, by))
        f64 = [False, False, False, True, False, False, False]
        compare_groups([got[0].astype(np.int64)], got[1:], [ref[0][3].astype(np.int64)],
                       [c[3] for c in ref[1:]], f64, by)
    eng.release(rt)
