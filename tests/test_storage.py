"""Column files and splayed tables (store/col.c, store/splay.c): the
reference's own files (tests/golden/splay_hc3k, written by the reference's
td_splay_save) load through td_splay_load / td_col_load with the generator's
values, and td_col_save / td_splay_save write them back byte for byte.  The
splayed forms resolve column names through the saved symbol table, so they
run in a fresh process whose symbol table starts empty, as the reference's
loader expects."""
import filecmp
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from teide_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
SPLAY = os.path.join(GOLD, "splay_hc3k")
SYMS = os.path.join(GOLD, "splay_hc3k.sym")
COLS = ["key", "sym", "v_i64", "v_f64"]


def data():
    argv = json.load(open(os.path.join(GOLD, "manifest.json")))["splay_hc3k"]["argv"]
    return W.high_card(int(argv[1]), int(argv[2]), int(argv[3]), 1.1, int(argv[4]))


def fresh(code, tmp_path):
    """Run `code` in a new interpreter with a host-only engine bound to e."""
    prog = ("import sys, ctypes as C, numpy as np\n"
            f"sys.path[:0] = [{REPO!r}, {HERE!r}]\n"
            "from teide_amd.td import Engine, LIB_PATH\n"
            "e = Engine(LIB_PATH, device=False)\n" + code)
    r = subprocess.run([sys.executable, "-c", prog], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_splay_load_reference_files(tmp_path):
    out = fresh(f"""
assert e.L.td_sym_load({SYMS.encode()!r}) == 0
t = e.L.td_splay_load({SPLAY.encode()!r})
assert t and t > 32, t
for name, typ, attrs, vals in e.read_table(t):
    np.save(name + ".npy", np.asarray(vals))
    print(name, typ)
""", tmp_path)
    assert out.split() == ["key", "6", "sym", "20", "v_i64", "6", "v_f64", "7"]
    d = data()
    for c in COLS:
        np.testing.assert_array_equal(np.load(tmp_path / (c + ".npy")), d[c], err_msg=c)


def test_splay_save_matches_reference_bytes(tmp_path):
    d = data()
    for c in COLS:
        np.save(tmp_path / (c + ".npy"), d[c])
    fresh(f"""
cols = {{c: np.load(c + ".npy") for c in {COLS!r}}}
t = e.table([("key", cols["key"]), ("sym", cols["sym"], 20, 2), ("v_i64", cols["v_i64"]),
             ("v_f64", cols["v_f64"])])
assert e.L.td_splay_save(t, b"out", b"out.sym") == 0
""", tmp_path)
    for f in [".d"] + COLS:
        assert filecmp.cmp(os.path.join(SPLAY, f), tmp_path / "out" / f, shallow=False), f
    assert filecmp.cmp(SYMS, tmp_path / "out.sym", shallow=False)


def test_col_load_and_errors(tmp_path):
    from teide_amd.td import Engine, LIB_PATH
    e = Engine(LIB_PATH, device=False)
    d = data()
    for c in COLS:
        v = e.L.td_col_load(os.path.join(SPLAY, c).encode())
        assert v and v > 32
        np.testing.assert_array_equal(e.read_vec(v), d[c], err_msg=c)
        e.release(v)
    raw = open(os.path.join(SPLAY, "v_i64"), "rb").read()
    (tmp_path / "short").write_bytes(raw[:1000])              # truncated data
    (tmp_path / "tiny").write_bytes(raw[:20])                  # truncated header
    (tmp_path / "strl").write_bytes(b"STRL" + b"\0" * 40)      # string-list form
    assert e.L.td_col_load(str(tmp_path / "short").encode()) == 10   # TD_ERR_CORRUPT
    assert e.L.td_col_load(str(tmp_path / "tiny").encode()) == 10
    assert e.L.td_col_load(str(tmp_path / "strl").encode()) == 7     # TD_ERR_NYI
    assert e.L.td_col_load(str(tmp_path / "missing").encode()) == 8  # TD_ERR_IO
    # a column written here reads back, header fields as the reference sets them
    v = e.vec(d["v_f64"][:777])
    path = str(tmp_path / "rt").encode()
    assert e.L.td_col_save(v, path) == 0
    hdr = open(path, "rb").read(32)
    assert hdr[:18] == b"\0" * 18 and hdr[18] == 7 and hdr[20:24] == b"\0" * 4
    assert int.from_bytes(hdr[24:32], "little") == 777
    w = e.L.td_col_load(path)
    np.testing.assert_array_equal(e.read_vec(w), d["v_f64"][:777])
    e.release(v)
    e.release(w)
