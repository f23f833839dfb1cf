"""Regenerate the golden fixtures under tests/golden/ from the reference.

Runs oracle/_ref/ref_harness (the reference library compiled from its own
sources by oracle/Makefile) on the shared workload recipes and stores the
reference's result tables (TDT1 files, a few KB to a few hundred KB each)
together with SHA-256 digests of every input column, so a later run can tell
a data change from a behaviour change.  Needs the reference tree to build the
harness; the fixtures it writes are committed and read by tests/ everywhere.

    python tests/golden/make_golden.py [CASE ...]

With case names, only those fixtures are rewritten and the other manifest
entries are kept as they are.
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

from teide_amd import workloads as W  # noqa: E402
from tdtio import canonical_digest, read_tdt  # noqa: E402

HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")

# name -> (harness argv, input-column generator for the digests)
CASES = {
    "fg_1m": (["fg", "1000000", "42"], lambda: W.filter_group(1000000, 42)),
    "fg_100k_s7": (["fg", "100000", "7"], lambda: W.filter_group(100000, 7)),
    "jg_2m_250k": (["jg", "2000000", "250000", "1000", "42"],
                   lambda: {**W.join_fact(2000000, 250000, 42),
                            **W.join_dim(250000, 1000, 42)}),
    "jg_100k_8k_c1m": (["jg", "100000", "8000", "1000000", "5"],
                       lambda: {**W.join_fact(100000, 8000, 5),
                                **W.join_dim(8000, 1000000, 5)}),
    "jg_expr_400k": (["jgexpr", "400000", "50000", "300", "9"],
                     lambda: {**W.join_fact(400000, 50000, 9), **W.join_dim(50000, 300, 9)}),
    "hc_60k": (["hc", "60000", "4096", "2048", "42"],
               lambda: W.high_card(60000, 4096, 2048, 1.1, 42)),
    "having_60k": (["having", "60000", "4096", "2048", "42", "29", "0.0", "10"],
                   lambda: W.high_card(60000, 4096, 2048, 1.1, 42)),
    "sort_8k": (["sort", "8000", "4096", "2048", "42", "25"],
                lambda: W.high_card(8000, 4096, 2048, 1.1, 42)),
    "expr_grp_60k": (["expr", "60000", "4096", "2048", "42", "grp"],
                     lambda: W.high_card(60000, 4096, 2048, 1.1, 42)),
    "expr_aff_60k": (["expr", "60000", "4096", "2048", "42", "aff"],
                     lambda: W.high_card(60000, 4096, 2048, 1.1, 42)),
    "expr_scalar_60k": (["expr", "60000", "4096", "2048", "42", "scalar"],
                        lambda: W.high_card(60000, 4096, 2048, 1.1, 42)),
    "expr_vec_4k": (["expr", "4000", "4096", "2048", "42", "vec"],
                    lambda: W.high_card(4000, 4096, 2048, 1.1, 42)),
    "parted_8k": (["partsave", "8000", "4096", "2048", "42", "5"],
                  lambda: W.high_card(8000, 4096, 2048, 1.1, 42)),
    "splay_hc3k": (["save", "3000", "4096", "2048", "42"],
                   lambda: W.high_card(3000, 4096, 2048, 1.1, 42)),
    "join_inner_small": (["join", "3000", "2000", "2500", "0", "42"], None),
    "join_left_small": (["join", "3000", "2000", "2500", "1", "42"], None),
    "join_full_small": (["join", "3000", "2000", "2500", "2", "42"], None),
    "join_inner_radix": (["join", "90000", "70000", "150000", "0", "11"], None),
    "join_full_radix": (["join", "90000", "70000", "150000", "2", "11"], None),
}


def digest(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: run make -C oracle ref")
    only = sys.argv[1:]
    mpath = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(mpath)) if only else {}
    for name, (argv, gen) in CASES.items():
        if only and name not in only:
            continue
        outs = []
        if argv[0] == "hc":
            outs = [f"{name}_key.tdt", f"{name}_sym.tdt"]
        elif argv[0] in ("having", "sort"):
            outs = [f"{name}.tdt", f"{name}_top.tdt"]
        elif argv[0] == "partsave":
            # a date-partitioned database written by the reference, and its
            # group-bys through td_read_parted (a second, fresh process)
            import shutil
            db = os.path.join(HERE, name + "_db")
            shutil.rmtree(db, ignore_errors=True)
            subprocess.run([HARNESS] + argv + [db], check=True)
            outs = [f"{name}_date.tdt", f"{name}_sym.tdt", f"{name}_date_sym.tdt"]
            subprocess.run([HARNESS, "partgroup", db] + [os.path.join(HERE, o) for o in outs],
                           check=True)
            tables = [read_tdt(os.path.join(HERE, o)) for o in outs]
            manifest[name] = {"argv": argv, "outputs": [name + "_db"] + outs,
                              "rows_out": [len(t[0][3]) for t in tables],
                              "inputs_sha256": {k: digest(v) for k, v in gen().items()}}
            print(name, manifest[name]["rows_out"])
            continue
        elif argv[0] == "save":
            # a splayed directory written by the reference, with its symbol
            # file: the on-disk format the loaders must read byte for byte
            import shutil
            shutil.rmtree(os.path.join(HERE, name), ignore_errors=True)
            subprocess.run([HARNESS] + argv + [os.path.join(HERE, name),
                                                os.path.join(HERE, name + ".sym")], check=True)
            manifest[name] = {"argv": argv, "outputs": [name, name + ".sym"],
                              "inputs_sha256": {k: digest(v) for k, v in gen().items()}}
            print(name, "splayed")
            continue
        else:
            outs = [f"{name}.tdt"]
        subprocess.run([HARNESS] + argv + [os.path.join(HERE, o) for o in outs], check=True)
        entry = {"argv": argv, "outputs": outs}
        if gen is not None:
            entry["inputs_sha256"] = {k: digest(v) for k, v in gen().items()}
        tables = [read_tdt(os.path.join(HERE, o)) for o in outs]
        entry["rows_out"] = [len(t[0][3]) for t in tables]
        if name.endswith("_radix"):
            # large join outputs are kept as an order-free digest only
            entry["digest"] = [canonical_digest([c[3] for c in t]) for t in tables]
            entry["columns"] = [[(c[0], c[1]) for c in t] for t in tables]
            for o in outs:
                os.remove(os.path.join(HERE, o))
            entry["outputs"] = []
        manifest[name] = entry
        print(name, entry["rows_out"])
    with open(mpath, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
