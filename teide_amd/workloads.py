"""Synthetic workloads for the filter -> join -> group-by path.

Thin ctypes front end of teide_amd/csrc/datagen.c (libtdgen.so): every column
is a pure function of (sizes, seed), identical to what the reference harness
generates for the same arguments.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtdgen.so")
NTHREADS = min(16, os.cpu_count() or 1)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P, I64, U64, I32, D = C.c_void_p, C.c_int64, C.c_uint64, C.c_int, C.c_double
        L.tdg_recipe_fg.argtypes = [I64, U64, I32, P, P, P]
        L.tdg_recipe_dim.argtypes = [I64, I64, U64, C.c_uint32, I32, P, P]
        L.tdg_recipe_fact.argtypes = [I64, I64, U64, C.c_uint32, I32, P, P, P]
        L.tdg_recipe_hc.argtypes = [I64, I64, D, C.c_uint32, U64, I32, P, P, P, P]
        L.tdg_i64_uniform.argtypes = [P, I64, I64, I64, U64, C.c_uint32, I32]
        L.tdg_i32_uniform.argtypes = [P, I64, C.c_int32, C.c_int32, U64, C.c_uint32, I32]
        L.tdg_f64_uniform.argtypes = [P, I64, D, D, U64, C.c_uint32, I32]
        L.tdg_i64_permutation.argtypes = [P, I64, I64, U64, C.c_uint32]
        L.tdg_i64_pick.argtypes = [P, I64, P, I64, U64, C.c_uint32, I32]
        L.tdg_unit.restype = D
        L.tdg_unit.argtypes = [U64, U64, U64]
        L.tdg_zipf_key.restype = I64
        L.tdg_zipf_key.argtypes = [I64]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def filter_group(n, seed=42):
    """flag i64 {0,1}, region i64 [0,16), amount f64 [0,1000)."""
    flag = np.empty(n, np.int64)
    region = np.empty(n, np.int64)
    amount = np.empty(n, np.float64)
    lib().tdg_recipe_fg(n, seed, NTHREADS, _p(flag), _p(region), _p(amount))
    return dict(flag=flag, region=region, amount=amount)


def join_dim(n_dim, n_cat=1000, seed=42, stream_base=200):
    """pk = permutation of [0, n_dim), cat i32 uniform [0, n_cat)."""
    pk = np.empty(n_dim, np.int64)
    cat = np.empty(n_dim, np.int32)
    lib().tdg_recipe_dim(n_dim, n_cat, seed, stream_base, NTHREADS, _p(pk), _p(cat))
    return dict(pk=pk, cat=cat)


def join_fact(n_fact, n_dim, seed=42, stream_base=200):
    """fk uniform [0, n_dim), qty i64 [1,100], price lognormal(3,1)."""
    fk = np.empty(n_fact, np.int64)
    qty = np.empty(n_fact, np.int64)
    price = np.empty(n_fact, np.float64)
    lib().tdg_recipe_fact(n_fact, n_dim, seed, stream_base, NTHREADS, _p(fk), _p(qty),
                          _p(price))
    return dict(fk=fk, qty=qty, price=price)


def high_card(n, n_values, n_sym, s=1.1, seed=42):
    """key = zipf(s) over n_values ranks, sym u32 [0,n_sym), v_i64, v_f64."""
    key = np.empty(n, np.int64)
    sym = np.empty(n, np.uint32)
    vi = np.empty(n, np.int64)
    vf = np.empty(n, np.float64)
    lib().tdg_recipe_hc(n, n_values, s, n_sym, seed, NTHREADS, _p(key), _p(sym), _p(vi),
                        _p(vf))
    return dict(key=key, sym=sym, v_i64=vi, v_f64=vf)


def i64_uniform(n, lo, hi, seed, stream):
    out = np.empty(n, np.int64)
    lib().tdg_i64_uniform(_p(out), n, lo, hi, seed, stream, NTHREADS)
    return out


def f64_uniform(n, lo, hi, seed, stream):
    out = np.empty(n, np.float64)
    lib().tdg_f64_uniform(_p(out), n, lo, hi, seed, stream, NTHREADS)
    return out


def i32_uniform(n, lo, hi, seed, stream):
    out = np.empty(n, np.int32)
    lib().tdg_i32_uniform(_p(out), n, lo, hi, seed, stream, NTHREADS)
    return out


def permutation(n, base, seed, stream):
    out = np.empty(n, np.int64)
    lib().tdg_i64_permutation(_p(out), n, base, seed, stream)
    return out


def unit(seed, a, b):
    return lib().tdg_unit(seed, a, b)


def batch_spec(n=4096, seed=42):
    """The batched workload's per-query sizes: fact rows log-uniform in
    [1K, 2M], group-key cardinality log-uniform in [4, 65536] and the
    `qty > thr` threshold uniform in [1, 100]; query q's dim has fact/8 rows
    and its tables come from join_fact / join_dim with seed + q."""
    rng = np.random.default_rng(seed)
    nf = np.exp(rng.uniform(np.log(1_000), np.log(2_000_000), n)).astype(np.int64)
    ncat = np.exp(rng.uniform(np.log(4), np.log(65_536), n)).astype(np.int64)
    thr = rng.integers(1, 101, n)
    return nf, ncat, thr
