"""tdd_minmax_noalloc, the allocation-free min / max, called through the C
ABI against tdd_minmax (the entry point it is a variant of) and numpy, on
two calls in a row with the second input smaller, for several column
types; the arena's bytes in use do not move across the calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Col(C.Structure):
    """tdd_col_t (tdd.h)."""
    _fields_ = [("data", C.c_void_p), ("type", C.c_int8), ("attrs", C.c_uint8),
                ("esz", C.c_uint8), ("ranged", C.c_uint8), ("pad", C.c_uint8 * 4),
                ("lo", C.c_int64), ("hi", C.c_int64)]


def dev_col(eng, v, type_tag, esz):
    """tdd_col_t of a resident vector (its td_data holds the device address),
    without column statistics so that the kernels run."""
    dptr = C.c_void_p.from_address(v + 32).value
    return Col(dptr, type_tag, 0, esz, 0, (C.c_uint8 * 4)(), 0, 0)


def call(fn, col, n):
    mn, mx = C.c_int64(), C.c_int64()
    rc = fn(C.byref(col), C.c_int64(n), None, C.byref(mn), C.byref(mx))
    return rc, mn.value, mx.value


@pytest.mark.parametrize("dtype", ["i64", "i32", "i16"])
def test_minmax_noalloc_matches_minmax(eng, dtype):
    L = eng.L
    for name in ("tdd_minmax", "tdd_minmax_noalloc"):
        getattr(L, name).argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        getattr(L, name).restype = C.c_int
    rng = np.random.default_rng(len(dtype))
    dt, tag = {"i64": (np.int64, 6), "i32": (np.int32, 5), "i16": (np.int16, 4)}[dtype]
    info = np.iinfo(dt)
    sizes = [5_000_003, 1_001]                         # the second call smaller
    for rep, n in enumerate(sizes):
        a = rng.integers(info.min // 2, info.max // 2, n).astype(dt)
        a[rng.integers(0, n)] = info.min + rep             # planted extremes
        a[rng.integers(0, n)] = info.max - rep
        v = eng.put(eng.vec(a))
        col = dev_col(eng, v, tag, a.itemsize)
        eng.L.td_dev_sync()
        before = eng.mem_stats()["in_use"]
        got = call(L.tdd_minmax_noalloc, col, n)
        eng.L.td_dev_sync()
        assert eng.mem_stats()["in_use"] == before       # nothing allocated
        ref = call(L.tdd_minmax, col, n)
        assert got == ref, (dtype, n, got, ref)
        assert got == (0, int(a.min()), int(a.max())), (dtype, n, got)
        # a prefix of the same column: fewer rows, fewer partials
        got = call(L.tdd_minmax_noalloc, col, 7)
        assert got == (0, int(a[:7].min()), int(a[:7].max()))
        eng.release(v)
