"""ctypes front end of the oracle library (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as a checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

TD_BOOL, TD_U8, TD_I16, TD_I32, TD_I64, TD_F64, TD_SYM = 1, 2, 4, 5, 6, 7, 20
OP = dict(EQ=25, NE=26, LT=27, LE=28, GT=29, GE=30, AND=31, OR=32, NOT=12,
          SUM=50, PROD=51, MIN=52, MAX=53, COUNT=54, AVG=55, FIRST=56, LAST=57,
          STDDEV=59, STDDEV_POP=73, VAR=74, VAR_POP=75)

_TYPE_OF = {np.dtype(np.uint8): (TD_U8, 0), np.dtype(np.bool_): (TD_BOOL, 0),
            np.dtype(np.int16): (TD_I16, 0), np.dtype(np.int32): (TD_I32, 0),
            np.dtype(np.int64): (TD_I64, 0), np.dtype(np.float64): (TD_F64, 0),
            np.dtype(np.uint16): (TD_SYM, 1), np.dtype(np.uint32): (TD_SYM, 2)}


class _Col(C.Structure):
    _fields_ = [("data", C.c_void_p), ("type", C.c_int8), ("attrs", C.c_uint8)]


class _GroupOut(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("vals", C.c_void_p), ("sum_abs", C.c_void_p),
                ("comp", C.c_void_p), ("out_types", C.c_int8 * 8), ("order", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_sel_payload_bytes.restype = C.c_int64
        L.orc_sel_payload_bytes.argtypes = [C.c_int64]
        L.orc_sel_from_bool.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_sel_and.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_sel_to_idx.restype = C.c_int64
        L.orc_sel_to_idx.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_cmp_const.argtypes = [_Col, C.c_int64, C.c_int, C.c_int, C.c_int64,
                                    C.c_double, C.c_void_p]
        L.orc_cmp_cols.argtypes = [_Col, _Col, C.c_int64, C.c_int, C.c_void_p]
        L.orc_bool_op.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_gather.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                 C.c_int]
        L.orc_join.restype = C.c_int64
        L.orc_join.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int64, C.c_int64,
                               C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_group.restype = C.c_int64
        L.orc_group.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                C.c_int64, C.c_void_p, C.c_int64, C.POINTER(_GroupOut)]
        _lib = L
    return _lib


def _col(a, type_tag=None, attrs=None):
    if a is None:
        return _Col(None, 0, 0)
    t, at = _TYPE_OF[a.dtype]
    if type_tag is not None:
        t = type_tag
    if attrs is not None:
        at = attrs
    return _Col(a.ctypes.data, t, at)


def _ptr(a):
    return a.ctypes.data if a is not None else None


def sel_from_bool(pred):
    pred = np.ascontiguousarray(pred, dtype=np.uint8)
    n = len(pred)
    out = np.zeros(lib().orc_sel_payload_bytes(n), np.uint8)
    lib().orc_sel_from_bool(_ptr(pred), n, _ptr(out))
    return out


def sel_and(a, b, n):
    out = np.zeros(lib().orc_sel_payload_bytes(n), np.uint8)
    lib().orc_sel_and(_ptr(a), _ptr(b), n, _ptr(out))
    return out


def sel_parts(payload, n):
    """Split a TD_SEL payload into (total, flags, popcnt, bits)."""
    nseg = (n + 1023) // 1024
    al8 = lambda x: (x + 7) & ~7
    total = int(payload[:8].view(np.int64)[0])
    flags = payload[16:16 + nseg]
    o = 16 + al8(nseg)
    pop = payload[o:o + 2 * nseg].view(np.uint16)
    o += al8(2 * nseg)
    bits = payload[o:o + 8 * ((n + 63) // 64)].view(np.uint64)
    return total, flags, pop, bits


def sel_to_idx(payload, n):
    idx = np.zeros(max(n, 1), np.int64)
    m = lib().orc_sel_to_idx(_ptr(payload), n, _ptr(idx))
    return idx[:m]


def cmp_const(col, op, const, type_tag=None, attrs=None):
    col = np.ascontiguousarray(col)
    out = np.zeros(len(col), np.uint8)
    is_f = isinstance(const, float)
    lib().orc_cmp_const(_col(col, type_tag, attrs), len(col), OP[op], int(is_f),
                        0 if is_f else int(const), float(const), _ptr(out))
    return out


def cmp_cols(a, b, op):
    out = np.zeros(len(a), np.uint8)
    lib().orc_cmp_cols(_col(a), _col(b), len(a), OP[op], _ptr(out))
    return out


def bool_op(op, a, b=None):
    out = np.zeros(len(a), np.uint8)
    lib().orc_bool_op(OP[op], _ptr(a), _ptr(b if b is not None else a), len(a), _ptr(out))
    return out


def gather(src, idx, nullable=False):
    src = np.ascontiguousarray(src)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    dst = np.zeros(len(idx), src.dtype)
    lib().orc_gather(_ptr(src), src.dtype.itemsize, _ptr(idx), len(idx), _ptr(dst),
                     int(nullable))
    return dst


def join(lkeys, rkeys, join_type=0):
    """Pairs (l_idx, r_idx) of the equi-join of key column lists."""
    nk = len(lkeys)
    lc = (_Col * nk)(*[_col(np.ascontiguousarray(k)) for k in lkeys])
    rc = (_Col * nk)(*[_col(np.ascontiguousarray(k)) for k in rkeys])
    nl, nr = len(lkeys[0]), len(rkeys[0])
    m = lib().orc_join(C.addressof(lc), C.addressof(rc), nk, nl, nr, join_type,
                       None, None, 0)
    if m < 0:
        raise ValueError("unsupported key types")
    li = np.zeros(max(m, 1), np.int64)
    ri = np.zeros(max(m, 1), np.int64)
    lib().orc_join(C.addressof(lc), C.addressof(rc), nk, nl, nr, join_type,
                   _ptr(li), _ptr(ri), m)
    return li[:m], ri[:m]


def group(keys, ops, aggs, mask_bits=None, key_types=None):
    """Group-by; returns dict(keys=[arrays], vals=[arrays], sum_abs, comp,
    out_types, order).  keys/aggs are numpy arrays (aggs may hold None for
    COUNT inputs that have no column)."""
    nk, na = len(keys), len(ops)
    n = len(keys[0])
    kt = key_types or [None] * nk
    kc = (_Col * nk)(*[_col(np.ascontiguousarray(k), *(kt[i] or (None, None)))
                       for i, k in enumerate(keys)])
    ac = (_Col * max(na, 1))(*[_col(np.ascontiguousarray(a) if a is not None else None)
                               for a in aggs])
    opsa = (C.c_uint16 * max(na, 1))(*[OP[o] for o in ops])
    cap = n + 1
    kbuf = np.zeros(nk * cap, np.int64)
    vbuf = np.zeros(max(na, 1) * cap, np.uint64)
    abuf = np.zeros(max(na, 1) * cap, np.float64)
    cbuf = np.zeros(max(na, 1) * cap, np.float64)
    out = _GroupOut(_ptr(kbuf), _ptr(vbuf), _ptr(abuf), _ptr(cbuf))
    mb = np.ascontiguousarray(mask_bits, dtype=np.uint64) if mask_bits is not None else None
    g = lib().orc_group(C.addressof(kc), nk, C.addressof(opsa), C.addressof(ac), na, n,
                        _ptr(mb), cap, C.byref(out))
    if g < 0:
        raise ValueError("unsupported aggregate")
    res = dict(order=out.order, out_types=[out.out_types[a] for a in range(na)])
    res["keys"] = [kbuf[k * cap:k * cap + g].copy() for k in range(nk)]
    vals, sabs, comp = [], [], []
    for a in range(na):
        raw = vbuf[a * cap:a * cap + g].copy()
        vals.append(raw.view(np.float64) if res["out_types"][a] == TD_F64
                    else raw.view(np.int64))
        sabs.append(abuf[a * cap:a * cap + g].copy())
        comp.append(cbuf[a * cap:a * cap + g].copy())
    res["vals"], res["sum_abs"], res["comp"] = vals, sabs, comp
    return res


def sort_words(col, desc=False, nulls_first=None, rank=None):
    """Order-preserving 64-bit words of a sort key column, as the
    reference's radix path encodes them (radix_encode_fn,
    exec.c:2739-2870): I64 with the sign bit flipped; I32 / I16 the same on
    their own width; F64 by IEEE bits with negatives inverted and NaN at
    0 or all-ones by (nulls_first xor desc); SYM by rank[id] (string order);
    BOOL / U8 as is; descending keys inverted."""
    nf = desc if nulls_first is None else nulls_first
    a = np.asarray(col)
    if rank is not None:
        e = np.asarray(rank, np.uint64)[a.astype(np.int64)]
    elif a.dtype == np.float64:
        bits = a.view(np.uint64)
        mask = np.where(bits >> np.uint64(63), np.uint64(0xFFFFFFFFFFFFFFFF),
                        np.uint64(1 << 63))
        e = bits ^ mask
        nan = np.isnan(a)
        nan_e = np.uint64(0) if (bool(nf) ^ bool(desc)) else np.uint64(0xFFFFFFFFFFFFFFFF)
        e = np.where(nan, ~nan_e if desc else nan_e, ~e if desc else e)
        return e.astype(np.uint64)
    elif a.dtype == np.int64:
        e = a.view(np.uint64) ^ np.uint64(1 << 63)
    elif a.dtype == np.int32:
        e = (a.view(np.uint32) ^ np.uint32(1 << 31)).astype(np.uint64)
    elif a.dtype == np.int16:
        e = (a.view(np.uint16) ^ np.uint16(1 << 15)).astype(np.uint64)
    else:
        e = a.astype(np.uint64)
    return ~e if desc else e


def sort_perm(cols, descs, nulls_first=None, ranks=None):
    """Row order of a stable multi-key sort: keys compared in order by their
    sort_words, equal rows kept in input order."""
    nk = len(cols)
    nfs = nulls_first or [None] * nk
    rks = ranks or [None] * nk
    words = [sort_words(c, d, f, r) for c, d, f, r in zip(cols, descs, nfs, rks)]
    return np.lexsort(words[::-1], axis=-1) if nk else np.arange(0)


def sym_rank(strings):
    """rank[id] of symbol strings (bytes) ordered by their bytes then
    length, ties by id (build_enum_rank, exec.c:2876-2970)."""
    order = sorted(range(len(strings)), key=lambda i: (strings[i], i))
    rank = np.empty(len(strings), np.uint32)
    rank[np.asarray(order, np.int64)] = np.arange(len(strings), dtype=np.uint32)
    return rank


# ---------------------------------------------------------------- expressions
# A restatement of the reference's fused expression path: typing as
# expr_compile (exec.c:760-918), arithmetic as the morsel evaluator
# (exec.c:980-1100).  An expression is a nested tuple:
#   ("col", array)  ("i64", v)  ("f64", v)  ("cast", e, "i64"|"f64")
#   (op, e) for NEG ABS NOT SQRT LOG EXP CEIL FLOOR
#   (op, e1, e2) for ADD SUB MUL DIV MOD MIN2 MAX2 EQ NE LT LE GT GE AND OR
# and evaluates to (type, array) with type "i64", "f64" or "bool".

_CMP = {"EQ": np.equal, "NE": np.not_equal, "LT": np.less, "LE": np.less_equal,
        "GT": np.greater, "GE": np.greater_equal}
_I64_MIN = np.int64(-2 ** 63)


def _as(t, v, to):
    """expr_ensure_type's CAST (exec.c:743-757) and the unary CAST rules."""
    if t == to:
        return v
    if to == "f64":
        return v.astype(np.float64)
    if t == "f64":                                   # clamp, exec.c:1072-1077
        # NaN fails both clamps and reaches the C cast, whose x86 result
        # (cvttsd2si's integer indefinite) is INT64_MIN
        out = np.where(v >= 9.223372036854775807e18, np.iinfo(np.int64).max,
                       np.iinfo(np.int64).min)
        mid = (v < 9.223372036854775807e18) & (v > -9.223372036854775808e18)
        out = out.astype(np.int64)
        out[mid] = np.trunc(v[mid]).astype(np.int64)
        return out
    return v.astype(np.int64)


def expr_eval(e):
    """(type, values) of expression `e` over its columns' rows."""
    kind = e[0]
    if kind == "col":
        a = np.asarray(e[1])
        return ("f64", a.astype(np.float64)) if a.dtype == np.float64 else \
            ("i64", a.astype(np.int64))
    if kind == "i64":
        return "i64", np.int64(e[1])
    if kind == "f64":
        return "f64", np.float64(e[1])
    if kind == "bool":                               # a folded constant: an i64 register
        return "i64", np.int64(e[1])
    if kind == "cast":
        t, v = expr_eval(e[1])
        return e[2], _as(t, np.asarray(v), e[2])
    op = kind
    t1, a = expr_eval(e[1])
    t2, b = expr_eval(e[2]) if len(e) > 2 else (t1, None)
    if op in _CMP or op in ("AND", "OR", "NOT"):
        ot = "bool"
    elif "f64" in (t1, t2) or op in ("DIV", "SQRT", "LOG", "EXP"):
        ot = "f64"
    else:
        ot = "i64"
    if ot == "f64":
        a = _as(t1, np.asarray(a), "f64")
        if b is not None:
            b = _as(t2, np.asarray(b), "f64")
    elif ot == "bool" and b is not None and t1 != t2:
        pt = "f64" if "f64" in (t1, t2) else "i64"
        a, b = _as(t1, np.asarray(a), pt), _as(t2, np.asarray(b), pt)
    with np.errstate(all="ignore"):
        if ot == "bool":
            if op == "NOT":
                return ot, np.logical_not(a)
            if op == "AND":
                return ot, np.logical_and(a, b)
            if op == "OR":
                return ot, np.logical_or(a, b)
            return ot, _CMP[op](a, b)
        if ot == "f64":
            a, b = np.asarray(a, np.float64), (None if b is None else np.asarray(b, np.float64))
            if op in ("NEG", "ABS", "SQRT", "LOG", "EXP", "CEIL", "FLOOR"):
                f = {"NEG": np.negative, "ABS": np.abs, "SQRT": np.sqrt, "LOG": np.log,
                     "EXP": np.exp, "CEIL": np.ceil, "FLOOR": np.floor}[op]
                return ot, f(a)
            a, b = np.broadcast_arrays(a, b)
            if op == "ADD":
                return ot, a + b
            if op == "SUB":
                return ot, a - b
            if op == "MUL":
                return ot, a * b
            if op == "DIV":
                return ot, np.where(b != 0, a / np.where(b != 0, b, 1), 0.0)
            if op == "MOD":
                return ot, np.where(b != 0, np.fmod(a, np.where(b != 0, b, 1)), 0.0)
            if op == "MIN2":
                return ot, np.where(a < b, a, b)
            if op == "MAX2":
                return ot, np.where(a > b, a, b)
        # i64: two's-complement wrapping (exec.c:1003-1013, 1060-1064)
        a = np.asarray(a, np.int64)
        if op == "NEG":
            return ot, (np.uint64(0) - a.astype(np.uint64)).astype(np.int64)
        if op == "ABS":
            return ot, np.where(a < 0, (np.uint64(0) - a.astype(np.uint64)).astype(np.int64), a)
        if op in ("CEIL", "FLOOR"):
            return ot, a            # no i64 case in the reference's unary switch
        a, b = np.broadcast_arrays(a, np.asarray(b, np.int64))
        ua, ub = a.astype(np.uint64), b.astype(np.uint64)
        if op == "ADD":
            return ot, (ua + ub).astype(np.int64)
        if op == "SUB":
            return ot, (ua - ub).astype(np.int64)
        if op == "MUL":
            return ot, (ua * ub).astype(np.int64)
        if op == "MOD":
            ok = (b != 0) & ~((a == _I64_MIN) & (b == -1))
            return ot, np.where(ok, np.fmod(a, np.where(ok, b, 1)), 0).astype(np.int64)
        if op == "MIN2":
            return ot, np.where(a < b, a, b)
        if op == "MAX2":
            return ot, np.where(a > b, a, b)
    raise ValueError(f"expression op {op}")


def _const(e):
    return e[0] in ("i64", "f64", "bool")


def _num(e):
    """(int value, float value, is_f64) of a constant, as atom_to_numeric."""
    if e[0] == "f64":
        f = float(e[1])
        return (int(f) if np.isfinite(f) else 0), f, True
    i = int(e[1])
    return i, float(i), False


def _wrap(v):
    return int(np.array(v % (1 << 64), np.uint64).astype(np.int64))


def expr_fold(e):
    """The tree after the reference optimizer's constant folding
    (opt.c:305-447): element-wise nodes over constants become constants,
    typed by what the graph builders declared (DIV f64, comparisons and
    logic bool, otherwise the wider operand type) and computed with IEEE
    division, fmod, fmin / fmax for f64 and wrapping guarded arithmetic for
    i64; a negation of INT64_MIN is left to the evaluator."""
    if e[0] in ("col", "i64", "f64", "bool"):
        return e
    if e[0] == "cast":
        return ("cast", expr_fold(e[1]), e[2])
    op = e[0]
    kids = [expr_fold(c) for c in e[1:]]
    if not all(_const(k) for k in kids):
        return (op,) + tuple(kids)
    with np.errstate(all="ignore"):
        if len(kids) == 1:
            i, f, isf = _num(kids[0])
            x = f if isf else float(i)
            if op == "NEG":
                return ("f64", -f) if isf else ((op,) + tuple(kids) if i == -2 ** 63
                                                 else ("i64", _wrap(-i)))
            if op == "ABS":
                return ("f64", abs(f)) if isf else ((op,) + tuple(kids) if i == -2 ** 63
                                                     else ("i64", abs(i)))
            if op == "NOT":
                return ("bool", int(f == 0.0 if isf else i == 0))
            if op in ("SQRT", "LOG", "EXP"):
                fn = {"SQRT": np.sqrt, "LOG": np.log, "EXP": np.exp}[op]
                return ("f64", float(fn(np.float64(x))))
            if op in ("CEIL", "FLOOR"):
                return ("f64", float((np.ceil if op == "CEIL" else np.floor)(f))) if isf \
                    else ("i64", i)
            return (op,) + tuple(kids)
        (li, lf, lfl), (ri, rf, rfl) = _num(kids[0]), _num(kids[1])
        types = {k[0] for k in kids}
        if op == "DIV":
            t = "f64"
        elif op in _CMP or op in ("AND", "OR"):
            t = "bool"
        else:
            t = "f64" if "f64" in types else "i64" if "i64" in types else "bool"
        a = np.float64(lf if lfl else li)
        b = np.float64(rf if rfl else ri)
        if t == "f64":
            r = {"ADD": lambda: a + b, "SUB": lambda: a - b, "MUL": lambda: a * b,
                 "DIV": lambda: a / b, "MOD": lambda: np.fmod(a, b),
                 "MIN2": lambda: np.fmin(a, b), "MAX2": lambda: np.fmax(a, b)}.get(op)
            return ("f64", float(r())) if r else (op,) + tuple(kids)
        if t == "i64":
            x, y = li, ri
            ok = y != 0 and not (x == -2 ** 63 and y == -1)
            if op == "ADD":
                return ("i64", _wrap(x + y))
            if op == "SUB":
                return ("i64", _wrap(x - y))
            if op == "MUL":
                return ("i64", _wrap(x * y))
            if op in ("DIV", "MOD"):
                if not ok:
                    return ("i64", 0)
                q = abs(x) // abs(y) * (1 if (x < 0) == (y < 0) else -1)
                return ("i64", _wrap(q if op == "DIV" else x - q * y))
            if op == "MIN2":
                return ("i64", min(x, y))
            if op == "MAX2":
                return ("i64", max(x, y))
            return (op,) + tuple(kids)
        if op in _CMP:
            return ("bool", int(_CMP[op](a, b)))
        if op == "AND":
            return ("bool", int(a != 0 and b != 0))
        if op == "OR":
            return ("bool", int(a != 0 or b != 0))
        return (op,) + tuple(kids)
