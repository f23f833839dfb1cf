"""ctypes mirror of the td.h C API.

`Engine(path)` binds any library with the td.h ABI: this package's
libteide_amd.so (the GPU engine) or, in tests, the reference engine built
from source.  It offers thin wrappers over the C calls plus numpy helpers to
build tables and read results, so a query DAG reads like the reference's
own C tests:

    e = engine()
    t = e.table([("flag", flag), ("region", region), ("amount", amount)])
    g = e.graph(t)
    pred = g.eq(g.scan("flag"), g.const_i64(0))
    ...
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libteide_amd.so")

# td.h type tags and opcodes (ABI values)
TD_BOOL, TD_U8, TD_CHAR, TD_I16, TD_I32, TD_I64, TD_F64 = 1, 2, 3, 4, 5, 6, 7
TD_DATE, TD_TIME, TD_TIMESTAMP, TD_TABLE, TD_SEL, TD_SYM = 9, 10, 11, 13, 16, 20
TD_ATOM_I64, TD_ATOM_F64 = -6, -7
OP = dict(SUM=50, PROD=51, MIN=52, MAX=53, COUNT=54, AVG=55, FIRST=56, LAST=57,
          COUNT_DISTINCT=58, STDDEV=59, STDDEV_POP=73, VAR=74, VAR_POP=75)
ERR_NAMES = ["OK", "OOM", "TYPE", "RANGE", "LENGTH", "RANK", "DOMAIN", "NYI", "IO",
             "SCHEMA", "CORRUPT", "CANCEL"]

_DT_TO_TYPE = {np.dtype(np.bool_): (TD_BOOL, 0), np.dtype(np.uint8): (TD_U8, 0),
               np.dtype(np.int16): (TD_I16, 0), np.dtype(np.int32): (TD_I32, 0),
               np.dtype(np.int64): (TD_I64, 0), np.dtype(np.float64): (TD_F64, 0),
               np.dtype(np.uint16): (TD_SYM, 1), np.dtype(np.uint32): (TD_SYM, 2)}
_TYPE_TO_DT = {TD_BOOL: np.uint8, TD_U8: np.uint8, TD_CHAR: np.uint8, TD_I16: np.int16,
               TD_I32: np.int32, TD_I64: np.int64, TD_F64: np.float64, TD_DATE: np.int32,
               TD_TIME: np.int32, TD_TIMESTAMP: np.int64}
_SYM_DT = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.int64}


class TdError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        name = ERR_NAMES[code] if 0 <= code < len(ERR_NAMES) else str(code)
        super().__init__(f"td error {name} ({code}) {what}".strip())


class Header(C.Structure):
    """The 32-byte td_t header (td.h:224-259)."""
    _fields_ = [("nullmap", C.c_uint8 * 16), ("mmod", C.c_uint8), ("order", C.c_uint8),
                ("type", C.c_int8), ("attrs", C.c_uint8), ("rc", C.c_uint32),
                ("val", C.c_int64)]


def dtype_of(type_tag, attrs):
    if type_tag == TD_SYM:
        return np.dtype(_SYM_DT[attrs & 3])
    return np.dtype(_TYPE_TO_DT[type_tag])


def is_err(p):
    return p is None or p < 32


class ProfRow(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("total_ms", C.c_double)]


class DevMem(C.Structure):
    _fields_ = [("in_use", C.c_uint64), ("peak", C.c_uint64), ("reserved", C.c_uint64),
                ("budget", C.c_uint64), ("device_total", C.c_uint64)]


class ServeCfg(C.Structure):
    _fields_ = [("fact", C.c_void_p), ("dim", C.c_void_p), ("fk", C.c_int64), ("pk", C.c_int64),
                ("pcol", C.c_int64), ("vals", C.c_int64 * 4), ("keys", C.c_int64 * 4),
                ("n_vals", C.c_int32), ("n_keys", C.c_int32), ("budget", C.c_uint64),
                ("chunk_rows", C.c_int64), ("max_active", C.c_int32), ("pad", C.c_int32)]


class ServeInfo(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("queued", "active", "ready", "steps", "chunks",
                                         "shared", "alone")] + \
               [(n, C.c_uint64) for n in ("held", "base", "budget")]


class Engine:
    """One td.h library.  Pointers are Python ints (0 / < 32: error)."""

    def __init__(self, path=LIB_PATH, device=True):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
        self.path = path
        self.L = C.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        self.device = device
        P, I64, U8, U16, I8 = C.c_void_p, C.c_int64, C.c_uint8, C.c_uint16, C.c_int8
        sig = {
            "td_heap_init": (None, []), "td_sym_init": (None, []),
            "td_vec_new": (P, [I8, I64]), "td_sym_vec_new": (P, [U8, I64]),
            "td_table_new": (P, [I64]), "td_table_add_col": (P, [P, I64, P]),
            "td_table_get_col": (P, [P, I64]), "td_table_get_col_idx": (P, [P, I64]),
            "td_table_col_name": (I64, [P, I64]), "td_table_ncols": (I64, [P]),
            "td_table_nrows": (I64, [P]), "td_sym_intern": (I64, [C.c_char_p, C.c_size_t]),
            "td_sym_find": (I64, [C.c_char_p, C.c_size_t]),
            "td_sym_str": (P, [I64]), "td_str_ptr": (C.c_char_p, [P]),
            "td_str_len": (C.c_size_t, [P]), "td_release": (None, [P]), "td_retain": (None, [P]),
            "td_graph_new": (P, [P]), "td_graph_free": (None, [P]),
            "td_scan": (P, [P, C.c_char_p]), "td_const_i64": (P, [P, I64]),
            "td_const_f64": (P, [P, C.c_double]), "td_const_bool": (P, [P, C.c_bool]),
            "td_const_str": (P, [P, C.c_char_p]), "td_const_table": (P, [P, P]),
            "td_const_vec": (P, [P, P]),
            "td_filter": (P, [P, P, P]), "td_group": (P, [P, P, U8, P, P, U8]),
            "td_sort_op": (P, [P, P, P, P, P, U8]),
            "td_col_save": (C.c_int, [P, C.c_char_p]), "td_col_load": (P, [C.c_char_p]),
            "td_col_mmap": (P, [C.c_char_p]),
            "td_splay_save": (C.c_int, [P, C.c_char_p, C.c_char_p]),
            "td_splay_load": (P, [C.c_char_p]), "td_read_splayed": (P, [C.c_char_p, C.c_char_p]),
            "td_sym_save": (C.c_int, [C.c_char_p]), "td_sym_load": (C.c_int, [C.c_char_p]),
            "td_part_load": (P, [C.c_char_p, C.c_char_p]),
            "td_read_parted": (P, [C.c_char_p, C.c_char_p]),
            "td_join": (P, [P, P, P, P, P, U8, U8]), "td_optimize": (P, [P, P]),
            "td_execute": (P, [P, P]), "td_head": (P, [P, P, I64]), "td_tail": (P, [P, P, I64]),
            "td_sel_from_pred": (P, [P]), "td_sel_and": (P, [P, P]), "td_sel_new": (P, [I64]),
            "td_sel_recompute": (None, [P]), "td_vec_from_raw": (P, [I8, P, I64]),
            "td_err_str": (C.c_char_p, [C.c_int]), "td_pool_init": (C.c_int, [C.c_uint32]),
            "td_thread_count": (C.c_uint32, []), "td_cancel": (None, []),
        }
        for op in ["eq", "ne", "lt", "le", "gt", "ge", "and", "or", "add", "sub", "mul", "div",
                   "mod", "min2", "max2"]:
            sig["td_" + op] = (P, [P, P, P])
        for op in ["not", "sum", "count", "avg", "min_op", "max_op", "first", "last", "neg",
                   "abs", "sqrt_op", "log_op", "exp_op", "ceil_op", "floor_op", "prod",
                   "count_distinct", "stddev", "stddev_pop", "var", "var_pop"]:
            sig["td_" + op] = (P, [P, P])
        sig["td_cast"] = (P, [P, P, I8])
        if device:
            sig.update({
                "td_dev_available": (C.c_int, []), "td_dev_put": (P, [P]),
                "td_dev_get": (P, [P]), "td_dev_is_resident": (C.c_int, [P]),
                "td_dev_sync": (C.c_int, []), "td_dev_prof_enable": (None, [C.c_int]),
                "td_dev_prof_reset": (None, []), "td_dev_prof_read": (C.c_int, [P, C.c_int]),
                "td_dev_force_paths": (None, [C.c_int, C.c_int, C.c_int]),
                "td_dev_last_plan": (C.c_char_p, []), "td_dev_mem_stats": (None, [P]),
                "td_dev_set_budget": (None, [C.c_uint64]),
                "td_dev_execute_batch": (C.c_int, [P, P, I64, P]),
                "td_dev_batch_prepare": (P, [P, P, I64]),
                "td_dev_batch_run": (C.c_int, [P, P]),
                "td_dev_batch_info": (I64, [P, P, P]),
                "td_dev_batch_radix": (I64, [P, P, P]),
                "td_dev_batch_free": (None, [P]),
                "td_dev_col_load": (P, [C.c_char_p]),
                "td_dev_splay_load": (P, [C.c_char_p, C.c_char_p]),
                "td_dev_parted_load": (P, [C.c_char_p, C.c_char_p]),
                "td_dev_serve_open": (P, [P]), "td_dev_serve_submit": (I64, [P, P, P]),
                "td_dev_serve_step": (C.c_int, [P]), "td_dev_serve_poll": (C.c_int, [P, P, P]),
                "td_dev_serve_info": (None, [P, P]), "td_dev_serve_close": (None, [P]),
                "td_dev_cancel_after": (None, [I64]), "td_dev_cancel_polls": (I64, []),
                "td_dev_fail_allocs": (None, [C.c_uint64, I64]),
                "td_heap_gc": (None, []),
            })
        for name, (res, args) in sig.items():
            f = getattr(self.L, name)
            f.restype = res
            f.argtypes = args
        self.L.td_heap_init()
        self.L.td_sym_init()

    # --------------------------------------------------------------- basics
    def hdr(self, p):
        return Header.from_address(p)

    def check(self, p, what=""):
        if is_err(p):
            raise TdError(p or 0, what)
        return p

    def sym(self, name):
        b = name.encode()
        return self.L.td_sym_intern(b, len(b))

    def sym_name(self, sid):
        s = self.L.td_sym_str(sid)
        if not s:
            return None
        return C.string_at(self.L.td_str_ptr(s), self.L.td_str_len(s)).decode()

    def release(self, p):
        if p and not is_err(p):
            self.L.td_release(p)

    # ---------------------------------------------------------- construction
    def vec(self, arr, type_tag=None, attrs=None):
        arr = np.ascontiguousarray(arr)
        t, a = _DT_TO_TYPE[arr.dtype]
        if type_tag is not None:
            t = type_tag
        if attrs is not None:
            a = attrs
        n = len(arr)
        if t == TD_SYM:
            v = self.check(self.L.td_sym_vec_new(a, n), "sym_vec_new")
        else:
            v = self.check(self.L.td_vec_new(t, n), "vec_new")
        self.hdr(v).val = n
        if n:
            C.memmove(v + 32, arr.ctypes.data, arr.nbytes)
        return v

    def table(self, cols):
        """cols: list of (name, ndarray) or (name, ndarray, type_tag[, attrs])."""
        t = self.check(self.L.td_table_new(len(cols)), "table_new")
        for c in cols:
            name, arr = c[0], c[1]
            v = self.vec(arr, *c[2:])
            t = self.check(self.L.td_table_add_col(t, self.sym(name), v), "add_col")
            self.L.td_release(v)
        return t

    def put(self, p):
        """Device-resident copy (td_dev_put)."""
        return self.check(self.L.td_dev_put(p), "dev_put")

    # --------------------------------------------------------------- reading
    def read_vec(self, v):
        h = self.hdr(v)
        owned = None
        if h.mmod == 3:
            owned = v = self.check(self.L.td_dev_get(v), "dev_get")
            h = self.hdr(v)
        dt = dtype_of(h.type, h.attrs)
        n = h.val
        out = np.empty(n, dt)
        if n:
            C.memmove(out.ctypes.data, v + 32, n * dt.itemsize)
        if owned:
            self.L.td_release(owned)
        return out

    def read_table(self, t):
        """[(name, type_tag, attrs, ndarray)] of a result table."""
        out = []
        for i in range(self.L.td_table_ncols(t)):
            col = self.L.td_table_get_col_idx(t, i)
            h = self.hdr(col)
            out.append((self.sym_name(self.L.td_table_col_name(t, i)), h.type, h.attrs,
                        self.read_vec(col)))
        return out

    def atom_value(self, p):
        h = self.hdr(p)
        if h.type == TD_ATOM_F64:
            return C.c_double.from_address(p + 24).value
        return h.val

    def graph(self, tbl):
        return Graph(self, tbl)

    # ------------------------------------------------------------ device ext
    def prof(self, on=True):
        self.L.td_dev_prof_enable(int(on))

    def prof_reset(self):
        self.L.td_dev_prof_reset()

    def prof_read(self):
        rows = (ProfRow * 128)()
        n = self.L.td_dev_prof_read(C.addressof(rows), 128)
        return {rows[i].name.decode(): (rows[i].launches, rows[i].total_ms) for i in range(n)}

    def force_paths(self, join=0, group=0, radix_bits=0):
        self.L.td_dev_force_paths(join, group, radix_bits)

    def last_plan(self):
        return self.L.td_dev_last_plan().decode()

    def execute_batch(self, graphs, roots):
        """td_dev_execute_batch over Graph objects and their root ops."""
        n = len(graphs)
        ga = (C.c_void_p * n)(*[g.g for g in graphs])
        ra = (C.c_void_p * n)(*roots)
        res = (C.c_void_p * n)()
        rc = self.L.td_dev_execute_batch(C.addressof(ga), C.addressof(ra), n, C.addressof(res))
        return rc, [res[i] for i in range(n)]

    def batch_prepare(self, graphs, roots):
        n = len(graphs)
        ga = (C.c_void_p * n)(*[g.g for g in graphs])
        ra = (C.c_void_p * n)(*roots)
        h = self.L.td_dev_batch_prepare(C.addressof(ga), C.addressof(ra), n)
        if not h:
            raise TdError(1, "td_dev_batch_prepare")
        return h

    def batch_run(self, h, n):
        res = (C.c_void_p * n)()
        rc = self.L.td_dev_batch_run(h, C.addressof(res))
        return rc, [res[i] for i in range(n)]

    def batch_info(self, h):
        tr, bl = C.c_int64(), C.c_int()
        nb = self.L.td_dev_batch_info(h, C.byref(tr), C.byref(bl))
        return nb, tr.value, bl.value

    def batch_radix(self, h):
        """(queries on the radix hash path, build rows per partition, max bits)"""
        pr, mb = C.c_int64(), C.c_int()
        nh = self.L.td_dev_batch_radix(h, C.byref(pr), C.byref(mb))
        return nh, pr.value, mb.value

    # -------------------------------------------------------------- serving
    def serve_open(self, fact, dim, fk, pk, pcol, vals, keys, budget=0, chunk_rows=0,
                   max_active=0):
        """td_dev_serve_open over resident tables; column arguments are names."""
        cfg = ServeCfg()
        cfg.fact, cfg.dim = fact, dim
        cfg.fk, cfg.pk, cfg.pcol = self.sym(fk), self.sym(pk), self.sym(pcol)
        for i, v in enumerate(vals):
            cfg.vals[i] = self.sym(v)
        for i, k in enumerate(keys):
            cfg.keys[i] = self.sym(k)
        cfg.n_vals, cfg.n_keys = len(vals), len(keys)
        cfg.budget, cfg.chunk_rows, cfg.max_active = int(budget), int(chunk_rows), int(max_active)
        h = self.L.td_dev_serve_open(C.byref(cfg))
        if not h:
            raise TdError(1, "td_dev_serve_open")
        return h

    def serve_submit(self, h, graph, root):
        rid = self.L.td_dev_serve_submit(h, graph.g, root)
        if rid < 0:
            raise TdError(-rid, "td_dev_serve_submit")
        return rid

    def serve_step(self, h):
        n = self.L.td_dev_serve_step(h)
        if n < 0:
            raise TdError(-n, "td_dev_serve_step")
        return n

    def serve_poll(self, h):
        """All finished requests as [(id, result pointer)]."""
        out = []
        rid, res = C.c_int64(), C.c_void_p()
        while self.L.td_dev_serve_poll(h, C.byref(rid), C.byref(res)):
            out.append((rid.value, res.value or 0))
        return out

    def serve_info(self, h):
        i = ServeInfo()
        self.L.td_dev_serve_info(h, C.byref(i))
        return {k: getattr(i, k) for k, _ in ServeInfo._fields_}

    def serve_close(self, h):
        self.L.td_dev_serve_close(h)

    def mem_stats(self):
        m = DevMem()
        self.L.td_dev_mem_stats(C.byref(m))
        return {k: getattr(m, k) for k, _ in DevMem._fields_}


class Graph:
    """A td_graph_t with node builders (pointers to td_op_t as ints)."""

    def __init__(self, eng, tbl):
        self.e = eng
        self.L = eng.L
        self.g = eng.check(self.L.td_graph_new(tbl), "graph_new")

    def free(self):
        if self.g:
            self.L.td_graph_free(self.g)
            self.g = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def scan(self, name):
        return self.L.td_scan(self.g, name.encode())

    def const_i64(self, v):
        return self.L.td_const_i64(self.g, int(v))

    def const_f64(self, v):
        return self.L.td_const_f64(self.g, float(v))

    def const_str(self, s):
        return self.L.td_const_str(self.g, s.encode())

    def const_table(self, t):
        return self.L.td_const_table(self.g, t)

    def __getattr__(self, op):
        # binary / unary builders: eq ne lt le gt ge and_ or_ not_ sum count ...
        name = {"and_": "and", "or_": "or", "not_": "not", "min_": "min_op",
                "max_": "max_op"}.get(op, op)
        f = getattr(self.L, "td_" + name, None)
        if f is None:
            raise AttributeError(op)
        return lambda *a: f(self.g, *a)

    def filter(self, inp, pred):
        return self.L.td_filter(self.g, inp, pred)

    def group(self, keys, aggs):
        """keys: [op]; aggs: [(OP name, op)]."""
        nk, na = len(keys), len(aggs)
        ka = (C.c_void_p * max(nk, 1))(*keys)
        ops = (C.c_uint16 * max(na, 1))(*[OP[o] for o, _ in aggs])
        ins = (C.c_void_p * max(na, 1))(*[i for _, i in aggs])
        return self.L.td_group(self.g, C.addressof(ka), nk, C.addressof(ops),
                               C.addressof(ins), na)

    def join(self, left, lkeys, right, rkeys, join_type=0):
        n = len(lkeys)
        la = (C.c_void_p * n)(*lkeys)
        ra = (C.c_void_p * n)(*rkeys)
        return self.L.td_join(self.g, left, C.addressof(la), right, C.addressof(ra), n,
                              join_type)

    def head(self, inp, n):
        return self.L.td_head(self.g, inp, n)

    def sort(self, inp, keys, descs, nulls_first=None):
        """td_sort_op: sort `inp` (a table node) by key nodes."""
        n = len(keys)
        ka = (C.c_void_p * n)(*keys)
        da = (C.c_uint8 * n)(*[int(d) for d in descs])
        nf = (C.c_uint8 * n)(*[int(f) for f in (nulls_first or descs)])
        return self.L.td_sort_op(self.g, inp, C.addressof(ka), C.addressof(da),
                                 C.addressof(nf), n)

    def optimize(self, root):
        return self.L.td_optimize(self.g, root)

    def execute(self, root):
        r = self.L.td_execute(self.g, root)
        if is_err(r):
            raise TdError(r or 0, "td_execute")
        return r

    SELECTION_OFFSET = 56          # offsetof(td_graph_t, selection)

    def selection(self):
        """Current g->selection pointer (td_graph_t.selection)."""
        return C.c_void_p.from_address(self.g + self.SELECTION_OFFSET).value

    def set_selection(self, sel):
        cur = self.selection()
        if sel:
            self.L.td_retain(sel)
        C.c_void_p.from_address(self.g + self.SELECTION_OFFSET).value = sel
        if cur:
            self.L.td_release(cur)


_engine = None


def engine():
    """The process-wide GPU engine (libteide_amd.so)."""
    global _engine
    if _engine is None:
        _engine = Engine(LIB_PATH, device=True)
    return _engine
