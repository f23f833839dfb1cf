"""Reader for the TDT1 column-table files written by oracle/ref_harness.c.

A TDT1 file holds one result table: magic "TDT1", u32 column count, then per
column its name, td.h type tag, attrs, row count and raw little-endian data.
"""
import struct

import numpy as np

# td.h type tag -> numpy dtype (TD_SYM width comes from attrs bits 0-1).
_DTYPES = {1: np.uint8, 2: np.uint8, 3: np.uint8, 4: np.int16, 5: np.int32,
           6: np.int64, 7: np.float64, 9: np.int32, 10: np.int32, 11: np.int64}
_SYM_DTYPES = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.int64}
TD_SYM = 20


def dtype_of(type_tag, attrs):
    if type_tag == TD_SYM:
        return _SYM_DTYPES[attrs & 3]
    return _DTYPES[type_tag]


def read_tdt(path):
    """Return an ordered list of (name, type_tag, attrs, ndarray)."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:4] != b"TDT1":
        raise ValueError(f"{path}: not a TDT1 file")
    (ncols,) = struct.unpack_from("<I", buf, 4)
    off = 8
    cols = []
    for _ in range(ncols):
        (nl,) = struct.unpack_from("<I", buf, off)
        off += 4
        name = buf[off:off + nl].decode()
        off += nl
        t, a, _pad, n = struct.unpack_from("<bBHq", buf, off)
        off += 12
        dt = np.dtype(dtype_of(t, a))
        arr = np.frombuffer(buf, dtype=dt, count=n, offset=off).copy()
        off += n * dt.itemsize
        cols.append((name, t, a, arr))
    return cols


def canonical_order(arrays):
    """Row order sorting rows lexicographically by all columns (first column
    most significant), comparing every column by its raw bits as uint64."""
    keys = [np.ascontiguousarray(a).astype(a.dtype).view(
        {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])
        .astype(np.uint64) for a in arrays]
    return np.lexsort(keys[::-1])


def canonical_digest(arrays):
    """SHA-256 of the columns after a canonical row sort (order-free identity
    of a result table)."""
    import hashlib
    order = canonical_order(arrays)
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a[order]).tobytes())
    return h.hexdigest()
