"""Result comparison rules for parity tests.

Rows are compared after a canonical sort on the key columns (raw bits).
Keys, row sets, COUNT, integer SUM and MIN/MAX must be bit-exact.  f64
SUM/AVG pass when |got - ref| <= 1e-12 * |ref|, or when both sit within
1e-12 * sum|x| of the exact (compensated) sum, which covers cancellation
in sums of signed values (the reference's own f64 sums vary from run to run
because its workers merge partials in claim order).
"""
import numpy as np

REL = 1e-12


def _u64(a):
    a = np.ascontiguousarray(a)
    return a.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
                  ).astype(np.uint64)


def key_order(keys):
    """Row order sorting by the key columns (first most significant)."""
    return np.lexsort([_u64(k) for k in keys][::-1])


def assert_f64_close(got, ref, exact=None, sum_abs=None, what=""):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, f"{what}: shape {got.shape} vs {ref.shape}"
    both_nan = np.isnan(got) & np.isnan(ref)
    ok = both_nan | (np.abs(got - ref) <= REL * np.abs(ref))
    if exact is not None and sum_abs is not None:
        tol = REL * np.asarray(sum_abs) + 1e-300
        ok |= (np.abs(got - exact) <= tol) & (np.abs(ref - exact) <= tol)
    if not ok.all():
        i = int(np.flatnonzero(~ok)[0])
        raise AssertionError(f"{what}: f64 mismatch at {i}: got {got[i]!r} ref {ref[i]!r}"
                             f" ({int((~ok).sum())} of {len(ok)} rows)")


def compare_groups(got_keys, got_vals, ref_keys, ref_vals, f64_mask, what="",
                   exact=None, sum_abs=None):
    """got/ref: lists of key arrays and value arrays; f64_mask[i] marks
    tolerance-compared value columns ("skip": checked by the caller).  Both
    sides are sorted by keys."""
    assert len(got_keys[0]) == len(ref_keys[0]), (
        f"{what}: {len(got_keys[0])} groups vs {len(ref_keys[0])}")
    go, ro = key_order(got_keys), key_order(ref_keys)
    for i, (a, b) in enumerate(zip(got_keys, ref_keys)):
        np.testing.assert_array_equal(_u64(a[go]), _u64(b[ro]), err_msg=f"{what}: key {i}")
    for i, (a, b) in enumerate(zip(got_vals, ref_vals)):
        if f64_mask[i] == "skip":
            continue
        if f64_mask[i]:
            ex = exact[i][ro] if exact is not None and exact[i] is not None else None
            sa = sum_abs[i][ro] if sum_abs is not None and sum_abs[i] is not None else None
            assert_f64_close(a[go], b[ro], ex, sa, what=f"{what}: value {i}")
        else:
            np.testing.assert_array_equal(np.asarray(a)[go], np.asarray(b)[ro],
                                          err_msg=f"{what}: value {i}")


def assert_var_close(got, ref, scale, sqrt_of=False, what=""):
    """VAR / STDDEV results.  The variance formula (sum of squares / n minus
    mean squared, exec.c:4450-4458) cancels, so both sides carry rounding at
    the scale of sum(x^2)/(n-1): |var_got - var_ref| <= 1e-12 * scale.
    STDDEV columns are compared as their squares."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    if sqrt_of:
        got, ref = got * got, ref * ref
    both_nan = np.isnan(got) & np.isnan(ref)
    ok = both_nan | (np.abs(got - ref) <= REL * np.asarray(scale) + 1e-300)
    if not ok.all():
        i = int(np.flatnonzero(~ok)[0])
        raise AssertionError(f"{what}: variance mismatch at {i}: got {got[i]!r} ref {ref[i]!r}"
                             f" scale {scale[i]!r} ({int((~ok).sum())} of {len(ok)} rows)")


def row_multiset_equal(got_cols, ref_cols, what=""):
    """Two tables hold the same multiset of rows (all columns, raw bits)."""
    assert len(got_cols) == len(ref_cols), f"{what}: column count"
    n = len(got_cols[0]) if got_cols else 0
    assert all(len(c) == n for c in got_cols), f"{what}: ragged"
    assert len(ref_cols[0]) == n, f"{what}: {n} rows vs {len(ref_cols[0])}"
    go, ro = key_order(got_cols), key_order(ref_cols)
    for i, (a, b) in enumerate(zip(got_cols, ref_cols)):
        np.testing.assert_array_equal(_u64(a)[go], _u64(b)[ro], err_msg=f"{what}: column {i}")
