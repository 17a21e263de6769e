"""Comparison helpers shared by the tests.

Value parity rule (FP64 path), per CRS entry k of row r:
    |a_k - b_k| <= rtol * max(|b_k|, 1e-3 * max_j |b_rj|) + atol_row * max_j |b_rj|
with rtol = 1e-12 and atol_row = 1e-14.  The row-scale terms exist because many
assembled entries cancel to exactly zero in exact arithmetic (e.g. 483 of the 2337
B^T entries at M=2); there the reference algorithm itself (5-point rule with a
negative weight) is ~1.2e-17 away from the exact rational value, 2.9e-12 of the
1e-3 row floor.

FP32-local rule (fp32_local_check): element matrices computed and stored in FP32, summed
in FP64.  Against the FP64 oracle, per entry k of row r, with S_k = sum_e max |K_e| over
the elements e contributing to the entry (the largest |value| of e's block: Navier-Stokes
A, B^T or B, the whole element matrix otherwise; pyoracle.element_scale):
    |a_k - b_k| <= 1e-5 * max(|b_k|, 1e-3 * max_j |b_rj|) + C * u32 * S_k,   u32 = 2^-24,
and per matrix max_k |a_k - b_k| / max_k |b_k| <= 1e-6.  The first term is the FP64 rule at
1e-5; the second is the forward error bound of element entries computed in FP32: each is
a short sum of terms of the size of the element's entries, rounded at u32 each.  It
matters only where the exact value is far below its row: the entries that cancel (within
an element, or across elements) to ~1e-3 of the row or to zero.  There no FP32 element path
meets 1e-5: rounding the exact FP64 element matrices to FP32 alone gives 3.3e-5 to
3.9e-5 on the Navier-Stokes A block at 4^3..16^3 x 6 (pyoracle.element_round32).
"""
import numpy as np


def row_max(ref, rowptr):
    nr = len(rowptr) - 1
    rows = np.repeat(np.arange(nr), np.diff(rowptr))
    rm = np.zeros(nr)
    np.maximum.at(rm, rows, np.abs(ref))
    return rows, rm


def assert_graph_equal(a, b):
    """Bit-exact CRS graph comparison: a, b = (rowptr, colind)."""
    assert np.array_equal(np.asarray(a[0], np.int64), np.asarray(b[0], np.int64)), "rowptr differs"
    assert np.array_equal(np.asarray(a[1], np.int64), np.asarray(b[1], np.int64)), "colind differs"


def parity_ratio(got, ref, rowptr, rtol=1e-12, atol_row=1e-14, floor=1e-3):
    """max_k |got-ref| / tolerance_k  (<= 1 means within the parity rule)."""
    ref = np.asarray(ref)
    got = np.asarray(got)
    if ref.size == 0:
        return 0.0
    rows, rm = row_max(ref, rowptr)
    tol = rtol * np.maximum(np.abs(ref), floor * rm[rows]) + atol_row * rm[rows]
    tol[tol == 0] = np.finfo(float).tiny
    return float(np.max(np.abs(got - ref) / tol))


def rel_err_rowfloor(got, ref, rowptr, floor=1e-3):
    """Reported figure: max_k |got-ref| / max(|ref_k|, floor*max_j|ref_row j|)."""
    ref = np.asarray(ref)
    got = np.asarray(got)
    if ref.size == 0:
        return 0.0
    rows, rm = row_max(ref, rowptr)
    den = np.maximum(np.abs(ref), floor * rm[rows])
    den[den == 0] = 1.0
    return float(np.max(np.abs(got - ref) / den))


def vec_rel_err(got, ref):
    s = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (s if s > 0 else 1.0))


def incident_elements(elems, nodes):
    """Elements that contain any of `nodes` (sorted ids)."""
    mask = np.isin(elems, nodes).any(axis=1)
    return np.nonzero(mask)[0]


def rows_of(rowptr, colind, vals, rows):
    """(row, col) -> value dict-free extraction: concatenated (cols, vals) of `rows`."""
    idx = np.concatenate([np.arange(rowptr[r], rowptr[r + 1]) for r in rows])
    return colind[idx], vals[idx], idx


def subset_rows_parity(got, ref, rows, rtol=1e-12, atol_row=1e-14):
    """Parity of selected rows.  got/ref = (rowptr, colind, vals) of two CRS matrices whose
    rows `rows` must hold the same columns (ref assembled from every element touching
    those rows, so they are complete); returns the max parity ratio over those rows."""
    gr, gc, gv = got
    rr, rc, rv = ref
    worst = 0.0
    for r in rows:
        a = slice(gr[r], gr[r + 1])
        b = slice(rr[r], rr[r + 1])
        assert np.array_equal(gc[a], rc[b]), "row %d: columns differ" % r
        x, y = gv[a], rv[b]
        rm = np.max(np.abs(y)) if y.size else 0.0
        tol = rtol * np.maximum(np.abs(y), 1e-3 * rm) + atol_row * rm
        tol[tol == 0] = np.finfo(float).tiny
        if y.size:
            worst = max(worst, float(np.max(np.abs(x - y) / tol)))
    return worst


def per_matrix_err(got, ref):
    """max |got - ref| / max |ref| over one matrix (or one set of rows)."""
    ref = np.asarray(ref)
    if ref.size == 0:
        return 0.0
    s = np.max(np.abs(ref))
    return float(np.max(np.abs(np.asarray(got) - ref)) / (s if s > 0 else 1.0))


def round32_floor(rounded, ref):
    """Row-floor error of `rounded` (the oracle's assembly with every element value rounded
    to FP32, pyoracle.element_round32) against ref = (rowptr, colind, vals)."""
    rp, ci, rv = ref
    return rel_err_rowfloor(rounded[2] if isinstance(rounded, tuple) else rounded, rv, rp)


U32 = 2.0 ** -24
FP32_C = 8.0  # measured need: at most 4.6 (3^3x6 Navier-Stokes with coeff (0.5, -2, 3, 1))


def fp32_local_check(got, ref, absref, rows=None, C=FP32_C):
    """FP32-local rule; got = values on ref's pattern, ref = (rowptr, colind, vals), absref
    = sum_e max |K_e| on the same pattern (tuple or values).  Asserts the rule on all rows or on
    `rows`; returns (row-floor error, per-matrix error, effective C = the smallest C the
    entries need)."""
    rp, ci, rv = ref
    s = np.asarray(absref[2] if isinstance(absref, tuple) else absref)
    got = np.asarray(got)
    rv = np.asarray(rv)
    rows_all, rm = row_max(rv, rp)
    idx = np.arange(len(rv)) if rows is None else np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows])
    d = np.abs(got[idx] - rv[idx])
    base = np.maximum(np.abs(rv[idx]), 1e-3 * rm[rows_all[idx]])
    rf = float(np.max(d / np.where(base > 0, base, 1.0))) if idx.size else 0.0
    pm = per_matrix_err(got[idx], rv[idx]) if idx.size else 0.0
    excess = np.maximum(d - 1e-5 * base, 0.0)
    sc = U32 * s[idx]
    ceff = float(np.max(np.where(sc > 0, excess / np.where(sc > 0, sc, 1.0), np.where(excess > 0, np.inf, 0.0)))) \
        if idx.size else 0.0
    assert pm <= 1e-6, ("per-matrix", pm)
    assert ceff <= C, ("entries beyond 1e-5 + C u32 sum|c_e|", ceff, rf)
    return rf, pm, ceff
