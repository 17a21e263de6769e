"""Comparison helpers shared by the tests.

Value parity rule (FP64 path), per CRS entry k of row r:
    |a_k - b_k| <= rtol * max(|b_k|, 1e-3 * max_j |b_rj|) + atol_row * max_j |b_rj|
with rtol = 1e-12 and atol_row = 1e-14.  The row-scale terms exist because many
assembled entries cancel to exactly zero in exact arithmetic (e.g. 483 of the 2337
B^T entries at M=2); there the reference algorithm itself (5-point rule with a
negative weight) is ~1.2e-17 away from the exact rational value, 2.9e-12 of the
1e-3 row floor.  The FP32-local path uses rtol = 1e-5, atol_row = 1e-6.
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
