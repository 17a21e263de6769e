"""Comparison helpers shared by the tests."""
import numpy as np


def csr_to_dict(rp, ci, v):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return rows, ci, v


def assert_graph_equal(a, b):
    """Bit-exact CRS graph comparison: a, b = (rowptr, colind)."""
    assert np.array_equal(np.asarray(a[0], np.int64), np.asarray(b[0], np.int64)), "rowptr differs"
    assert np.array_equal(np.asarray(a[1], np.int64), np.asarray(b[1], np.int64)), "colind differs"


def rel_err_rowfloor(got, ref, rowptr, floor=1e-3):
    """max_k |got-ref| / max(|ref_k|, floor*max_j|ref_row j|)  (row-scale floor for entries
    that cancel to ~0)."""
    ref = np.asarray(ref)
    got = np.asarray(got)
    nr = len(rowptr) - 1
    rows = np.repeat(np.arange(nr), np.diff(rowptr))
    rowmax = np.zeros(nr)
    np.maximum.at(rowmax, rows, np.abs(ref))
    den = np.maximum(np.abs(ref), floor * rowmax[rows])
    den[den == 0] = 1.0
    return float(np.max(np.abs(got - ref) / den)) if ref.size else 0.0


def vec_rel_err(got, ref):
    s = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (s if s > 0 else 1.0))
