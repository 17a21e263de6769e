"""The FP32 element-storage floor (pyoracle.element_round32) that the FP32-local rule of
helpers.fp32_local_check is measured against.

- The switch is pinned by an independent path: the oracle's single-element function
  (orc_ns_element), each element matrix rounded to FP32 in numpy and summed by scipy,
  gives the switched assembly to 1e-15 (only the order of the FP64 sums differs).
- The switch is off again after the context, and the FP64 assembly is bit-identical.
- The floor itself: storing the exact element matrices in FP32 puts the Navier-Stokes A
  block at 3e-5..4e-5 under the row-floor rule (above 1e-5), while the per-matrix error
  max|d| / max|A| stays below 1e-7, and B^T, B stay below 1e-7 under both.
"""
import numpy as np
import scipy.sparse as sp

import pyoracle as O
from feddlib_amd import workloads as W
from helpers import per_matrix_err, rel_err_rowfloor, round32_floor

NU, RHO = 0.01, 1.0


def _ns(M):
    p2, _, e2 = O.mesh(M, 2)
    p1, _, e1 = O.mesh(M, 1)
    u = W.ns_velocity(p2, 0)
    return p2, e2, p1, e1, u, np.zeros(len(p1))


def test_switch_equals_elementwise_rounding():
    p2, e2, p1, e1, u, p = _ns(2)
    with O.element_round32():
        rnd = O.ns_assembly(p2, e2, e1, len(p1), u, p, NU, RHO, True, 0, False)
    n = 3 * len(p2)
    rows, cols, vals = [], [], []
    for T in range(len(e2)):
        sol = np.concatenate([u[e2[T]].ravel(), p[e1[T]]])
        K = O.ns_element(p2[e2[T]], sol, NU, RHO, True)[0]
        d = (3 * e2[T][:, None] + np.arange(3)).ravel()
        rows.append(np.repeat(d, 30))
        cols.append(np.tile(d, 30))
        vals.append(K[:30, :30].astype(np.float32).astype(np.float64).ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    rp, ci, v = rnd["A"]
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, ci)
    assert np.max(np.abs(A.data - v)) <= 1e-15 * np.max(np.abs(v))
    # every element value was rounded: the plain assembly differs
    ref = O.ns_assembly(p2, e2, e1, len(p1), u, p, NU, RHO, True, 0, False)
    assert not np.array_equal(ref["A"][2], v)


def test_switch_resets():
    p2, e2, p1, e1, u, p = _ns(2)
    a = O.ns_assembly(p2, e2, e1, len(p1), u, p, NU, RHO, True, 0, False)
    with O.element_round32():
        O.ns_assembly(p2, e2, e1, len(p1), u, p, NU, RHO, True, 0, False)
    b = O.ns_assembly(p2, e2, e1, len(p1), u, p, NU, RHO, True, 0, False)
    for k in ("A", "BT", "B"):
        assert np.array_equal(a[k][2], b[k][2])


def test_ns_fp32_storage_floor():
    p2, e2, p1, e1, u, p = _ns(6)
    ref = O.ns_assembly(p2, e2, e1, len(p1), u, p, NU, RHO, True, 0, False)
    with O.element_round32():
        rnd = O.ns_assembly(p2, e2, e1, len(p1), u, p, NU, RHO, True, 0, False)
    fa = round32_floor(rnd["A"], ref["A"])
    assert 1e-5 < fa < 1e-4, fa
    assert per_matrix_err(rnd["A"][2], ref["A"][2]) < 1e-7
    for k in ("BT", "B"):
        assert rel_err_rowfloor(rnd[k][2], ref[k][2], ref[k][0]) < 1e-7
        assert per_matrix_err(rnd[k][2], ref[k][2]) < 1e-7
