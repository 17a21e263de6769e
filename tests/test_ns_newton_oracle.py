"""Pins of oracle/nsnewton.py (NonLinearSolver::solveNewton for NavierStokesAssFE restated).

- From u = 0 the Navier-Stokes Jacobian is the Stokes matrix (N = W = 0) and the residual
  is g at the Dirichlet dofs, 0 elsewhere, so the first Newton iterate is the Stokes
  solution: it must reproduce the file stokesUnitTest.cpp stores (1e-9, :230).
- Newton converges quadratically, Picard ("FixedPoint") linearly, to the same solution,
  whose residual (recomputed from scratch) vanishes.
"""
import numpy as np
import pytest

import nsnewton as NN
import pyoracle as O
import reference_problems as RP
from feddlib_amd import native as N


def _stokes_case():
    fv, fp, M = RP.STOKES
    p2, _, e2 = O.mesh(M, 2)
    p1, _, e1 = O.mesh(M, 1)
    rows, vals = RP.stokes_dirichlet(p2)
    g = np.zeros(3 * len(p2))
    g[rows] = vals
    return p2, e2, p1, e1, rows, g, fv, fp


def test_first_newton_iterate_is_the_stored_stokes_solution():
    p2, e2, p1, e1, rows, g, fv, fp = _stokes_case()
    u, p, hist, its = NN.newton(p2, e2, e1, len(p1), np.zeros(3 * len(p2)), np.zeros(len(p1)), RP.STOKES_NU,
                                1.0, rows, g, max_its=1)
    assert its == 1
    gv = N.hdf5_read_vector(RP.golden(fv), "velocity")
    gp = N.hdf5_read_vector(RP.golden(fp), "pressure")
    assert np.abs(u - gv).max() <= 1e-9 * np.abs(gv).max()
    assert np.abs(p - gp).max() <= 1e-9 * np.abs(gp).max()


@pytest.mark.parametrize("linearization", [1, 0])
def test_newton_and_picard_converge_to_one_solution(linearization):
    M = 2
    p2, _, e2 = O.mesh(M, 2)
    p1, _, e1 = O.mesh(M, 1)
    rows, vals = RP.stokes_dirichlet(p2)
    g = np.zeros(3 * len(p2))
    g[rows] = vals
    nu = 0.02
    u, p, hist, its = NN.newton(p2, e2, e1, len(p1), np.zeros(3 * len(p2)), np.zeros(len(p1)), nu, 1.0, rows, g,
                                linearization=linearization, tol=1e-11, max_its=60)
    assert hist[-1] <= 1e-11 * hist[0]
    h = np.array(hist)
    if linearization:
        # quadratic: each ratio at most the square of the previous one (up to a constant)
        r = h[1:] / h[:-1]
        assert its <= 8
        assert r[-1] < 1e-2
    else:
        assert its > 8  # linear convergence needs more updates
    # the final residual recomputed from scratch
    res = O.ns_assembly(p2, e2, e1, len(p1), u, p, nu, 1.0, True, 2, True)
    free = np.ones(3 * len(p2), bool)
    free[rows] = False
    assert np.abs(res["res_u"][free]).max() <= 1e-10 * np.abs(res["res_u"]).max() + 1e-12
    assert np.allclose(u[rows], g[rows], rtol=0, atol=1e-12)
    # both linearisations reach the same solution
    u2, p2_, _, _ = NN.newton(p2, e2, e1, len(p1), np.zeros(3 * len(p2)), np.zeros(len(p1)), nu, 1.0, rows, g,
                              linearization=1 - linearization, tol=1e-11, max_its=60)
    assert np.abs(u - u2).max() <= 1e-9 * np.abs(u).max()
    assert np.abs(p - p2_).max() <= 1e-8 * np.abs(p).max()
