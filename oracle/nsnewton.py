"""CPU restatement of FEDDLib's Newton loop for the P2-P1 Navier-Stokes problem.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of fedd_newton_navier_stokes
and fedd_ns_linear_solve.  The product never imports it.

Follows NonLinearSolver::solveNewton (feddlib/problems/Solver/NonLinearSolver_def.hpp:
356-414) with NavierStokesAssFE (feddlib/problems/specific/NavierStokesAssFE_def.hpp:
225-313):
  - calculateNonLinResidualVec("reverse"): residual = rhs - F(u) with F(u) the "Rhs"
    assembly (constant matrix + rho N) [u; p] (:272-313), then setBCMinusVector: g - u
    at the Dirichlet dofs (BCBuilder_def.hpp:234-300);
  - Criterion "Residual": ||r|| / ||r_0|| < relNonLinTol before the update (:367-393),
    "Update": ||delta|| < relNonLinTol after it (:395-402, NonLinearProblem_def.hpp:148-162);
  - assemble("Newton") (Jacobian; "FixedPoint" for Picard), setBoundariesSystem
    (BCBuilder::setSystem, oracle/bcbuilder.py), then a linear solve (here a direct one).
The element matrices and residuals come from the oracle's restated assembly (pyoracle).
As in fedd_newton_navier_stokes, the residual of the final iterate is evaluated after
the last update.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import bcbuilder as BC
import pyoracle as O


def system(blocks, nv, npr):
    A = sp.csr_matrix((blocks[(0, 0)][2], blocks[(0, 0)][1], blocks[(0, 0)][0]), shape=(nv, nv))
    BT = sp.csr_matrix((blocks[(0, 1)][2], blocks[(0, 1)][1], blocks[(0, 1)][0]), shape=(nv, npr))
    B = sp.csr_matrix((blocks[(1, 0)][2], blocks[(1, 0)][1], blocks[(1, 0)][0]), shape=(npr, nv))
    return sp.bmat([[A, BT], [B, None]], format="csc")


def newton(p2, e2, e1, npr, u0, p0, nu, rho, rows, g, linearization=1, criterion=0, tol=1e-10, max_its=20,
           fext=None):
    """Returns (u, p, history of ||r||, number of updates)."""
    nv = 3 * len(p2)
    u = np.array(u0, dtype=np.float64).ravel().copy()
    p = np.array(p0, dtype=np.float64).ravel().copy()
    rhs_u = np.zeros(nv) if fext is None else np.asarray(fext, np.float64).ravel()
    hist = []
    r0 = None
    its = 0
    mode = 0 if linearization else 1
    while True:
        res = O.ns_assembly(p2, e2, e1, npr, u, p, nu, rho, bool(linearization), 2, True)
        ru = BC.bc_vector(1, rhs_u - res["res_u"], rows, g, u)
        rp = -res["res_p"]
        r = np.concatenate([ru, rp])
        rn = float(np.linalg.norm(r))
        hist.append(rn)
        if r0 is None:
            r0 = rn
        if criterion == 0 and (rn == 0.0 or rn / r0 < tol):
            break
        if its >= max_its:
            break
        J = O.ns_assembly(p2, e2, e1, npr, u, p, nu, rho, bool(linearization), mode, False)
        blocks = {(0, 0): J["A"], (0, 1): J["BT"], (1, 0): J["B"]}
        vals = BC.set_system(blocks, {0: rows})
        blocks = {k: (b[0], b[1], vals[k]) for k, b in blocks.items()}
        d = spla.spsolve(system(blocks, nv, npr), r)
        u += d[:nv]
        p += d[nv:]
        its += 1
        if criterion == 1 and np.linalg.norm(d) < tol:
            break
    return u, p, hist, its
