"""fedd_newton_navier_stokes and fedd_ns_linear_solve on the GPU.

- The first Newton iterate from u = 0 is the Stokes solution: it reproduces the file
  stokesUnitTest.cpp stores (1e-9 relative, :230), all on the device (fused assembly,
  setBCMinusVector, setSystem, GMRES).
- Newton and Picard iterations match the oracle's restated loop (oracle/nsnewton.py, with
  direct solves) iterate by iterate: same number of updates, residual histories within
  the GMRES tolerance, the same converged solution to 1e-8.
- The saddle-point GMRES solve equals a direct solve of the same system.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import nsnewton as NN
import reference_problems as RP
from feddlib_amd import native as N

pytestmark = pytest.mark.gpu


def _setup(ctx, M):
    p2, _, e2 = N.mesh_structured(M, 2)
    p1, _, e1 = N.mesh_structured(M, 1)
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    ctx.graph_build(N.GRAPH_NS)
    rows, vals = RP.stokes_dirichlet(p2)
    g = np.zeros(3 * len(p2))
    g[rows] = vals
    ctx.dirichlet_set(rows.astype(np.int32))
    ctx.dirichlet_values(g)
    ctx.vector_upload(N.VEC_U, np.zeros(3 * len(p2)))
    ctx.vector_upload(N.VEC_P, np.zeros(len(p1)))
    return p2, e2, p1, e1, rows, g


def test_first_newton_iterate_reproduces_stored_stokes_solution(gpu_ctx):
    fv, fp, M = RP.STOKES
    p2, e2, p1, e1, rows, g = _setup(gpu_ctx, M)
    st, hist = gpu_ctx.newton_navier_stokes(RP.STOKES_NU, RP.STOKES_RHO, max_its=1, lin_rtol=1e-13,
                                            lin_maxit=20000)
    assert st["newton_iters"] == 1 and st["last_lin_relres"] <= 1e-13
    u = gpu_ctx.vector_download(N.VEC_U, 3 * len(p2))
    p = gpu_ctx.vector_download(N.VEC_P, len(p1))
    gv = N.hdf5_read_vector(RP.golden(fv), "velocity")
    gp = N.hdf5_read_vector(RP.golden(fp), "pressure")
    assert np.abs(u - gv).max() <= 1e-9 * np.abs(gv).max()
    assert np.abs(p - gp).max() <= 1e-9 * np.abs(gp).max()
    assert hist[0] == pytest.approx(np.linalg.norm(g), rel=1e-12)  # r_0 = g - 0 at the Dirichlet dofs


@pytest.mark.parametrize("linearization", [1, 0])
def test_newton_matches_oracle_loop(gpu_ctx, linearization):
    M, nu, rho = 2, 0.02, 1.0
    p2, e2, p1, e1, rows, g = _setup(gpu_ctx, M)
    st, hist = gpu_ctx.newton_navier_stokes(nu, rho, linearization=linearization, rel_tol=1e-11, max_its=60,
                                            lin_rtol=1e-13, lin_maxit=20000)
    u_ref, p_ref, h_ref, its_ref = NN.newton(p2, e2, e1, len(p1), np.zeros(3 * len(p2)), np.zeros(len(p1)), nu,
                                             rho, rows, g, linearization=linearization, tol=1e-11, max_its=60)
    assert st["converged"] == 1, (st, hist)
    assert st["newton_iters"] == its_ref, (st, hist, h_ref)
    n = min(len(hist), len(h_ref))
    # histories agree while the residual is well above the GMRES tolerance
    for k in range(n):
        if h_ref[k] > 1e-8 * h_ref[0]:
            assert abs(hist[k] - h_ref[k]) <= 1e-6 * h_ref[k] + 1e-12 * h_ref[0], k
    u = gpu_ctx.vector_download(N.VEC_U, 3 * len(p2))
    p = gpu_ctx.vector_download(N.VEC_P, len(p1))
    assert np.abs(u - u_ref).max() <= 1e-8 * np.abs(u_ref).max()
    assert np.abs(p - p_ref).max() <= 1e-8 * np.abs(p_ref).max()


def test_update_criterion_and_iteration_cap(gpu_ctx):
    M, nu = 2, 0.02
    p2, e2, p1, e1, rows, g = _setup(gpu_ctx, M)
    st, hist = gpu_ctx.newton_navier_stokes(nu, 1.0, criterion=1, rel_tol=1e-9, max_its=40, lin_rtol=1e-13)
    assert st["converged"] == 1 and st["last_update"] < 1e-9
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * len(p2)))
    gpu_ctx.vector_upload(N.VEC_P, np.zeros(len(p1)))
    st, hist = gpu_ctx.newton_navier_stokes(nu, 1.0, linearization=0, rel_tol=1e-14, max_its=3)
    assert st["converged"] == 0 and st["newton_iters"] == 3 and len(hist) == 4


def test_ns_linear_solve_equals_direct_solve(gpu_ctx):
    """GMRES on the Dirichlet-rewritten Navier-Stokes Jacobian at a nonzero iterate."""
    M = 3
    p2, e2, p1, e1, rows, g = _setup(gpu_ctx, M)
    rng = np.random.default_rng(2)
    u = rng.normal(size=3 * len(p2))
    u[rows] = g[rows]
    gpu_ctx.vector_upload(N.VEC_U, u)
    gpu_ctx.assembly_navier_stokes(0.05, 1.0, True, N.NS_JACOBIAN, reassemble=False)
    gpu_ctx.dirichlet_set_system()
    nv, npr = 3 * len(p2), len(p1)
    b = rng.normal(size=nv + npr)
    gpu_ctx.vector_upload(N.VEC_B, b[:nv])
    gpu_ctx.vector_upload(N.VEC_B_P, b[nv:])
    it, rel = gpu_ctx.ns_linear_solve(1e-12, restart=60, maxit=20000)
    assert rel <= 1e-12, (it, rel)
    x = np.concatenate([gpu_ctx.vector_download(N.VEC_X, nv), gpu_ctx.vector_download(N.VEC_X_P, npr)])
    rp, ci, v = gpu_ctx.ns_merged()
    K = sp.csr_matrix((v, ci, rp), shape=(nv + npr, nv + npr))
    xd = spla.spsolve(K.tocsc(), b)
    assert np.abs(x - xd).max() <= 1e-8 * np.abs(xd).max()


def test_solves_ignore_stale_device_memory(gpu_ctx):
    """Fresh work space must not leak into the solve: device memory that held NaN bit
    patterns (freed vectors) is handed out again for the Krylov basis and the residual
    vectors; the solve must write them before any read (y = a x must not read y)."""
    p2, _, e2 = N.mesh_structured(10, 2)
    p1, _, e1 = N.mesh_structured(10, 1)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    gpu_ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    gpu_ctx.graph_build(N.GRAPH_NS)
    for w, n in ((N.VEC_U, 3 * len(p2)), (N.VEC_B, 3 * len(p2)), (N.VEC_P, len(p1)), (N.VEC_B_P, len(p1))):
        gpu_ctx.vector_upload(w, np.full(n, np.nan))
    # the smaller problems below free these buffers when they resize the vectors
    test_ns_linear_solve_equals_direct_solve(gpu_ctx)
    test_newton_matches_oracle_loop(gpu_ctx, 0)
