"""GPU Newton solve of the clamped, top-loaded Neo-Hookean cube vs a CPU Newton built
from the oracle's assembly (FE::assemblyNonLinearElasticity restated) and a sparse
direct solve, plus the Krylov solver on its own."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W

pytestmark = pytest.mark.gpu

EMOD, NU = 1000.0, 0.3


def _problem(M):
    pts, _, el = N.mesh_structured(M, 2)
    el = np.ascontiguousarray(el)
    N.correct_orientation(pts, el)
    base = np.where(np.abs(pts[:, 2]) < 1e-12)[0]
    dofs = np.sort((3 * base[:, None] + np.arange(3)).ravel()).astype(np.int32)
    return pts, el, dofs


def _traction_vector(pts, el, t):
    f = np.zeros(pts.shape)
    faces = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    edge = {(0, 1): 4, (1, 2): 5, (0, 2): 6, (0, 3): 7, (1, 3): 8, (2, 3): 9}
    for T in el:
        for fc in faces:
            if np.all(np.abs(pts[T[list(fc)], 2] - 1.0) < 1e-12):
                a, b, c = pts[T[list(fc)]]
                area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
                for i, j in [(fc[0], fc[1]), (fc[1], fc[2]), (fc[0], fc[2])]:
                    f[T[edge[(min(i, j), max(i, j))]]] += area / 3 * np.asarray(t)
    return f.ravel()


def _cpu_newton(pts, el, dofs, fext, steps, tol):
    n = 3 * pts.shape[0]
    u = np.zeros(n)
    free = np.ones(n, bool)
    free[dofs] = False
    for k in range(1, steps + 1):
        lam = k / steps
        for _ in range(50):
            (rp, ci, v), res = O.elasticity(1, EMOD, NU, pts, el, u)
            R = -res - lam * fext
            R[~free] = 0
            if np.linalg.norm(R) < tol:
                break
            K = sp.csr_matrix((v, ci, rp), shape=(n, n))[free][:, free]
            du = np.zeros(n)
            du[free] = spla.spsolve(K.tocsc(), -R[free])
            u += du
    return u


def test_traction_load_matches_closed_form(gpu_ctx):
    pts, el, dofs = _problem(2)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    t = np.array([10.0, -4.0, 30.0])
    gpu_ctx.traction_load(2, 1.0, t)
    f = gpu_ctx.vector_download(N.VEC_FEXT, 3 * pts.shape[0])
    ref = _traction_vector(pts, el, t)
    assert np.max(np.abs(f - ref)) < 1e-14
    assert np.allclose(f.reshape(-1, 3).sum(0), t * 1.0)  # total force = t * area(top) = t


def test_pcg_solves_spd_system(gpu_ctx):
    pts, el, dofs = _problem(3)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    gpu_ctx.assembly_nonlinear_elasticity(1, EMOD, NU)
    gpu_ctx.dirichlet_set(dofs)
    gpu_ctx.apply_dirichlet_symmetric()
    rng = np.random.default_rng(0)
    b = rng.normal(size=3 * pts.shape[0])
    b[dofs] = 0
    gpu_ctx.vector_upload(N.VEC_B, b)
    it, rr = gpu_ctx.linear_solve(1e-10, 5000)
    x = gpu_ctx.vector_download(N.VEC_X, b.size)
    rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
    A = sp.csr_matrix((gpu_ctx.block_values(N.BLOCK_A), ci, rp), shape=(b.size, b.size))
    assert rr <= 1e-10 and it > 0
    assert np.linalg.norm(A @ x - b) <= 1.01e-10 * np.linalg.norm(b)


def test_newton_matches_cpu_newton(gpu_ctx):
    M = 2
    pts, el, dofs = _problem(M)
    t = np.array([40.0, 0.0, 60.0])
    fext = _traction_vector(pts, el, t)
    uref = _cpu_newton(pts, el, dofs, fext, 3, 1e-9)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    gpu_ctx.dirichlet_set(dofs)
    gpu_ctx.traction_load(2, 1.0, t)
    st, hist = gpu_ctx.newton_elasticity(model=1, E=EMOD, nu=NU, load_steps=3, abs_tol=1e-9)
    u = gpu_ctx.vector_download(N.VEC_U, 3 * pts.shape[0])
    assert st["converged"] == 1 and st["final_residual"] < 1e-9
    assert np.max(np.abs(u - uref)) < 1e-8 * np.max(np.abs(uref))
    assert np.max(np.abs(uref)) > 0.01  # a genuinely nonlinear load
    # quadratic convergence at the end of each load step
    assert st["newton_iters"] <= 3 * 8


def test_dirichlet_rows_and_columns(gpu_ctx):
    """fedd_apply_dirichlet_symmetric: constrained rows and columns become identity rows / zero
    columns (the symmetric elimination CG needs; the reference's row-only setSystem is
    fedd_dirichlet_set_system, tested in test_gpu_dirichlet.py), every other entry untouched.
    Two different constraint sets on one graph, then a rebuilt graph: each must use its
    own list of touched rows."""
    pts, el, _ = _problem(2)
    n = 3 * pts.shape[0]
    u = W.smoothed_displacement(pts, el, 5, 0.05)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    rng = np.random.default_rng(4)
    for trial in range(3):
        if trial == 2:
            gpu_ctx.graph_build(N.GRAPH_VECTOR)
        gpu_ctx.vector_upload(N.VEC_U, u)
        gpu_ctx.assembly_nonlinear_elasticity(1, EMOD, NU, N.FP64)
        rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
        before = gpu_ctx.block_values(N.BLOCK_A)
        dofs = np.sort(rng.choice(n, size=n // (5 + trial), replace=False)).astype(np.int32)
        gpu_ctx.dirichlet_set(dofs)
        gpu_ctx.apply_dirichlet_symmetric()
        after = gpu_ctx.block_values(N.BLOCK_A)
        flag = np.zeros(n, bool)
        flag[dofs] = True
        rows = np.repeat(np.arange(n), np.diff(rp))
        hit = flag[rows] | flag[ci]
        expect = np.where(hit, (rows == ci).astype(np.float64), before)
        assert np.array_equal(after, expect), trial
    gpu_ctx.dirichlet_set(np.zeros(0, np.int32))


@pytest.mark.parametrize("criterion", [1, 2])
def test_newton_relative_criteria(gpu_ctx, criterion):
    """Criterion "Residual" (||R_k|| / ||R_0|| < relNonLinTol per load step) and "Update"
    (||du_k|| < relNonLinTol), NonLinearSolver_def.hpp:367-402: they end the load steps
    earlier than the absolute test alone, and the reported figures obey the rule."""
    M = 2
    pts, el, dofs = _problem(M)
    t = np.array([40.0, 0.0, 60.0])
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.dirichlet_set(dofs)
    gpu_ctx.traction_load(2, 1.0, t)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    base, hb = gpu_ctx.newton_elasticity(model=1, E=EMOD, nu=NU, load_steps=3, abs_tol=1e-14, max_newton=12)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    st, hist = gpu_ctx.newton_elasticity(model=1, E=EMOD, nu=NU, load_steps=3, abs_tol=1e-14, max_newton=12,
                                         criterion=criterion, rel_tol=1e-6)
    assert st["converged"] == 1
    assert st["newton_iters"] < base["newton_iters"]
    if criterion == 1:
        assert st["rel_residual"] < 1e-6
    else:
        assert 0 < st["last_update"] < 1e-6
