"""Follower pressure load on the GPU against the oracle (FE::assemblySurfaceIntegralExternal,
FE::assemblyNonlinearSurfaceIntegralExternal, NonLinElasAssFE's A + Kext), and a
load-stepped Newton solve under it.

- Load (FEDD_VEC_FEXT) on the jittered 3^3 x 6 cube at a smooth displacement, pressures per
  flag, against orc_assembly_surface_pressure: 1e-12 of the largest entry.
- Tangent: the elasticity matrix plus Kext on the vector graph against the oracle's
  elasticity matrix plus its Kext: the FP64 parity rule.
- Newton (skr2, z = 0 clamped, pressure 40 on z = 1 ramped in 5 load steps): converges to
  ||r|| < 1e-8 with BiCGStab under multigrid and under block Jacobi, in about 3 updates per
  load step; the oracle's residual -p_int - f_pressure at the GPU's displacement vanishes
  (< 1e-8) on every free row.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W
from helpers import parity_ratio, vec_rel_err

pytestmark = pytest.mark.gpu

EMOD, NU = 1000.0, 0.3
PRESS = np.zeros(7)
PRESS[1], PRESS[4], PRESS[5] = 0.7, 2.0, -1.5


def _setup(ctx, M=3):
    pts, _, el = W.jittered_p2_mesh(M, 0)
    u = W.smoothed_displacement(pts, el, 2, 0.02)
    faces, flags = N.structured_surfaces(pts, el)
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    ctx.vector_upload(N.VEC_U, u)
    ctx.surface_set(faces, flags)
    return pts, el, u, faces, flags


def test_surface_load_matches_oracle(gpu_ctx):
    pts, el, u, faces, flags = _setup(gpu_ctx)
    gpu_ctx.assembly_surface_pressure(PRESS, with_tangent=False, out=N.VEC_FEXT)
    f = gpu_ctx.vector_download(N.VEC_FEXT, 3 * len(pts))
    fo, _ = O.surface_pressure(pts, faces, flags, u, PRESS)
    assert vec_rel_err(f, fo) <= 1e-12
    assert np.count_nonzero(f) > 0
    # at zero displacement, with a table of 5 flags: flag 5 lies beyond it and carries no load
    gpu_ctx.vector_upload(N.VEC_U, np.zeros_like(u))
    gpu_ctx.assembly_surface_pressure(PRESS[:5], out=N.VEC_FEXT)
    f0 = gpu_ctx.vector_download(N.VEC_FEXT, 3 * len(pts))
    p5 = PRESS.copy()
    p5[5:] = 0.0
    fo0, _ = O.surface_pressure(pts, faces, flags, None, p5)
    assert vec_rel_err(f0, fo0) <= 1e-12


def test_surface_tangent_adds_kext(gpu_ctx):
    pts, el, u, faces, flags = _setup(gpu_ctx)
    gpu_ctx.assembly_nonlinear_elasticity(1, EMOD, NU, N.FP64)
    gpu_ctx.assembly_surface_pressure(PRESS, with_tangent=True, out=N.VEC_FPRES)
    rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
    v = gpu_ctx.block_values(N.BLOCK_A)
    (arp, aci, av), _ = O.elasticity(1, EMOD, NU, pts, el, u)
    fo, (krp, kci, kv) = O.surface_pressure(pts, faces, flags, u, PRESS, tangent=True)
    n = 3 * len(pts)
    ref = sp.csr_matrix((av, aci, arp), shape=(n, n)) + sp.csr_matrix((kv, kci, krp), shape=(n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    refv = np.asarray(ref[rows, ci]).ravel()
    assert parity_ratio(v, refv, rp) <= 1.0
    assert vec_rel_err(gpu_ctx.vector_download(N.VEC_FPRES, n), fo) <= 1e-12


def _clamped_problem(M):
    pts, _, el = N.mesh_structured(M, 2)
    el = np.ascontiguousarray(el)
    N.correct_orientation(pts, el)
    base = np.where(np.abs(pts[:, 2]) < 1e-12)[0]
    dofs = np.sort((3 * base[:, None] + np.arange(3)).ravel()).astype(np.int32)
    return pts, el, dofs


@pytest.mark.parametrize("precond", [1, 0])
def test_newton_with_follower_pressure(gpu_ctx, precond):
    M = 4
    pts, el, dofs = _clamped_problem(M)
    faces, flags = N.structured_surfaces(pts, el)
    p = np.zeros(7)
    p[4] = 40.0  # top face, against its outward normal
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * len(pts)))
    gpu_ctx.dirichlet_set(dofs)
    gpu_ctx.surface_set(faces, flags)
    gpu_ctx.newton_pressure(p)
    try:
        st, hist = gpu_ctx.newton_elasticity(model=1, E=EMOD, nu=NU, load_steps=5, abs_tol=1e-8, max_newton=30,
                                             lin_rtol=1e-10, lin_maxit=5000, preconditioner=precond)
    finally:
        gpu_ctx.newton_pressure([])
    u = gpu_ctx.vector_download(N.VEC_U, 3 * len(pts))
    print("\nfollower pressure, preconditioner %d: %d Newton, %d Krylov, ||r|| %.1e" % (
        precond, st["newton_iters"], st["krylov_iters"], st["final_residual"]))
    assert st["converged"] == 1 and st["final_residual"] < 1e-8
    # the full Kext is in the Newton matrix: about 3 updates per load step (measured 15 with
    # multigrid, 20 with block Jacobi, whose inexact solves stop earlier)
    assert st["newton_iters"] <= 22
    assert np.abs(u[2::3]).max() > 0.01  # the top face moves
    # the oracle's residual at the GPU's displacement
    _, res = O.elasticity(1, EMOD, NU, pts, el, u)
    fo, _ = O.surface_pressure(pts, faces, flags, u, p)
    R = -res - fo
    R[dofs] = 0.0
    assert np.abs(R).max() < 1e-8
