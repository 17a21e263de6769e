"""End-to-end pins of the GPU path against the solutions FEDDLib's unit tests store.

The problems of laplaceUnitTest.cpp, linElasUnitTest.cpp and stokesUnitTest.cpp (see
reference_problems.py) are assembled on the GPU through the C ABI.  Laplace and Stokes
are solved directly on the host from the downloaded CRS blocks; linear elasticity is
solved on the GPU (Dirichlet elimination + CG with the multigrid or block-Jacobi
preconditioner).  FP64 results must meet the reference tests' own tolerances; the
FP32-local variant is held to 1e-5 relative.
"""
import numpy as np
import pytest

import reference_problems as RP
from feddlib_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fe", [1, 2])
@pytest.mark.parametrize("prec", [N.FP64, N.FP32_LOCAL])
def test_gpu_laplace_reproduces_reference_solution(gpu_ctx, fe, prec):
    fname, M = RP.LAPLACE[fe]
    pts, _, el = N.mesh_structured(M, fe)
    n = len(pts)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_SCALAR)
    gpu_ctx.assembly_laplace(prec)
    gpu_ctx.assembly_rhs(1.0)
    rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
    A = RP.csr((rp, ci, gpu_ctx.block_values(N.BLOCK_A)), (n, n))
    b = gpu_ctx.vector_download(N.VEC_RHS, n)
    u = RP.solve_with_dirichlet(A, b, np.where(RP.flags(pts) > 0)[0], 0.0)
    g = N.hdf5_read_vector(RP.golden(fname), "solution")
    err = np.abs(u - g).max()
    if prec == N.FP64:
        assert err <= 1e-11, err  # laplaceUnitTest.cpp:170
    else:
        assert err <= 1e-5 * np.abs(g).max(), err


@pytest.mark.parametrize("model,pc", [(0, 1), (1, 1), (2, 0)])
def test_gpu_linelas_reproduces_reference_solution(gpu_ctx, model, pc):
    """Tangent at u=0 of skr1 / skr2 / skr3 = assemblyLinElasXDim; GPU Dirichlet + CG."""
    fname, M = RP.LINELAS
    pts, _, el = N.mesh_structured(M, 2)
    el = np.ascontiguousarray(el)
    N.correct_orientation(pts, el)
    n = 3 * len(pts)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_SCALAR)
    gpu_ctx.assembly_rhs(1.0)
    bs = gpu_ctx.vector_download(N.VEC_RHS, len(pts))
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(n))
    if model == 0:
        gpu_ctx.assembly_linear_elasticity(RP.LINELAS_E, RP.LINELAS_NU)
    else:
        gpu_ctx.assembly_nonlinear_elasticity(model, RP.LINELAS_E, RP.LINELAS_NU)
    rows = RP.node_dofs(np.where(RP.flags(pts) == 1)[0]).astype(np.int32)
    gpu_ctx.dirichlet_set(rows)
    gpu_ctx.apply_dirichlet_symmetric()
    b = np.zeros(n)
    b[0::3] = RP.LINELAS_FORCE * bs
    b[rows] = 0.0
    gpu_ctx.set_preconditioner(pc)
    try:
        gpu_ctx.vector_upload(N.VEC_B, b)
        it, rr = gpu_ctx.linear_solve(1e-13, 5000)
    finally:
        gpu_ctx.set_preconditioner(0)
    u = gpu_ctx.vector_download(N.VEC_X, n)
    g = N.hdf5_read_vector(RP.golden(fname), "solution")
    assert rr <= 1e-13
    assert np.abs(u - g).max() <= 1e-11  # linElasUnitTest.cpp tolerance


@pytest.mark.parametrize("prec", [N.FP64, N.FP32_LOCAL])
def test_gpu_stokes_reproduces_reference_solution(gpu_ctx, prec):
    """Stokes::assemble = Navier-Stokes Jacobian at u = 0, merged block export."""
    fv, fp, M = RP.STOKES
    p2, _, e2 = N.mesh_structured(M, 2)
    p1, _, e1 = N.mesh_structured(M, 1)
    nv, npr = 3 * len(p2), len(p1)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    gpu_ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    gpu_ctx.graph_build(N.GRAPH_NS)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(nv))
    gpu_ctx.vector_upload(N.VEC_P, np.zeros(npr))
    gpu_ctx.assembly_navier_stokes(RP.STOKES_NU, RP.STOKES_RHO, True, N.NS_JACOBIAN, reassemble=False,
                                   precision=prec)
    rp, ci, v = gpu_ctx.ns_merged()
    K = RP.csr((rp, ci, v), (nv + npr, nv + npr))
    rows, vals = RP.stokes_dirichlet(p2)
    x = RP.solve_with_dirichlet(K, np.zeros(nv + npr), rows, vals)
    gv = N.hdf5_read_vector(RP.golden(fv), "velocity")
    gp = N.hdf5_read_vector(RP.golden(fp), "pressure")
    tol = 1e-9 if prec == N.FP64 else 1e-5  # stokesUnitTest.cpp:230
    assert np.abs(x[:nv] - gv).max() <= tol * np.abs(gv).max()
    assert np.abs(x[nv:] - gp).max() <= tol * np.abs(gp).max()


@pytest.mark.parametrize("fe", [1, 2])
@pytest.mark.parametrize("prec", [N.FP64, N.FP32_LOCAL])
def test_gpu_linelas_xdim_matches_oracle(gpu_ctx, fe, prec):
    """fedd_assembly_linear_elasticity_xdim against the oracle's FE::assemblyLinElasXDim,
    entry by entry on the vector graph (FP64 1e-12 parity rule / helpers.fp32_local_check)."""
    import linelas as LE
    from helpers import fp32_local_check, parity_ratio
    pts, _, el = N.mesh_structured(3, fe)
    n = 3 * len(pts)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.assembly_linear_elasticity_xdim(RP.LINELAS_LAMBDA, RP.LINELAS_MU, prec)
    rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
    v = gpu_ctx.block_values(N.BLOCK_A)
    K = LE.linelas_xdim(fe, pts, el, RP.LINELAS_LAMBDA, RP.LINELAS_MU)
    rows = np.repeat(np.arange(n), np.diff(rp))
    ref = np.asarray(K[rows, ci]).ravel()
    assert K.nnz <= len(ci)  # no oracle entry outside the GPU graph
    if prec == N.FP64:
        assert parity_ratio(v, ref, rp) <= 1.0
    else:
        Kscale = LE.linelas_xdim(fe, pts, el, RP.LINELAS_LAMBDA, RP.LINELAS_MU, element="scale")
        scale = np.asarray(Kscale[rows, ci]).ravel()
        fp32_local_check(v, (rp, ci, ref), scale)


@pytest.mark.parametrize("fe,pc", [(1, 0), (2, 1)])
def test_gpu_linelas_xdim_reproduces_reference_solution(gpu_ctx, fe, pc):
    """linElasUnitTest.cpp end to end on the GPU: assemblyLinElasXDim, volume force,
    Dirichlet rows, CG (block Jacobi for P1, multigrid for P2)."""
    _, M = RP.LINELAS
    pts, _, el = N.mesh_structured(M, fe)
    n = 3 * len(pts)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_SCALAR)
    gpu_ctx.assembly_rhs(1.0)
    bs = gpu_ctx.vector_download(N.VEC_RHS, len(pts))
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.assembly_linear_elasticity_xdim(RP.LINELAS_LAMBDA, RP.LINELAS_MU)
    rows = RP.node_dofs(np.where(RP.flags(pts) == 1)[0]).astype(np.int32)
    gpu_ctx.dirichlet_set(rows)
    gpu_ctx.apply_dirichlet_symmetric()
    b = np.zeros(n)
    b[0::3] = RP.LINELAS_FORCE * bs
    b[rows] = 0.0
    gpu_ctx.set_preconditioner(pc)
    try:
        gpu_ctx.vector_upload(N.VEC_B, b)
        it, rr = gpu_ctx.linear_solve(1e-13, 5000)
    finally:
        gpu_ctx.set_preconditioner(0)
    u = gpu_ctx.vector_download(N.VEC_X, n)
    g = N.hdf5_read_vector(RP.golden(RP.LINELAS_FILES[fe]), "solution")
    assert rr <= 1e-13
    assert np.abs(u - g).max() <= 1e-11  # linElasUnitTest.cpp tolerance


def test_gpu_hyperelastic_needs_p2(gpu_ctx):
    pts, _, el = N.mesh_structured(2, 1)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * len(pts)))
    with pytest.raises(N.FeddError, match="P2"):
        gpu_ctx.assembly_nonlinear_elasticity(1, 1000.0, 0.3)
