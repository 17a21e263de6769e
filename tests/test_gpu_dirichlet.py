"""fedd_dirichlet_set_system / fedd_dirichlet_values / fedd_dirichlet_vector against the
oracle's restatement of BCBuilder (oracle/bcbuilder.py), and end to end against the
Stokes and Laplace solutions FEDDLib's unit tests store.

Integer-like work (which entries become 0 or 1, which stay) is compared bit for bit: the
GPU's own assembled values before the rewrite go through the oracle's setSystem and must
equal what the GPU leaves behind.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import bcbuilder as BC
import reference_problems as RP
from feddlib_amd import native as N

pytestmark = pytest.mark.gpu


def _stokes(ctx, prec=N.FP64):
    fv, fp, M = RP.STOKES
    p2, _, e2 = N.mesh_structured(M, 2)
    p1, _, e1 = N.mesh_structured(M, 1)
    nv, npr = 3 * len(p2), len(p1)
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    ctx.graph_build(N.GRAPH_NS)
    ctx.vector_upload(N.VEC_U, np.zeros(nv))
    ctx.vector_upload(N.VEC_P, np.zeros(npr))
    ctx.assembly_navier_stokes(RP.STOKES_NU, RP.STOKES_RHO, True, N.NS_JACOBIAN, reassemble=False, precision=prec)
    rows, vals = RP.stokes_dirichlet(p2)
    g = np.zeros(nv)
    g[rows] = vals
    return p2, nv, npr, rows, g


def _blocks(ctx):
    out = {}
    for key, b in (((0, 0), N.BLOCK_A), ((0, 1), N.BLOCK_BT), ((1, 0), N.BLOCK_B)):
        rp, ci = ctx.block_graph(b)
        out[key] = (rp, ci, ctx.block_values(b))
    return out


def test_set_system_bit_exact_against_oracle(gpu_ctx):
    p2, nv, npr, rows, g = _stokes(gpu_ctx)
    before = _blocks(gpu_ctx)
    gpu_ctx.dirichlet_set(rows.astype(np.int32))
    gpu_ctx.dirichlet_set_system()
    after = _blocks(gpu_ctx)
    expect = BC.set_system(before, {0: rows})
    for key in before:
        assert np.array_equal(after[key][2], expect[key]), key
    # a second call is idempotent
    gpu_ctx.dirichlet_set_system()
    assert np.array_equal(gpu_ctx.block_values(N.BLOCK_A), expect[(0, 0)])


def test_bc_vector_modes_bit_exact(gpu_ctx):
    p2, nv, npr, rows, g = _stokes(gpu_ctx)
    gpu_ctx.dirichlet_set(rows.astype(np.int32))
    rng = np.random.default_rng(7)
    x = rng.normal(size=nv)
    r = rng.normal(size=nv)
    # values absent: homogeneous
    gpu_ctx.vector_upload(N.VEC_RES_U, r)
    gpu_ctx.dirichlet_vector(N.BC_RHS, N.VEC_RES_U)
    assert np.array_equal(gpu_ctx.vector_download(N.VEC_RES_U, nv), BC.bc_vector(0, r, rows, np.zeros(nv)))
    gpu_ctx.dirichlet_values(g)
    gpu_ctx.vector_upload(N.VEC_U, x)
    for mode in (N.BC_RHS, N.BC_MINUS_VECTOR, N.BC_VECTOR_MINUS_BC):
        gpu_ctx.vector_upload(N.VEC_RES_U, r)
        gpu_ctx.dirichlet_vector(mode, N.VEC_RES_U, N.VEC_U)
        got = gpu_ctx.vector_download(N.VEC_RES_U, nv)
        assert np.array_equal(got, BC.bc_vector(mode, r, rows, g, x)), mode
    with pytest.raises(N.FeddError, match="one entry per dof"):
        gpu_ctx.dirichlet_vector(N.BC_RHS, N.VEC_RES_P)


@pytest.mark.parametrize("prec", [N.FP64, N.FP32_LOCAL])
def test_stokes_golden_through_gpu_set_system(gpu_ctx, prec):
    fv, fp, _ = RP.STOKES
    p2, nv, npr, rows, g = _stokes(gpu_ctx, prec)
    gpu_ctx.dirichlet_set(rows.astype(np.int32))
    gpu_ctx.dirichlet_values(g)
    gpu_ctx.dirichlet_set_system()
    gpu_ctx.vector_upload(N.VEC_B, np.zeros(nv))
    gpu_ctx.dirichlet_vector(N.BC_RHS, N.VEC_B)
    rhs = np.concatenate([gpu_ctx.vector_download(N.VEC_B, nv), np.zeros(npr)])
    rp, ci, v = gpu_ctx.ns_merged()
    x = spla.spsolve(sp.csr_matrix((v, ci, rp), shape=(nv + npr, nv + npr)).tocsc(), rhs)
    gv = N.hdf5_read_vector(RP.golden(fv), "velocity")
    gp = N.hdf5_read_vector(RP.golden(fp), "pressure")
    tol = 1e-9 if prec == N.FP64 else 1e-5  # stokesUnitTest.cpp:230
    assert np.abs(x[:nv] - gv).max() <= tol * np.abs(gv).max()
    assert np.abs(x[nv:] - gp).max() <= tol * np.abs(gp).max()


@pytest.mark.parametrize("fe", [1, 2])
def test_laplace_golden_through_gpu_set_system(gpu_ctx, fe):
    fname, M = RP.LAPLACE[fe]
    pts, _, el = N.mesh_structured(M, fe)
    n = len(pts)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_SCALAR)
    gpu_ctx.assembly_laplace(N.FP64)
    gpu_ctx.assembly_rhs(1.0)
    rows = np.where(RP.flags(pts) > 0)[0].astype(np.int32)
    rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
    before = gpu_ctx.block_values(N.BLOCK_A)
    gpu_ctx.dirichlet_set(rows)  # scalar graph: dof = node
    gpu_ctx.dirichlet_set_system()
    gpu_ctx.dirichlet_vector(N.BC_RHS, N.VEC_RHS)
    vals = gpu_ctx.block_values(N.BLOCK_A)
    assert np.array_equal(vals, BC.set_system({(0, 0): (rp, ci, before)}, {0: rows})[(0, 0)])
    b = gpu_ctx.vector_download(N.VEC_RHS, n)
    u = spla.spsolve(sp.csr_matrix((vals, ci, rp), shape=(n, n)).tocsc(), b)
    g = N.hdf5_read_vector(RP.golden(fname), "solution")
    assert np.abs(u - g).max() <= 1e-11  # laplaceUnitTest.cpp:170


def test_dirichlet_set_follows_graph_kind(gpu_ctx):
    pts, _, el = N.mesh_structured(2, 1)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_SCALAR)
    gpu_ctx.dirichlet_set(np.array([0, 1], np.int32))
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    with pytest.raises(N.FeddError, match="another graph kind"):
        gpu_ctx.dirichlet_set_system()
    with pytest.raises(N.FeddError, match="out of range"):
        gpu_ctx.dirichlet_set(np.array([3 * len(pts)], np.int32))
    gpu_ctx.dirichlet_set(np.array([0, 1, 2], np.int32))
    gpu_ctx.dirichlet_set_system()
    gpu_ctx.dirichlet_set(np.zeros(0, np.int32))
