"""GPU parity: P2 hyperelastic assembly (skr1 / skr2 / skr3 laws) vs the oracle.

The oracle's element (oracle/oracle.cpp nh_skr) agrees with the reference's own
AceGen Fortran to ~1e-15 (tests/test_oracle_pinning.py); here the GPU tangent and
residual are compared with the oracle's FE::assemblyNonLinearElasticity /
assemblyLinearElasticity on jittered, orientation-corrected P2 cubes.
"""
import numpy as np
import pytest

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W
from helpers import assert_graph_equal, fp32_local_check, parity_ratio, rel_err_rowfloor, vec_rel_err

pytestmark = pytest.mark.gpu

EMOD, NU = 1000.0, 0.3


def _setup(ctx, M, seed=0, orient=True, scale=0.05):
    pts, _, el = W.jittered_p2_mesh(M, seed, orient=orient)
    u = W.smoothed_displacement(pts, el, seed + 1, scale)
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    ctx.vector_upload(N.VEC_U, u)
    return pts, el, u


@pytest.mark.parametrize("M", [1, 2])
def test_vector_graph_bit_exact(gpu_ctx, M):
    pts, el, u = _setup(gpu_ctx, M)
    ref, _ = O.elasticity(0, EMOD, NU, pts, el, u)
    assert_graph_equal(gpu_ctx.block_graph(N.BLOCK_A), ref[:2])
    assert gpu_ctx.block_info(N.BLOCK_A)[2] == 9 * (230 * M ** 3 + 138 * M ** 2 + 24 * M + 1)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_elasticity_fp64(gpu_ctx, model):
    M = 2
    pts, el, u = _setup(gpu_ctx, M)
    ref, rref = O.elasticity(model, EMOD, NU, pts, el, u)
    if model == 0:
        gpu_ctx.assembly_linear_elasticity(EMOD, NU, N.FP64)
    else:
        gpu_ctx.assembly_nonlinear_elasticity(1 if model == 1 else 2, EMOD, NU, N.FP64)
    v = gpu_ctx.block_values(N.BLOCK_A)
    ratio = parity_ratio(v, ref[2], ref[0])
    assert ratio <= 1.0, (ratio, rel_err_rowfloor(v, ref[2], ref[0]))
    if model != 0:
        r = gpu_ctx.vector_download(N.VEC_RES_U, 3 * pts.shape[0])
        assert vec_rel_err(r, rref) <= 1e-12


def test_elasticity_unoriented_mesh(gpu_ctx):
    """The raw structured mesh keeps FEDDLib's alternating orientation; the reference
    multiplies by the signed AceGen determinant, and so must the GPU path."""
    pts, el, u = _setup(gpu_ctx, 2, orient=False)
    ref, rref = O.elasticity(1, EMOD, NU, pts, el, u)
    gpu_ctx.assembly_nonlinear_elasticity(1, EMOD, NU, N.FP64)
    v = gpu_ctx.block_values(N.BLOCK_A)
    assert parity_ratio(v, ref[2], ref[0]) <= 1.0
    r = gpu_ctx.vector_download(N.VEC_RES_U, 3 * pts.shape[0])
    assert vec_rel_err(r, rref) <= 1e-12


@pytest.mark.parametrize("model", [1, 2])
def test_elasticity_fp32_local(gpu_ctx, model):
    pts, el, u = _setup(gpu_ctx, 3)
    ref, rref = O.elasticity(model, EMOD, NU, pts, el, u)
    gpu_ctx.assembly_nonlinear_elasticity(1 if model == 1 else 2, EMOD, NU, N.FP32_LOCAL)
    v = gpu_ctx.block_values(N.BLOCK_A)
    with O.element_scale():
        scale, _ = O.elasticity(model, EMOD, NU, pts, el, u)
    fp32_local_check(v, ref, scale)
    # the residual stays FP64 in the FP32-local path
    r = gpu_ctx.vector_download(N.VEC_RES_U, 3 * pts.shape[0])
    assert vec_rel_err(r, rref) <= 1e-12


def test_zero_displacement_is_linear_elasticity(gpu_ctx):
    """At u = 0 all three laws give the small-strain operator 2 mu eps:eps + lam tr tr;
    the residual vanishes."""
    pts, el, _ = _setup(gpu_ctx, 2)
    gpu_ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    vals = []
    for model in (1, 2):
        gpu_ctx.assembly_nonlinear_elasticity(model, EMOD, NU, N.FP64)
        vals.append(gpu_ctx.block_values(N.BLOCK_A))
        r = gpu_ctx.vector_download(N.VEC_RES_U, 3 * pts.shape[0])
        assert np.max(np.abs(r)) < 1e-12 * np.max(np.abs(vals[-1]))
    rp, _ = gpu_ctx.block_graph(N.BLOCK_A)
    assert parity_ratio(vals[0], vals[1], rp, rtol=1e-11) <= 1.0


@pytest.mark.parametrize("model", [1, 2])
def test_elasticity_efield(gpu_ctx, model):
    """Per-element Young's modulus (FE_def.hpp:538-614): two E values on a checker of
    elements; the oracle assembles each element group with its own E and the two
    partial systems are summed."""
    import scipy.sparse as sp
    pts, el, u = _setup(gpu_ctx, 2)
    nel = el.shape[0]
    group = (np.arange(nel) % 3 == 0)
    E = np.where(group, 2500.0, 700.0)
    n = 3 * pts.shape[0]
    mats, res = [], np.zeros(n)
    for mask, Ev in ((group, 2500.0), (~group, 700.0)):
        (rp, ci, v), r = O.elasticity(model, Ev, NU, pts, np.ascontiguousarray(el[mask]), u)
        mats.append(sp.csr_matrix((v, ci, rp), shape=(n, n)))
        res += r
    ref = (mats[0] + mats[1]).tocsr()
    gpu_ctx.assembly_nonlinear_elasticity_efield(1 if model == 1 else 2, E, NU, N.FP64)
    rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
    v = gpu_ctx.block_values(N.BLOCK_A)
    rows = np.repeat(np.arange(n), np.diff(rp))
    refv = np.asarray(ref[rows, ci]).ravel()
    assert parity_ratio(v, refv, rp) <= 1.0
    r = gpu_ctx.vector_download(N.VEC_RES_U, n)
    assert vec_rel_err(r, res) <= 1e-12
    # constant field == the constant-E entry point
    gpu_ctx.assembly_nonlinear_elasticity_efield(1 if model == 1 else 2, np.full(nel, EMOD), NU, N.FP64)
    v1 = gpu_ctx.block_values(N.BLOCK_A)
    gpu_ctx.assembly_nonlinear_elasticity(1 if model == 1 else 2, EMOD, NU, N.FP64)
    v0 = gpu_ctx.block_values(N.BLOCK_A)
    # equal up to the order of the LDS atomics
    assert np.max(np.abs(v1 - v0)) <= 1e-13 * np.max(np.abs(v0))


def test_efield_rejects_bad_field(gpu_ctx):
    pts, el, u = _setup(gpu_ctx, 1)
    with pytest.raises(N.FeddError):
        gpu_ctx.assembly_nonlinear_elasticity_efield(1, np.full(el.shape[0] - 1, EMOD), NU)
    bad = np.full(el.shape[0], EMOD)
    bad[3] = -1.0
    with pytest.raises(N.FeddError):
        gpu_ctx.assembly_nonlinear_elasticity_efield(1, bad, NU)


_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]


def _iso_jacobian_dets(pts, el, xi):
    """det of the isoparametric P2 map x(xi) = sum_k x_k phi_k(xi) at reference point xi."""
    lam = np.array([1.0 - xi.sum(), xi[0], xi[1], xi[2]])
    dlam = np.vstack([-np.ones(3), np.eye(3)])  # d lambda_m / d xi
    grads = [(4 * lam[i] - 1) * dlam[i] for i in range(4)]
    grads += [4 * (lam[a] * dlam[b] + lam[b] * dlam[a]) for a, b in _EDGES]
    G = np.array(grads)  # (10, 3)
    J = np.einsum("ekd,kj->edj", pts[el], G)
    return np.linalg.det(J)


def _curved_shuffled(M, seed=11, amp=0.08):
    """Isoparametric (curved) P2 elements: the jittered cube with every edge node moved off
    its midpoint by U(+-amp*h), then node ids permuted, elements shuffled and each
    element's vertex order permuted (edge slots follow, orientation flips for odd
    permutations)."""
    pts, _, el = W.jittered_p2_mesh(M, seed, orient=True)
    rng = np.random.default_rng(seed)
    h = 1.0 / M
    edge_nodes = np.unique(el[:, 4:])
    pts = pts.copy()
    pts[edge_nodes] += rng.uniform(-amp * h, amp * h, size=(len(edge_nodes), 3))
    slot = {}
    for m, (a, b) in enumerate(_EDGES):
        slot[(a, b)] = slot[(b, a)] = 4 + m
    v = np.empty_like(el)
    for e in range(el.shape[0]):
        pi = rng.permutation(4)
        v[e, :4] = el[e, pi]
        for m, (a, b) in enumerate(_EDGES):
            v[e, 4 + m] = el[e, slot[(pi[a], pi[b])]]
    v = v[rng.permutation(el.shape[0])]
    q = rng.permutation(pts.shape[0])
    inv = np.argsort(q)
    pts = pts[q]
    el = np.ascontiguousarray(inv[v].astype(np.int32))
    mid = np.stack([0.5 * (pts[el[:, a]] + pts[el[:, b]]) for a, b in _EDGES], axis=1)
    assert np.max(np.abs(pts[el[:, 4:]] - mid)) > 1e-2 * h  # not affine
    # the map stays regular with one sign per element at every point of both rules
    for xi in ([0.25, 0.25, 0.25], [1 / 6, 1 / 6, 1 / 6], [0.5, 1 / 6, 1 / 6], [0.1, 0.2, 0.6], [0, 0, 0]):
        d = _iso_jacobian_dets(pts, el, np.array(xi))
        d0 = _iso_jacobian_dets(pts, el, np.array([0.25, 0.25, 0.25]))
        assert np.all(d * d0 > 0)
    d0 = _iso_jacobian_dets(pts, el, np.array([0.25, 0.25, 0.25]))
    assert (d0 > 0).any() and (d0 < 0).any()
    return pts, el


@pytest.mark.parametrize("model", [0, 1, 2])
def test_elasticity_curved_shuffled_mesh(gpu_ctx, model):
    """Curved (isoparametric) P2 elements in any numbering, element order and vertex order:
    the laws integrate with the geometry Jacobian of all 10 nodes at every point
    (NeoHookQuadraticTets3.f:80-175), so the row gather's relabelling must carry each
    point's own Jacobian."""
    pts, el = _curved_shuffled(2)
    u = W.smoothed_displacement(pts, el, 3, 0.05)
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    gpu_ctx.vector_upload(N.VEC_U, u)
    ref, rref = O.elasticity(model, EMOD, NU, pts, el, u)
    assert_graph_equal(gpu_ctx.block_graph(N.BLOCK_A), ref[:2])
    with O.element_scale():
        scale, _ = O.elasticity(model, EMOD, NU, pts, el, u)
    for prec in (N.FP64, N.FP32_LOCAL):
        if model == 0:
            gpu_ctx.assembly_linear_elasticity(EMOD, NU, prec)
        else:
            gpu_ctx.assembly_nonlinear_elasticity(model, EMOD, NU, prec)
        v = gpu_ctx.block_values(N.BLOCK_A)
        if prec == N.FP64:
            ratio = parity_ratio(v, ref[2], ref[0])
            assert ratio <= 1.0, (ratio, rel_err_rowfloor(v, ref[2], ref[0]))
        else:
            fp32_local_check(v, ref, scale)
        if model != 0:
            r = gpu_ctx.vector_download(N.VEC_RES_U, 3 * pts.shape[0])
            assert vec_rel_err(r, rref) <= 1e-12
