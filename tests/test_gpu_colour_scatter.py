"""GPU parity of the coloured element scatter (FEDD_METHOD_COLOUR_SCATTER).

The scatter indices (node-level CRS position of every element block) are bit-exact
against positions looked up in the oracle's graph; the matrices it assembles meet the
oracle's parity rule of their precision (FP64: helpers.parity_ratio at 1e-12; FP32-local:
helpers.fp32_local_check), equal the row gather's to rounding, and repeat bit for bit
(each value receives its contributions colour by colour in launch order).
"""
import numpy as np
import pytest

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W
from helpers import fp32_local_check, parity_ratio, vec_rel_err

pytestmark = pytest.mark.gpu

NU, RHO = 0.01, 1.0


@pytest.fixture
def scatter_ctx(gpu_ctx):
    gpu_ctx.set_assembly_method(N.METHOD_COLOUR_SCATTER)
    yield gpu_ctx
    gpu_ctx.set_assembly_method(N.METHOD_ROW_GATHER)


def _setup(ctx, M, seed=0):
    p2, e2, p1, e1, u, p = W.ns_problem(M, seed)
    p = np.random.default_rng(seed + 7).standard_normal(p1.shape[0])
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    ctx.graph_build(N.GRAPH_NS)
    ctx.vector_upload(N.VEC_U, u)
    ctx.vector_upload(N.VEC_P, p)
    return p2, e2, p1, e1, u, p


def _node_keys(rowptr, colind, nrows_node, rows_per_node, cols_per_node, ncols_node):
    """Sorted keys row * ncols + col of the node-level graph of a dof-level CRS graph, and
    the node-level row pointer (position of a node row's first block)."""
    starts = rowptr[0:rows_per_node * nrows_node + 1:rows_per_node]
    k0 = starts // (rows_per_node * cols_per_node)
    keys = []
    for i in range(nrows_node):
        cols = colind[starts[i]:rowptr[rows_per_node * i + 1]:cols_per_node] // cols_per_node
        keys.append(i * ncols_node + cols.astype(np.int64))
    return np.concatenate(keys), k0


def _expected(keys, rows, cols, ncols_node):
    want = rows.astype(np.int64) * ncols_node + cols
    k = np.searchsorted(keys, want)
    assert np.array_equal(keys[k], want), "an element block is missing from the graph"
    return k


@pytest.mark.parametrize("M", [1, 3])
def test_scatter_indices_bit_exact(scatter_ctx, M):
    p2, e2, p1, e1, u, p = _setup(scatter_ctx, M)
    ref = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, NU, RHO, True, 0, False)
    ne, nv, npr = e2.shape[0], p2.shape[0], p1.shape[0]
    # A: velocity node a, velocity node b
    keys, _ = _node_keys(ref["A"][0], ref["A"][1], nv, 3, 3, nv)
    rows = np.repeat(e2[:, :, None], 10, axis=2)
    cols = np.repeat(e2[:, None, :], 10, axis=1)
    exp = _expected(keys, rows.ravel(), cols.ravel(), nv)
    assert np.array_equal(scatter_ctx.scatter_indices(0, ne), exp)
    # B^T: velocity node a, pressure vertex m
    keys, _ = _node_keys(ref["BT"][0], ref["BT"][1], nv, 3, 1, npr)
    rows = np.repeat(e2[:, :, None], 4, axis=2)
    cols = np.repeat(e1[:, None, :], 10, axis=1)
    exp = _expected(keys, rows.ravel(), cols.ravel(), npr)
    assert np.array_equal(scatter_ctx.scatter_indices(1, ne), exp)
    # B: pressure vertex m, velocity node b
    keys, _ = _node_keys(ref["B"][0], ref["B"][1], npr, 1, 3, nv)
    rows = np.repeat(e1[:, :, None], 10, axis=2)
    cols = np.repeat(e2[:, None, :], 4, axis=1)
    exp = _expected(keys, rows.ravel(), cols.ravel(), nv)
    assert np.array_equal(scatter_ctx.scatter_indices(2, ne), exp)
    # the schedule runs over the library's colouring: exported, conflict-free
    nc = scatter_ctx.colour_build()
    cptr, perm = scatter_ctx.colour_export(nc, ne)
    for k in range(nc):
        nodes = e2[perm[cptr[k]:cptr[k + 1]]].ravel()
        assert len(np.unique(nodes)) == len(nodes)


@pytest.mark.parametrize("mode,newton,prec,reassemble", [
    (N.NS_JACOBIAN, True, N.FP64, False),
    (N.NS_FIXEDPOINT, True, N.FP64, False),
    (N.NS_JACOBIAN, False, N.FP64, True),
    (N.NS_JACOBIAN, True, N.FP32_LOCAL, False),
])
def test_scatter_matrices_match_the_oracle(scatter_ctx, mode, newton, prec, reassemble):
    M = 3
    p2, e2, p1, e1, u, p = _setup(scatter_ctx, M, seed=2)
    mat_mode = 0 if mode == N.NS_JACOBIAN else 1
    ref = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, NU, RHO, newton, mat_mode, reassemble)
    with O.element_scale():
        scale = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, NU, RHO, newton, mat_mode, reassemble)
    scatter_ctx.assembly_navier_stokes(NU, RHO, newton, N.NS_JACOBIAN, reassemble=False)  # B, B^T first
    before_bt = scatter_ctx.block_values(N.BLOCK_BT)
    scatter_ctx.assembly_navier_stokes(NU, RHO, newton, mode, reassemble=reassemble, precision=prec)
    names = [(N.BLOCK_A, "A")] + ([] if reassemble else [(N.BLOCK_BT, "BT"), (N.BLOCK_B, "B")])
    for blk, name in names:
        v = scatter_ctx.block_values(blk)
        if prec == N.FP64:
            assert parity_ratio(v, ref[name][2], ref[name][0]) <= 1.0, name
        else:
            fp32_local_check(v, ref[name], scale[name])
    if reassemble:  # addFeBlock only: B^T keeps its values
        assert np.array_equal(scatter_ctx.block_values(N.BLOCK_BT), before_bt)


def test_scatter_fused_mode_and_residual(scatter_ctx):
    """A fused mode: the scattered matrix and the residual of FEDD_NS_RHS."""
    M = 3
    coeff = (0.5, -2.0, 3.0, 1.0)
    p2, e2, p1, e1, u, p = _setup(scatter_ctx, M, seed=4)
    ref = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, NU, RHO, True, 0, False, coeff)
    res = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, NU, RHO, True, 2, True, coeff)
    scatter_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN_RHS, reassemble=False, coeff=coeff)
    for blk, name in [(N.BLOCK_A, "A"), (N.BLOCK_BT, "BT"), (N.BLOCK_B, "B")]:
        assert parity_ratio(scatter_ctx.block_values(blk), ref[name][2], ref[name][0]) <= 1.0, name
    assert vec_rel_err(scatter_ctx.vector_download(N.VEC_RES_U, 3 * p2.shape[0]), res["res_u"]) <= 1e-12
    assert vec_rel_err(scatter_ctx.vector_download(N.VEC_RES_P, p1.shape[0]), res["res_p"]) <= 1e-12


def test_scatter_equals_row_gather_and_repeats_bitwise(gpu_ctx):
    """At 8^3 x 6: the scattered matrices equal the row gather's to 1e-14 of each block's
    largest value, and five scattered assemblies give identical bits."""
    M = 8
    _setup(gpu_ctx, M, seed=1)
    gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False)
    gather = [gpu_ctx.block_values(b) for b in (N.BLOCK_A, N.BLOCK_BT, N.BLOCK_B)]
    gpu_ctx.set_assembly_method(N.METHOD_COLOUR_SCATTER)
    try:
        _setup(gpu_ctx, M, seed=1)
        runs = []
        for _ in range(5):
            gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False)
            runs.append([gpu_ctx.block_values(b) for b in (N.BLOCK_A, N.BLOCK_BT, N.BLOCK_B)])
    finally:
        gpu_ctx.set_assembly_method(N.METHOD_ROW_GATHER)
    for g, s in zip(gather, runs[0]):
        assert np.max(np.abs(g - s)) <= 1e-14 * np.max(np.abs(g))
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_scatter_indices_need_the_schedule(gpu_ctx):
    _setup(gpu_ctx, 1)  # built with the row gather
    with pytest.raises(N.FeddError) as e:
        gpu_ctx.scatter_indices(0, 6)
    assert e.value.code == -3
    with pytest.raises(N.FeddError):
        gpu_ctx.set_assembly_method(2)


@pytest.mark.parametrize("curved", [False, True])
def test_scatter_on_an_unstructured_mesh(scatter_ctx, curved):
    """Jittered, renumbered, shuffled and reoriented elements (edge nodes off the midpoints
    when curved): every NS mode by the coloured scatter meets the oracle, and the scatter
    indices are the oracle graph's positions."""
    from test_gpu_ns import _unstructured_modes, _unstructured_ns
    p2, e2, p1, e1, u, p = _unstructured_ns(3, curved=curved)
    _unstructured_modes(scatter_ctx, p2, e2, p1, e1, u, p)
    scatter_ctx.graph_build(N.GRAPH_NS)  # the helper ends on a scalar (Laplace) graph
    ref = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, NU, RHO, True, 0, False)
    keys, _ = _node_keys(ref["A"][0], ref["A"][1], p2.shape[0], 3, 3, p2.shape[0])
    rows = np.repeat(e2[:, :, None], 10, axis=2)
    cols = np.repeat(e2[:, None, :], 10, axis=1)
    exp = _expected(keys, rows.ravel(), cols.ravel(), p2.shape[0])
    assert np.array_equal(scatter_ctx.scatter_indices(0, e2.shape[0]), exp)
