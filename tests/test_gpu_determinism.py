"""fedd_set_reproducible: repeated assemblies of the same state give bit-identical values.

With the option on, the row-gather plans lay a chunk's (node, element) pairs out in
segments of 64, one wave's worth, and never let one node's pairs straddle two segments;
every row of a chunk is then accumulated by one wave in a fixed order, and the matrices,
residuals and load vectors repeat bit for bit (Tpetra's fillComplete sums duplicates in a
fixed order too).  Each case assembles the state several times on one context and
compares the raw bits of every value with the first assembly, and checks the values
against the oracle.  The Navier-Stokes, pressure and Laplace plans are segmented; the
hyperelastic plan is not (its case is skipped, see the reason).
"""
import numpy as np
import pytest

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W

from helpers import fp32_local_check, parity_ratio

pytestmark = pytest.mark.gpu


@pytest.fixture
def repro_ctx(gpu_ctx):
    gpu_ctx.set_reproducible(True)
    yield gpu_ctx
    gpu_ctx.set_reproducible(False)

REPEATS = 6


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def ns_runs(ctx, M, precision, repeats=REPEATS):
    """Values of A, B^T, B and both residuals from `repeats` fused Jacobian + residual passes."""
    p2, e2, p1, e1, u, p = W.ns_problem(M, 0)
    p = np.random.default_rng(3).standard_normal(p.shape[0])
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    ctx.graph_build(N.GRAPH_NS)
    ctx.vector_upload(N.VEC_U, u)
    ctx.vector_upload(N.VEC_P, p)
    out = []
    for _ in range(repeats):
        ctx.assembly_navier_stokes(0.01, 1.0, True, N.NS_JACOBIAN_RHS, reassemble=False, precision=precision,
                                   coeff=(1.0, 0.5, -1.0, 1.0))
        out.append([ctx.block_values(b) for b in (N.BLOCK_A, N.BLOCK_BT, N.BLOCK_B)] +
                   [ctx.vector_download(N.VEC_RES_U, 3 * p2.shape[0]), ctx.vector_download(N.VEC_RES_P, p1.shape[0])])
    return out


def elasticity_runs(ctx, M, precision, scale, repeats=REPEATS):
    """Tangent values and residual of `repeats` skr2 assemblies on the jittered P2 cube,
    plus the FP32-local guard counts of the last one."""
    pts, _, el = W.jittered_p2_mesh(M, seed=0)
    u = W.smoothed_displacement(pts, el, seed=1, scale=scale)
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    ctx.vector_upload(N.VEC_U, u.ravel())
    out = []
    for _ in range(repeats):
        ctx.assembly_nonlinear_elasticity(1, 1000.0, 0.3, precision=precision)
        out.append([ctx.block_values(N.BLOCK_A), ctx.vector_download(N.VEC_RES_U, 3 * pts.shape[0])])
    counts = ctx.elasticity_fp64_counts() if precision == N.FP32_LOCAL else (0, 0)
    return out, counts


def laplace_runs(ctx, M, repeats=REPEATS):
    p1, _, e1 = N.mesh_structured(M, 1)
    ctx.mesh_upload(N.MESH_PRIMARY, p1, e1)
    ctx.graph_build(N.GRAPH_SCALAR)
    out = []
    for _ in range(repeats):
        ctx.assembly_laplace()
        ctx.assembly_rhs(1.0)
        out.append([ctx.block_values(N.BLOCK_A), ctx.vector_download(N.VEC_RHS, p1.shape[0])])
    return out


def mismatches(runs):
    """Number of values (over all arrays) whose bits differ from the first run's."""
    first = runs[0]
    return sum(int(np.count_nonzero(_bits(a) != _bits(b))) for r in runs[1:] for a, b in zip(first, r))


@pytest.mark.parametrize("precision", [N.FP64, N.FP32_LOCAL])
def test_ns_jacobian_and_residual_repeat_bitwise(repro_ctx, precision):
    runs = ns_runs(repro_ctx, 16, precision)
    assert mismatches(runs) == 0


@pytest.mark.parametrize("precision", [N.FP64, N.FP32_LOCAL])
def test_reproducible_plans_match_the_oracle(repro_ctx, precision):
    """The segmented plans (padding pairs included) assemble the same values as always."""
    M = 3
    p2, e2, p1, e1, u, p = W.ns_problem(M, 0)
    p = np.random.default_rng(3).standard_normal(p.shape[0])
    out = ns_runs(repro_ctx, M, precision, repeats=1)[0]
    ref = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, 0.01, 1.0, True, 0, False)
    res = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, 0.01, 1.0, True, 2, False, coeff=(1.0, 0.5, -1.0, 1.0))
    with O.element_scale():
        scale = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, 0.01, 1.0, True, 0, False)
    for v, name in zip(out[:3], ("A", "BT", "B")):
        if precision == N.FP64:
            assert parity_ratio(v, ref[name][2], ref[name][0]) <= 1.0, name
        else:
            fp32_local_check(v, ref[name], scale[name])
    for v, b in ((out[3], res["res_u"]), (out[4], res["res_p"])):
        assert np.max(np.abs(v - b)) <= 1e-12 * np.max(np.abs(b))


@pytest.mark.skip(reason="the hyperelastic row gather keeps the rank-interleaved plan, whose rows race between "
                         "waves: the segmented plan faulted the GPU once (k_elas_rows<double>, FE facade test, 2^3 P2 "
                         "cube) and that fault's cause has not been found, so the path is switched off")
@pytest.mark.parametrize("precision", [N.FP64, N.FP32_LOCAL])
def test_elasticity_repeats_bitwise(repro_ctx, precision):
    runs, counts = elasticity_runs(repro_ctx, 12, precision, scale=0.15)
    if precision == N.FP32_LOCAL:
        assert counts[1] > 0, "the state should flag elements for the FP64 guard pass"
    assert mismatches(runs) == 0


def test_laplace_and_load_repeat_bitwise(repro_ctx):
    assert mismatches(laplace_runs(repro_ctx, 16)) == 0
