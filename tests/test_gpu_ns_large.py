"""GPU parity at full size and HIP-graph replay of the NS assembly.

Full-size matrices are too large for the CPU oracle, but a row is complete as soon
as every element containing its node is assembled: the oracle assembles only the
elements incident to a random sample of nodes (same global numbering) and those
rows must match the GPU's rows of the full-size assembly, by the same parity rule
as the small cases.  Covered: the 40^3x6 Jacobian of the throughput workload (FP64)
and the 64^3x6 FP32-local Jacobian + FP64 residual replayed from a captured HIP
graph over 10 Newton-like iterate updates.
"""
import numpy as np
import pytest

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W
from helpers import (assert_graph_equal, fp32_local_check, incident_elements, parity_ratio,
                     subset_rows_parity)

pytestmark = pytest.mark.gpu

NU, RHO = 0.01, 1.0


def _upload(ctx, p2, e2, p1, e1, u, p):
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    ctx.graph_build(N.GRAPH_NS)
    ctx.vector_upload(N.VEC_U, u)
    ctx.vector_upload(N.VEC_P, p)


def _check_rows(ctx, p2, e2, p1, e1, u, p, vel_nodes, pres_nodes, fp32=False):
    """Oracle on the elements touching the sampled nodes; compare their rows: the FP64
    parity rule, or the FP32-local rule (helpers.fp32_local_check).  Returns the worst
    parity ratio (FP64) or the (row-floor error, per-matrix error, C needed) per block."""
    S = incident_elements(e2, vel_nodes)
    Sp = incident_elements(e1, pres_nodes)
    rows = np.sort((3 * vel_nodes[:, None] + np.arange(3)).ravel())
    worst = {}
    for blk, name, sel, rr in ((N.BLOCK_A, "A", S, rows), (N.BLOCK_BT, "BT", S, rows),
                               (N.BLOCK_B, "B", Sp, np.sort(pres_nodes))):
        sub = (p2, np.ascontiguousarray(e2[sel]), np.ascontiguousarray(e1[sel]), p1.shape[0], u, p, NU, RHO, True, 0,
               False)
        ref = O.ns_assembly(*sub)[name]
        rp, ci = ctx.block_graph(blk)
        v = ctx.block_values(blk)
        if not fp32:
            worst[name] = subset_rows_parity((rp, ci, v), ref, rr)
            continue
        got = np.zeros(len(ref[2]))  # the GPU's values of the sampled rows on the oracle's pattern
        for r in rr:
            assert np.array_equal(ci[rp[r]:rp[r + 1]], ref[1][ref[0][r]:ref[0][r + 1]]), "row %d: columns differ" % r
            got[ref[0][r]:ref[0][r + 1]] = v[rp[r]:rp[r + 1]]
        with O.element_scale():
            scale = O.ns_assembly(*sub)[name]
        worst[name] = fp32_local_check(got, ref, scale, rows=rr)
    return worst


def _sample(n, k, seed):
    return np.sort(np.random.default_rng(seed).choice(n, k, replace=False))


def test_ns_jacobian_full_size_rows(gpu_ctx):
    """40^3x6 = 384,000 tets, FP64: 400 velocity and 200 pressure rows vs the oracle."""
    p2, e2, p1, e1, u, p = W.ns_problem(40, 0)
    _upload(gpu_ctx, p2, e2, p1, e1, u, p)
    gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False)
    vn, pn = _sample(p2.shape[0], 400, 1), _sample(p1.shape[0], 200, 2)
    worst = _check_rows(gpu_ctx, p2, e2, p1, e1, u, p, vn, pn)
    assert max(worst.values()) <= 1.0, worst
    # sizes are the closed-form polynomials
    M = 40
    assert gpu_ctx.block_info(N.BLOCK_A)[2] == 9 * (230 * M ** 3 + 138 * M ** 2 + 24 * M + 1)
    assert gpu_ctx.block_info(N.BLOCK_BT)[2] == 3 * (65 * M ** 3 + 57 * M ** 2 + 15 * M + 1)


def test_ns_jacobian_full_size_rows_fp32_local(gpu_ctx):
    """40^3x6, FP32-local: 400 velocity and 200 pressure rows vs the FP64 oracle, per
    matrix within 1e-6 and per entry within the FP32-local rule of helpers.fp32_local_check."""
    p2, e2, p1, e1, u, p = W.ns_problem(40, 0)
    _upload(gpu_ctx, p2, e2, p1, e1, u, p)
    gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False, precision=N.FP32_LOCAL)
    vn, pn = _sample(p2.shape[0], 400, 1), _sample(p1.shape[0], 200, 2)
    worst = _check_rows(gpu_ctx, p2, e2, p1, e1, u, p, vn, pn, fp32=True)
    for name, (rf, pm, ceff) in worst.items():
        print("\n40^3 FP32-local %s: row-floor %.2e, per-matrix %.2e, C needed %.2f" % (name, rf, pm, ceff))


def _res_rows(ctx, p2, e2, p1, e1, u, p, vel_nodes, pres_nodes):
    Sv = incident_elements(e2, vel_nodes)
    rv = O.ns_assembly(p2, np.ascontiguousarray(e2[Sv]), np.ascontiguousarray(e1[Sv]), p1.shape[0], u, p, NU, RHO,
                       True, 2, False)
    Sp = incident_elements(e1, pres_nodes)
    rpres = O.ns_assembly(p2, np.ascontiguousarray(e2[Sp]), np.ascontiguousarray(e1[Sp]), p1.shape[0], u, p, NU,
                          RHO, True, 2, False)
    gu = ctx.vector_download(N.VEC_RES_U, 3 * p2.shape[0]).reshape(-1, 3)[vel_nodes]
    gp = ctx.vector_download(N.VEC_RES_P, p1.shape[0])[pres_nodes]
    ru = rv["res_u"].reshape(-1, 3)[vel_nodes]
    rp = rpres["res_p"][pres_nodes]
    su, sp_ = np.max(np.abs(ru)), np.max(np.abs(rp))
    return np.max(np.abs(gu - ru)) / su, np.max(np.abs(gp - rp)) / sp_


@pytest.mark.parametrize("fused", [False, True])
def test_graph_replay_small(gpu_ctx, fused):
    """Captured Jacobian + residual (two calls, or the fused mode) replayed with new
    iterates: every replay checked against the oracle, with eager work and
    synchronisations between replays."""
    p2, e2, p1, e1, u, p = W.ns_problem(3, 0)
    _upload(gpu_ctx, p2, e2, p1, e1, u, p)
    gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False)
    gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_RHS, reassemble=False)  # allocates the residuals
    step0 = gpu_ctx.newton_step()
    gpu_ctx.capture_begin()
    if fused:
        gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN_RHS, reassemble=False)
    else:
        gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN, reassemble=False)
        gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_RHS, reassemble=False)
    nodes = gpu_ctx.capture_end()
    assert nodes >= 2  # B rows (+ pressure residual), velocity rows (+ velocity residual)
    assert gpu_ctx.newton_step() == step0 + 1
    for rep in range(3):
        ur = u * (1.0 + 0.25 * rep)
        pr = np.linspace(-1, 1, p1.shape[0]) * rep
        gpu_ctx.vector_upload(N.VEC_U, ur)
        gpu_ctx.vector_upload(N.VEC_P, pr)
        gpu_ctx.graph_launch()
        gpu_ctx.synchronize()
        ref = O.ns_assembly(p2, e2, e1, p1.shape[0], ur, pr, NU, RHO, True, 0, False)
        for blk, name in ((N.BLOCK_A, "A"), (N.BLOCK_BT, "BT"), (N.BLOCK_B, "B")):
            rp, ci = gpu_ctx.block_graph(blk)
            assert_graph_equal((rp, ci), ref[name][:2])
            assert parity_ratio(gpu_ctx.block_values(blk), ref[name][2], rp) <= 1.0, (rep, name)
        rr = O.ns_assembly(p2, e2, e1, p1.shape[0], ur, pr, NU, RHO, True, 2, False)
        gu = gpu_ctx.vector_download(N.VEC_RES_U, 3 * p2.shape[0])
        assert np.max(np.abs(gu - rr["res_u"])) <= 1e-12 * np.max(np.abs(rr["res_u"]))
        # eager work between replays: a FixedPoint assembly overwrites A, the next
        # replay must restore the Newton Jacobian
        gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_FIXEDPOINT, reassemble=False)
        gpu_ctx.synchronize()
    with pytest.raises(N.FeddError):  # host transfers are refused inside a capture
        gpu_ctx.capture_begin()
        try:
            gpu_ctx.vector_upload(N.VEC_U, u)
        finally:
            gpu_ctx.capture_end()


def test_ns_large_fp32_graph_iterations(gpu_ctx):
    """64^3x6 = 1,572,864 tets, FP32-local Jacobian + FP64 residual in one fused pass,
    captured once and replayed for 10 iterate updates; sampled rows vs the oracle at
    iterations 0 and 9."""
    M = 64
    p2, e2, p1, e1, u, p = W.ns_problem(M, 0)
    _upload(gpu_ctx, p2, e2, p1, e1, u, p)
    gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN_RHS, reassemble=False, precision=N.FP32_LOCAL)
    gpu_ctx.capture_begin()
    gpu_ctx.assembly_navier_stokes(NU, RHO, True, N.NS_JACOBIAN_RHS, reassemble=False, precision=N.FP32_LOCAL)
    gpu_ctx.capture_end()
    vn, pn = _sample(p2.shape[0], 200, 3), _sample(p1.shape[0], 100, 4)
    rng = np.random.default_rng(7)
    for it in range(10):
        ui = u * (1.0 + 0.05 * it)
        pi = p + 0.1 * it * rng.standard_normal(p.shape[0])
        gpu_ctx.vector_upload(N.VEC_U, ui)
        gpu_ctx.vector_upload(N.VEC_P, pi)
        gpu_ctx.graph_launch()
        gpu_ctx.synchronize()
        if it in (0, 9):
            _check_rows(gpu_ctx, p2, e2, p1, e1, ui, pi, vn, pn, fp32=True)
            eu, ep = _res_rows(gpu_ctx, p2, e2, p1, e1, ui, pi, vn, pn)
            assert eu <= 1e-12 and ep <= 1e-12, (it, eu, ep)  # the residual stays FP64
