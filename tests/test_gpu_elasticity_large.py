"""Full-size parity of the hyperelastic assembly: 32^3x6 = 196,608 jittered P2 tets
(interior vertices moved by up to 0.2 h, smoothed displacement iterate).

As for the NS matrices, sampled rows are complete once every element touching
their node is assembled: the oracle (FE::assemblyNonLinearElasticity restated,
AceGen-pinned element) assembles only those elements.

Two iterates:
* u scaled by 2/M, so that grad u stays as small as in the small-mesh cases
  (max ~0.05): FP64 rows satisfy the 1e-12 rule, FP32-local rows helpers.fp32_local_check
  (1e-6 per matrix, per entry the 1e-5 rule plus the FP32 summation bound), the residual
  (always FP64) 1e-12;
* the configured iterate, 0.05 N(0,1) per dof and one smoothing pass, not scaled with
  h: at this size grad u reaches 3 and 1.4% of the elements invert.  Single elements
  (the most strained ones) still agree with the oracle to 1e-9 of their largest entry
  (measured <= 1.3e-10; AceGen against its own restatement differs by the same order
  at such strains).  Assembled rows sum large contributions of strongly deformed
  neighbours that cancel, which amplifies that level relative to the row: rows sampled
  among nodes whose incident elements keep 0.5 <= det F <= 2 (vertex-gradient
  estimate) are bounded by |a - b| <= 1e-5 max_row |b| (measured up to 3.4e-6).  The
  FP32-local rows of the same iterate are reported as quantiles of that figure: the
  element tangents there cancel by orders of magnitude (B^T CC B with CC from the FP64
  law), which FP32 products cannot resolve; the median row is bounded by 5e-4.
"""
import numpy as np
import pytest

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W
from helpers import fp32_local_check, incident_elements, subset_rows_parity

pytestmark = pytest.mark.gpu

EMOD, NU = 1000.0, 0.3
M = 32


@pytest.fixture(scope="module")
def big_problem():
    pts, _, el = W.jittered_p2_mesh(M, 0)
    u = W.smoothed_displacement(pts, el, 1)
    X, U = pts[el[:, :4]], u[el[:, :4]]
    Dx = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], 2)
    Du = np.stack([U[:, 1] - U[:, 0], U[:, 2] - U[:, 0], U[:, 3] - U[:, 0]], 2)
    J = np.linalg.det(np.eye(3) + Du @ np.linalg.inv(Dx))
    bad = np.zeros(pts.shape[0], bool)
    bad[el[(J < 0.5) | (J > 2.0)].ravel()] = True
    ok = np.nonzero(~bad)[0]
    assert ok.size > 0.5 * pts.shape[0]
    nodes = np.sort(np.random.default_rng(5).choice(ok, 300, replace=False))
    S = incident_elements(el, nodes)
    return pts, el, u, nodes, np.ascontiguousarray(el[S])


def _row_errs(got, ref, rows):
    """max |a - b| / max_row |b| of each of `rows`."""
    gr, gc, gv = got
    rr, rc, rv = ref
    out = np.zeros(len(rows))
    for k, r in enumerate(rows):
        a, b = slice(gr[r], gr[r + 1]), slice(rr[r], rr[r + 1])
        assert np.array_equal(gc[a], rc[b])
        out[k] = np.max(np.abs(gv[a] - rv[b])) / np.max(np.abs(rv[b]))
    return out


def _assemble(ctx, model, prec):
    if model == 0:
        ctx.assembly_linear_elasticity(EMOD, NU, prec)
    else:
        ctx.assembly_nonlinear_elasticity(1 if model == 1 else 2, EMOD, NU, prec)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_elasticity_full_size_rows(gpu_ctx, big_problem, model):
    pts, el, u, nodes, sub = big_problem
    gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    gpu_ctx.graph_build(N.GRAPH_VECTOR)
    rows = np.sort((3 * nodes[:, None] + np.arange(3)).ravel())
    rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
    assert gpu_ctx.block_info(N.BLOCK_A)[2] == 9 * (230 * M ** 3 + 138 * M ** 2 + 24 * M + 1)
    # iterate with small-mesh strains: the full parity rules
    us = u * (2.0 / M)
    gpu_ctx.vector_upload(N.VEC_U, us)
    ref, rres = O.elasticity(model, EMOD, NU, pts, sub, us)
    with O.element_scale():
        scale, _ = O.elasticity(model, EMOD, NU, pts, sub, us)
    for prec in (N.FP64, N.FP32_LOCAL):
        _assemble(gpu_ctx, model, prec)
        v = gpu_ctx.block_values(N.BLOCK_A)
        if prec == N.FP64:
            worst = subset_rows_parity((rp, ci, v), ref, rows)
            assert worst <= 1.0, (model, worst)
        else:
            got = np.zeros(len(ref[2]))  # the GPU's sampled rows on the oracle's pattern
            for r in rows:
                assert np.array_equal(ci[rp[r]:rp[r + 1]], ref[1][ref[0][r]:ref[0][r + 1]])
                got[ref[0][r]:ref[0][r + 1]] = v[rp[r]:rp[r + 1]]
            fp32_local_check(got, ref, scale, rows=rows)
        if model != 0:
            r = gpu_ctx.vector_download(N.VEC_RES_U, 3 * pts.shape[0])[rows]
            assert np.max(np.abs(r - rres[rows])) <= 1e-12 * np.max(np.abs(rres[rows])), (model, prec)
    # the configured iterate (large strains)
    gpu_ctx.vector_upload(N.VEC_U, u)
    ref, rres = O.elasticity(model, EMOD, NU, pts, sub, u)
    errs = {}
    for prec in (N.FP64, N.FP32_LOCAL):
        _assemble(gpu_ctx, model, prec)
        v = gpu_ctx.block_values(N.BLOCK_A)
        assert np.all(np.isfinite(v))
        errs[prec] = _row_errs((rp, ci, v), ref, rows)
    q = (50, 90, 99, 100)
    for prec, name in ((N.FP64, "FP64"), (N.FP32_LOCAL, "FP32-local")):
        print("model %d, configured iterate, %s: max |a-b| / max_row |b| per row, quantiles %s: %s" % (
            model, name, q, " ".join("%.1e" % x for x in np.percentile(errs[prec], q))))
    assert np.max(errs[N.FP64]) <= 1e-5, model
    # FP32-local rows: median measured 7.4e-5..8.0e-5 of the row (90th percentile 1e-2)
    assert np.percentile(errs[N.FP32_LOCAL], 50) <= 5e-4, model


@pytest.mark.parametrize("model", [0, 1, 2])
def test_most_strained_elements(gpu_ctx, big_problem, model):
    """Element by element at the configured (large-strain) iterate."""
    pts, el, u, _, _ = big_problem
    X, U = pts[el[:, :4]], u[el[:, :4]]
    Dx = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], 2)
    Du = np.stack([U[:, 1] - U[:, 0], U[:, 2] - U[:, 0], U[:, 3] - U[:, 0]], 2)
    gmax = np.abs(Du @ np.linalg.inv(Dx)).reshape(len(el), -1).max(1)
    for e in np.argsort(-gmax)[:6]:
        sub = np.ascontiguousarray(el[[e]])
        gpu_ctx.mesh_upload(N.MESH_PRIMARY, pts, sub)
        gpu_ctx.graph_build(N.GRAPH_VECTOR)
        gpu_ctx.vector_upload(N.VEC_U, u)
        _assemble(gpu_ctx, model, N.FP64)
        (orp, oci, ov), _ = O.elasticity(model, EMOD, NU, pts, sub, u)
        rp, ci = gpu_ctx.block_graph(N.BLOCK_A)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        err = np.max(np.abs(gpu_ctx.block_values(N.BLOCK_A) - ov)) / np.max(np.abs(ov))
        assert err <= 1e-9, (model, int(e), gmax[e], err)
