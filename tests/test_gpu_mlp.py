"""GPU tests of the learned tangent surrogate (BF16-MFMA MLP kernels).

Forward and gradient kernels are compared with oracle/mlp.py (float64 network with
the same bfloat16 operand roundings; the oracle's float64 mode is pinned by finite
differences and torch autograd in tests/test_mlp_oracle.py).  Training runs on the
device with the library's own forward/backward/Adam kernels; the trained network is
measured against the exact law on held-out deformation gradients and inside Newton,
where it must beat the network-free surrogate (the constant small-strain tangent
CC0, i.e. a zero network) and fall back to the exact tangent when it fails.
"""
import numpy as np
import pytest

import mlp as R
from feddlib_amd import native as N

pytestmark = pytest.mark.gpu

EMOD, NU = 1000.0, 0.3


@pytest.fixture()
def mlp_ctx(gpu_ctx):
    gpu_ctx.mlp_init(1, EMOD, NU, 0)
    return gpu_ctx


@pytest.mark.parametrize("n", [1, 17, 1003])
def test_forward_matches_oracle(mlp_ctx, n):
    p = R.init_params(7).astype(np.float32)
    mlp_ctx.mlp_set_params(p)
    x = np.random.default_rng(n).normal(size=(n, 9)).astype(np.float32)
    y = mlp_ctx.mlp_forward(x)
    ref = R.forward(p.astype(np.float64), x.astype(np.float64), bf=True)[0]
    scale = np.max(np.abs(ref))
    # tolerance: FP32 accumulation order and an occasional bfloat16 rounding flip of
    # a hidden activation (one bf16 ulp = 2^-8 relative on one of 64 terms)
    assert np.max(np.abs(y - ref)) <= 2e-3 * scale, np.max(np.abs(y - ref)) / scale
    ref64 = R.forward(p.astype(np.float64), x.astype(np.float64))[0]
    assert np.max(np.abs(y - ref64)) <= 5e-2 * scale


def test_params_round_trip_and_padding(mlp_ctx):
    p = mlp_ctx.mlp_params()
    W1, b1, W2, b2, W3, b3 = R.unpack(p)
    assert np.all(W1[:, 9:] == 0) and np.all(W3[21:] == 0) and np.all(b3[21:] == 0)
    assert np.all(b1 == 0) and np.std(W2) > 0.05
    q = R.init_params(3).astype(np.float32)
    mlp_ctx.mlp_set_params(q)
    assert np.array_equal(mlp_ctx.mlp_params(), q)


def test_gradient_matches_oracle(mlp_ctx):
    p = R.init_params(11).astype(np.float32)
    mlp_ctx.mlp_set_params(p)
    rng = np.random.default_rng(1)
    n = 1000  # 62.5 tiles: a ragged last tile
    x = rng.normal(size=(n, 9)).astype(np.float32)
    t = rng.normal(size=(n, 21)).astype(np.float32)
    g, loss = mlp_ctx.mlp_gradient(x, t)
    rl, rg = R.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss - rl) <= 1e-3 * rl
    blocks = [(R.P_W1, R.P_B1), (R.P_B1, R.P_W2), (R.P_W2, R.P_B2), (R.P_B2, R.P_W3), (R.P_W3, R.P_B3),
              (R.P_B3, R.P_TOTAL)]
    for a, b in blocks:
        err = np.linalg.norm(g[a:b] - rg[a:b]) / np.linalg.norm(rg[a:b])
        assert err <= 5e-3, (a, b, err)
    # padding entries carry no gradient
    W1g, _, _, _, W3g, b3g = R.unpack(g)
    assert np.all(W1g[:, 9:] == 0) and np.all(W3g[21:] == 0) and np.all(b3g[21:] == 0)
    # and it is the gradient of the float64 network to bfloat16 accuracy
    _, g64 = R.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert np.linalg.norm(g - g64) <= 3e-2 * np.linalg.norm(g64)


def _held_out(n, seed, sigma=0.1):
    rng = np.random.default_rng(seed)
    F = np.eye(3) + sigma * rng.normal(size=(n, 3, 3))
    return F[np.linalg.det(F) > 0.2]


def test_training_learns_the_tangent(mlp_ctx):
    ctx = mlp_ctx
    ctx.mlp_add_samples_random(400000, 0.1, 2)
    assert ctx.mlp_num_samples() == 400000
    cc0, s_out, s_in = ctx.mlp_norm()
    mu, lam = R.lame(EMOD, NU)
    assert s_in == 0.1 and abs(s_out - (lam + 2 * mu)) < 1e-3
    assert np.allclose(cc0, R.stress_tangent(1, EMOD, NU, np.eye(3)[None])[1][0], rtol=1e-6)
    hist = ctx.mlp_train(5000, 8192, 3e-3, 0)
    assert np.all(np.isfinite(hist))
    assert hist[-50:].mean() < 0.01 * hist[:10].mean()
    F = _held_out(4000, 99)
    x, y = R.normalise(1, EMOD, NU, F)
    pred = ctx.mlp_forward(x.astype(np.float32))
    # the network-free surrogate (y = 0, CC = CC0) has relative error 1 by construction
    err = np.linalg.norm(pred - y) / np.linalg.norm(y)
    assert err < 0.05, err
    # the tangent the assembly uses, CC = CC0 + s_out * y, against the exact law
    _, cc = R.stress_tangent(1, EMOD, NU, F)
    cc_pred = cc0[None].astype(np.float64) + s_out * pred
    assert np.linalg.norm(cc_pred - cc) / np.linalg.norm(cc) < 0.025


def _problem(M):
    pts, _, el = N.mesh_structured(M, 2)
    el = np.ascontiguousarray(el)
    N.correct_orientation(pts, el)
    base = np.where(np.abs(pts[:, 2]) < 1e-12)[0]
    dofs = np.sort((3 * base[:, None] + np.arange(3)).ravel()).astype(np.int32)
    return pts, el, dofs


def _newton(ctx, pts, el, dofs, t, tangent, steps=3, stall=0.5, max_newton=30):
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    ctx.dirichlet_set(dofs)
    ctx.traction_load(2, 1.0, t)
    st, hist = ctx.newton_elasticity(model=1, E=EMOD, nu=NU, load_steps=steps, abs_tol=1e-8, tangent=tangent,
                                     stall_ratio=stall, max_newton=max_newton)
    return st, hist, ctx.vector_download(N.VEC_U, 3 * pts.shape[0])


def test_newton_with_surrogate(mlp_ctx):
    ctx = mlp_ctx
    pts, el, dofs = _problem(3)
    t = np.array([40.0, 0.0, 60.0])
    st_e, _, u_e = _newton(ctx, pts, el, dofs, t, 0)
    ctx.mlp_add_samples_qp()  # solver iterates (final state of the exact solve)
    ctx.mlp_add_samples_random(400000, 0.1, 2)
    ctx.mlp_train(5000, 8192, 3e-3, 0)
    trained = ctx.mlp_params()
    # surrogate on its own (fallback only if a step increases the residual)
    st_m, _, u_m = _newton(ctx, pts, el, dofs, t, 1, stall=1.0, max_newton=60)
    # method without the learned part: zero network, CC = CC0 everywhere
    ctx.mlp_set_params(np.zeros(N.MLP_PARAMS, np.float32))
    st_0, _, u_0 = _newton(ctx, pts, el, dofs, t, 1, stall=1.0, max_newton=60)
    ctx.mlp_set_params(trained)
    st_f, _, u_f = _newton(ctx, pts, el, dofs, t, 1)  # default stall detector
    print("\nexact", st_e, "\ntrained", st_m, "\nzero", st_0, "\ntrained+fallback", st_f)
    for st, u in ((st_m, u_m), (st_f, u_f)):
        assert st["converged"] == 1 and st["final_residual"] < 1e-8
        # exact residual -> the same converged solution
        assert np.max(np.abs(u - u_e)) <= 1e-6 * np.max(np.abs(u_e))
    assert st_m["fallbacks"] == 0
    # what the trained network gains over the zero network: far fewer Newton updates
    assert st_0["converged"] == 0 or st_0["newton_iters"] > 2 * st_m["newton_iters"]
    # and what it costs against the exact tangent (linear instead of quadratic convergence)
    assert st_e["newton_iters"] <= st_m["newton_iters"] <= 3 * st_e["newton_iters"]


def test_newton_falls_back_on_broken_network(mlp_ctx):
    """NaN network: every surrogate step makes the iterate NaN; the step is rejected
    and the exact tangent finishes each load step with the exact run's solution."""
    ctx = mlp_ctx
    pts, el, dofs = _problem(2)
    t = np.array([40.0, 0.0, 60.0])
    st_e, _, u_e = _newton(ctx, pts, el, dofs, t, 0)
    ctx.mlp_set_params(np.full(N.MLP_PARAMS, np.nan, np.float32))
    st_b, _, u_b = _newton(ctx, pts, el, dofs, t, 1)
    assert st_b["fallbacks"] == 3 and st_b["rejected"] == 3 and st_b["converged"] == 1
    assert np.max(np.abs(u_b - u_e)) <= 1e-6 * np.max(np.abs(u_e))
    # one wasted update per load step, otherwise the exact run (+-1 for the order of
    # the floating-point atomics near the tolerance)
    assert abs(st_b["newton_iters"] - (st_e["newton_iters"] + 3)) <= 1


def test_surrogate_rejects_mismatched_law(mlp_ctx):
    pts, el, dofs = _problem(1)
    mlp_ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    mlp_ctx.graph_build(N.GRAPH_VECTOR)
    mlp_ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    with pytest.raises(N.FeddError) as e:
        mlp_ctx.newton_elasticity(model=2, E=EMOD, nu=NU, tangent=1)
    assert e.value.code == -3


def test_standalone_learned_tangent_assembly(mlp_ctx):
    """fedd_assembly_nonlinear_elasticity_tangent: the learned tangent assembled on its
    own, beside the exact one, on the same graph.  The residual is the exact law's in both;
    the trained network's matrix is close to the exact tangent, much closer than the
    network-free CC0 tangent (zero network), and the surrogate refuses another law."""
    ctx = mlp_ctx
    pts, el, dofs = _problem(3)
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    # a smooth deformation with |grad u| <= 0.1, inside the training distribution F = I + N(0, 0.1)
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    u = 0.1 * np.stack([z * x, -0.5 * z * y, 0.3 * x * y], axis=1).ravel()
    ctx.vector_upload(N.VEC_U, u)
    ctx.assembly_nonlinear_elasticity_tangent(1, EMOD, NU, N.FP64, tangent=0)
    K_e = ctx.block_values(N.BLOCK_A)
    r_e = ctx.vector_download(N.VEC_RES_U, u.size)
    ctx.assembly_nonlinear_elasticity(1, EMOD, NU, N.FP64)  # the plain entry point is the exact tangent
    assert np.max(np.abs(ctx.block_values(N.BLOCK_A) - K_e)) <= 1e-12 * np.max(np.abs(K_e))
    ctx.mlp_add_samples_random(400000, 0.1, 2)
    ctx.mlp_train(5000, 8192, 3e-3, 0)
    ctx.assembly_nonlinear_elasticity_tangent(1, EMOD, NU, N.FP64, tangent=1)
    K_m = ctx.block_values(N.BLOCK_A)
    r_m = ctx.vector_download(N.VEC_RES_U, u.size)
    assert np.max(np.abs(r_m - r_e)) <= 1e-12 * np.max(np.abs(r_e))
    err_m = np.linalg.norm(K_m - K_e) / np.linalg.norm(K_e)
    ctx.mlp_set_params(np.zeros(N.MLP_PARAMS, np.float32))
    ctx.assembly_nonlinear_elasticity_tangent(1, EMOD, NU, N.FP64, tangent=1)
    err_0 = np.linalg.norm(ctx.block_values(N.BLOCK_A) - K_e) / np.linalg.norm(K_e)
    print("\nlearned tangent rel. Frobenius error %.3e, zero network %.3e" % (err_m, err_0))
    assert err_m < 0.05 and err_m < 0.5 * err_0
    with pytest.raises(N.FeddError) as e:
        ctx.assembly_nonlinear_elasticity_tangent(2, EMOD, NU, N.FP64, tangent=1)
    assert e.value.code == -3
