"""CPU pinning of oracle/mlp.py: the hyperelastic laws and the reference MLP.

Laws: S is checked against finite differences of the strain energy W (dW/dF = F S),
CC against finite differences of S (CC_IJ = dS_I / d eps_J, engineering shear),
and F = I against the small-strain tangent lam d d + 2 mu I.  The MLP's float64
backward is checked against central differences and against torch autograd
(an independent implementation); the bfloat16 mode against the float64 mode.
"""
import numpy as np
import pytest
import torch

import mlp as R

E, NU = 1000.0, 0.3


def _samples(n, seed, sigma=0.15):
    rng = np.random.default_rng(seed)
    F = np.eye(3) + sigma * rng.normal(size=(n, 3, 3))
    return F[np.linalg.det(F) > 0.3]


@pytest.mark.parametrize("model", [0, 1, 2])
def test_stress_is_energy_derivative(model):
    F = _samples(20, model)
    S, _ = R.stress_tangent(model, E, NU, F)
    Sm = np.zeros((len(F), 3, 3))
    for k, (A, B) in enumerate(R.VOIGT):
        Sm[:, A, B] = Sm[:, B, A] = S[:, k]
    P = F @ Sm
    h = 1e-6
    for i in range(3):
        for A in range(3):
            Fp, Fm = F.copy(), F.copy()
            Fp[:, i, A] += h
            Fm[:, i, A] -= h
            fd = (R.energy(model, E, NU, Fp) - R.energy(model, E, NU, Fm)) / (2 * h)
            assert np.allclose(fd, P[:, i, A], rtol=1e-6, atol=1e-6 * E)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_tangent_is_stress_derivative(model):
    F = _samples(10, 10 + model)
    C = np.swapaxes(F, 1, 2) @ F
    _, cc = R.stress_tangent(model, E, NU, F)
    T = R.full_tangent(cc)
    h = 1e-6

    def S_of(Cm):
        Fm = np.swapaxes(np.linalg.cholesky(Cm), 1, 2)  # F^T F = C
        return R.stress_tangent(model, E, NU, Fm)[0]

    for J, (A, B) in enumerate(R.VOIGT):
        dC = np.zeros((3, 3))
        # eps_J += h  <=>  C_AA += 2h (normal) or C_AB = C_BA += h (engineering shear)
        if A == B:
            dC[A, A] = 2 * h
        else:
            dC[A, B] = dC[B, A] = h
        fd = (S_of(C + dC) - S_of(C - dC)) / (2 * h)
        assert np.allclose(fd, T[:, :, J], rtol=1e-6, atol=1e-6 * E)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_reference_configuration_is_linear_elasticity(model):
    S, cc = R.stress_tangent(model, E, NU, np.eye(3)[None])
    mu, lam = R.lame(E, NU)
    T = R.full_tangent(cc)[0]
    ref = np.zeros((6, 6))
    ref[:3, :3] = lam
    ref[np.arange(6), np.arange(6)] += [2 * mu] * 3 + [mu] * 3
    assert np.max(np.abs(S)) < 1e-12 * E
    assert np.allclose(T, ref, rtol=1e-13, atol=1e-12 * E)


def _torch_loss(p, x, t):
    W1, b1, W2, b2, W3, b3 = [torch.tensor(a) for a in R.unpack(p)]
    for a in (W1, b1, W2, b2, W3, b3):
        a.requires_grad_(True)
    X = torch.zeros((x.shape[0], 16), dtype=torch.float64)
    X[:, :9] = torch.tensor(x)
    g = torch.nn.functional.gelu
    y = g(g(X @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    loss = ((y[:, :21] - torch.tensor(t)) ** 2).sum() / x.shape[0]
    loss.backward()
    grads = [a.grad.numpy() for a in (W1, b1, W2, b2, W3, b3)]
    grads[0][:, 9:] = 0
    grads[4][21:] = 0
    grads[5][21:] = 0
    return loss.item(), y[:, :21].detach().numpy(), R.pack(*grads)


def test_mlp_float64_matches_torch_autograd():
    rng = np.random.default_rng(0)
    p = R.init_params(1)
    x, t = rng.normal(size=(37, 9)), rng.normal(size=(37, 21))
    loss, g = R.loss_grad(p, x, t)
    tl, ty, tg = _torch_loss(p, x, t)
    assert np.allclose(R.forward(p, x)[0], ty, rtol=1e-13, atol=1e-13)
    assert abs(loss - tl) < 1e-12 * tl
    assert np.max(np.abs(g - tg)) < 1e-12 * np.max(np.abs(tg))


def test_mlp_float64_gradient_central_differences():
    rng = np.random.default_rng(3)
    p = R.init_params(2)
    x, t = rng.normal(size=(11, 9)), rng.normal(size=(11, 21))
    _, g = R.loss_grad(p, x, t)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(R.P_W1, R.P_B1).reshape(64, 16)[:, :9].ravel(), 6),
                          rng.integers(R.P_B1, R.P_W2, 3), rng.integers(R.P_W2, R.P_B2, 6),
                          rng.integers(R.P_B2, R.P_W3, 3), rng.choice(np.arange(R.P_W3, R.P_W3 + 21 * 64), 6),
                          rng.integers(R.P_B3, R.P_B3 + 21, 3)])
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_grad(pp, x, t)[0] - R.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_bf16_rounding_and_mode():
    a = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0, 65504.0], np.float32)
    r = R.bf16(a)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12 and r[4] == 0.0 and r[5] == 65536.0
    rng = np.random.default_rng(5)
    p = R.init_params(4)
    x, t = rng.normal(size=(64, 9)), rng.normal(size=(64, 21))
    y64, y16 = R.forward(p, x)[0], R.forward(p, x, bf=True)[0]
    assert np.max(np.abs(y64 - y16)) < 3e-2 * np.max(np.abs(y64))
    g64, g16 = R.loss_grad(p, x, t)[1], R.loss_grad(p, x, t, bf=True)[1]
    assert np.linalg.norm(g64 - g16) < 2e-2 * np.linalg.norm(g64)
