"""CPU reference of the learned tangent surrogate and of the hyperelastic laws it learns.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checks, as the checker.
The product (libfeddgpu.so) never imports or calls it.

Contents
  * the three strain energies of the reference's AceGen elements
      skr1  NeoHookQuadraticTets.f   W = mu/2 (I1-3) - mu lnJ + lam/4 (J^2 - 1 - 2 lnJ)
      skr2  NeoHookQuadraticTets2.f  W = mu/2 (I1-3) - mu lnJ + lam/2 (lnJ)^2
      skr3  NeoHookQuadraticTets3.f  W = mu/2 (I1 I3^(-1/3) - 3) + K/2 (J-1)^2,  K = E / (3 (1-2 nu))
    (FEDDLib's feddlib/core/AceFemAssembly/AceInterface/NeoHookQuadraticTets*.f,
    the Fortran sources; internal model index 0/1/2 as in the library),
    with the second Piola-Kirchhoff stress S = 2 dW/dC and tangent CC = dS/dE in
    closed form (Voigt order 11,22,33,12,13,23, engineering shear), pinned in
    tests/test_mlp_oracle.py by finite differences (S against W, CC against S);
  * the 9 -> 64 -> 64 -> 21 GELU MLP with the parameter layout of the C ABI
    (W1[64][16] b1[64] W2[64][64] b2[64] W3[32][64] b3[32]), forward and
    backward in float64, optionally rounding every MFMA operand to bfloat16 as
    the device kernels do (weights, inputs, hidden activations, back-propagated
    errors; accumulation and biases exact).  The float64 mode is pinned by finite
    differences and by torch autograd; the bfloat16 mode is the same code with
    the roundings inserted.
"""
import numpy as np
from scipy.special import erf

VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
P_W1, P_B1 = 0, 64 * 16
P_W2 = P_B1 + 64
P_B2 = P_W2 + 64 * 64
P_W3 = P_B2 + 64
P_B3 = P_W3 + 32 * 64
P_TOTAL = P_B3 + 32


def lame(E, nu):
    mu = E / (2 * (1 + nu))
    return mu, 2 * mu * nu / (1 - 2 * nu)


def energy(model, E, nu, F):
    """Strain energy W(F) of law `model` (0 skr1, 1 skr2, 2 skr3); F (..., 3, 3)."""
    mu, lam = lame(E, nu)
    C = np.swapaxes(F, -1, -2) @ F
    I1 = np.trace(C, axis1=-2, axis2=-1)
    I3 = np.linalg.det(C)
    J = np.sqrt(I3)
    if model == 0:
        return mu / 2 * (I1 - 3) - mu * np.log(J) + lam / 4 * (J ** 2 - 1 - 2 * np.log(J))
    if model == 1:
        return mu / 2 * (I1 - 3) - mu * np.log(J) + lam / 2 * np.log(J) ** 2
    K = E / (3 * (1 - 2 * nu))
    return mu / 2 * (I1 * I3 ** (-1.0 / 3) - 3) + K / 2 * (J - 1) ** 2


def stress_tangent(model, E, nu, F):
    """S (..., 6) and CC (..., 21, upper triangle row-major of the 6x6 Voigt tangent).

    Closed forms: with Ci = C^-1, S = a_id I + a_ci Ci and
    CC_ABCD = c_cc Ci_AB Ci_CD + c_ii (Ci_AC Ci_BD + Ci_AD Ci_BC)/2 + c_ic (d_AB Ci_CD + Ci_AB d_CD).
    """
    F = np.asarray(F, np.float64)
    mu, lam = lame(E, nu)
    C = np.swapaxes(F, -1, -2) @ F
    Ci = np.linalg.inv(C)
    I1 = np.trace(C, axis1=-2, axis2=-1)
    I3 = np.linalg.det(C)
    J = np.sqrt(I3)
    z = np.zeros_like(J)
    if model == 0:
        a_id, a_ci = mu + z, lam / 2 * I3 - mu - lam / 2
        c_cc, c_ii, c_ic = lam * I3, 2 * mu + lam - lam * I3, z
    elif model == 1:
        lnJ = np.log(J)
        a_id, a_ci = mu + z, lam * lnJ - mu
        c_cc, c_ii, c_ic = lam + z, 2 * (mu - lam * lnJ), z
    else:
        K = E / (3 * (1 - 2 * nu))
        f = I3 ** (-1.0 / 3)
        a_id, a_ci = mu * f, -mu / 3 * I1 * f + K * (J - 1) * J
        c_cc = 2.0 / 9 * mu * f * I1 + K * (2 * J - 1) * J
        c_ii = 2.0 / 3 * mu * f * I1 - 2 * K * (J - 1) * J
        c_ic = -2.0 / 3 * mu * f
    eye = np.eye(3)
    S = np.stack([a_id * eye[A, B] + a_ci * Ci[..., A, B] for A, B in VOIGT], -1)
    cc = []
    for i, (A, B) in enumerate(VOIGT):
        for (Cc, D) in VOIGT[i:]:
            cc.append(c_cc * Ci[..., A, B] * Ci[..., Cc, D]
                      + c_ii * 0.5 * (Ci[..., A, Cc] * Ci[..., B, D] + Ci[..., A, D] * Ci[..., B, Cc])
                      + c_ic * (eye[A, B] * Ci[..., Cc, D] + eye[Cc, D] * Ci[..., A, B]))
    return S, np.stack(cc, -1)


def full_tangent(cc21):
    """6x6 symmetric matrix from the 21 upper-triangle entries."""
    T = np.zeros(cc21.shape[:-1] + (6, 6))
    k = 0
    for i in range(6):
        for j in range(i, 6):
            T[..., i, j] = T[..., j, i] = cc21[..., k]
            k += 1
    return T


# ------------------------------------------------------------------------ MLP
def bf16(a):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[P_W1:P_B1].reshape(64, 16), p[P_B1:P_W2], p[P_W2:P_B2].reshape(64, 64), p[P_B2:P_W3],
            p[P_W3:P_B3].reshape(32, 64), p[P_B3:P_TOTAL])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def gelu(z):
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


def gelu_d(z):
    return 0.5 * (1 + erf(z / np.sqrt(2))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, np.float64)


def forward(p, x, bf=False):
    """y (n, 21) for inputs x (n, 9); with bf=True every MFMA operand is rounded to bfloat16."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.zeros((x.shape[0], 16))
    X[:, :9] = x
    z1 = _r(X, bf) @ _r(W1, bf).T + b1
    h1 = _r(gelu(z1), bf)
    z2 = h1 @ _r(W2, bf).T + b2
    h2 = _r(gelu(z2), bf)
    y = h2 @ _r(W3, bf).T + b3
    return y[:, :21], (X, z1, h1, z2, h2, y)


def loss_grad(p, x, t, bf=False):
    """Loss sum_o (y - t)^2 / n and its gradient w.r.t. the parameter vector."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = x.shape[0]
    y21, (X, z1, h1, z2, h2, y) = forward(p, x, bf)
    e = y21 - t
    loss = np.sum(e * e) / n
    d3 = np.zeros((n, 32))
    d3[:, :21] = 2 * e / n
    gW3 = _r(d3, bf).T @ h2
    gb3 = d3.sum(0)
    d2 = (_r(d3, bf) @ _r(W3, bf)) * gelu_d(z2)
    gW2 = _r(d2, bf).T @ h1
    gb2 = d2.sum(0)
    d1 = (_r(d2, bf) @ _r(W2, bf)) * gelu_d(z1)
    gW1 = _r(d1, bf).T @ _r(X, bf)
    gb1 = d1.sum(0)
    gW1[:, 9:] = 0
    gW3[21:] = 0
    gb3[21:] = 0
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def init_params(seed):
    """Random parameters in the layout of the C ABI (padding zero), for tests."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((64, 16))
    W1[:, :9] = rng.normal(0, np.sqrt(2 / 73), (64, 9))
    W2 = rng.normal(0, np.sqrt(2 / 128), (64, 64))
    W3 = np.zeros((32, 64))
    W3[:21] = rng.normal(0, np.sqrt(2 / 85), (21, 64))
    b1, b2 = rng.normal(0, 0.1, 64), rng.normal(0, 0.1, 64)
    b3 = np.zeros(32)
    b3[:21] = rng.normal(0, 0.1, 21)
    return pack(W1, b1, W2, b2, W3, b3)


def normalise(model, E, nu, F, s_in=0.1):
    """Network input x and target y of deformation gradients F (n, 3, 3)."""
    mu, lam = lame(E, nu)
    _, cc0 = stress_tangent(model, E, nu, np.eye(3)[None])
    _, cc = stress_tangent(model, E, nu, F)
    x = (F - np.eye(3)).reshape(-1, 9) / s_in
    return x, (cc - cc0) / (lam + 2 * mu)
