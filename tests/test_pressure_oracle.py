"""Pins of the oracle's follower pressure load (AceGen PressureTriangle3D6 restated:
orc_pressure_triangle, orc_assembly_surface_pressure) and of the surface list of
MeshStructured::buildSurfaces(3, "P2") (orc_structured_surfaces, and the library's own
fedd_structured_surfaces, host code).

The AceGenInterface library that holds PressureTriangle3D6 is not part of the reference
(FindTPLAceGENInterface.cmake names only its headers and libraries), so no vector of the
element itself exists: parity with it is unpinned, and its integration rule (code 35) is
unknown.  The restatement is pinned by what the mathematics fixes:
  - a flat straight-sided triangle: vertex loads 0, edge loads -p A n / 3 (the consistent
    load of a quadratic triangle);
  - a closed surface (every boundary face of the cube) under uniform pressure, deformed
    by any continuous displacement: total force and total moment vanish (closed-surface
    integrals of n and x x n), and the assembled Kext is symmetric (a uniform pressure on a
    closed surface is conservative);
  - Kext = -df/du by central differences; covariance under rotation and translation;
    linearity in p.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W


def _rot(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.diag(r))


def test_flat_triangle_closed_form():
    rng = np.random.default_rng(0)
    v = rng.normal(size=(3, 3))
    X = np.vstack([v, (v[0] + v[1]) / 2, (v[1] + v[2]) / 2, (v[0] + v[2]) / 2])
    an = 0.5 * np.cross(v[1] - v[0], v[2] - v[0])  # A n
    f, _ = O.pressure_triangle(2.5, X, np.zeros((6, 3)))
    assert np.abs(f[:3]).max() <= 1e-15 * np.abs(an).max()
    assert np.allclose(f[3:], -2.5 * an / 3, rtol=0, atol=1e-14 * np.abs(an).max())


def test_tangent_is_minus_the_load_derivative():
    rng = np.random.default_rng(1)
    v = rng.normal(size=(3, 3))
    X = np.vstack([v, (v[0] + v[1]) / 2, (v[1] + v[2]) / 2, (v[0] + v[2]) / 2]) + 0.05 * rng.normal(size=(6, 3))
    U = 0.1 * rng.normal(size=(6, 3))
    _, K = O.pressure_triangle(1.7, X, U)
    h = 1e-6
    Kfd = np.zeros((18, 18))
    for k in range(18):
        up, um = U.ravel().copy(), U.ravel().copy()
        up[k] += h
        um[k] -= h
        Kfd[:, k] = -(O.pressure_triangle(1.7, X, up.reshape(6, 3))[0] -
                      O.pressure_triangle(1.7, X, um.reshape(6, 3))[0]).ravel() / (2 * h)
    assert np.abs(Kfd - K).max() <= 1e-8 * np.abs(K).max()


def test_covariance_and_linearity():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(6, 3))
    U = 0.1 * rng.normal(size=(6, 3))
    f, K = O.pressure_triangle(1.0, X, U)
    R, t = _rot(3), rng.normal(size=3)
    fr, Kr = O.pressure_triangle(1.0, X @ R.T + t, U @ R.T)
    assert np.abs(fr - f @ R.T).max() <= 1e-13 * np.abs(f).max()
    Rb = np.kron(np.eye(6), R)
    assert np.abs(Kr - Rb @ K @ Rb.T).max() <= 1e-13 * np.abs(K).max()
    f3, K3 = O.pressure_triangle(-3.0, X, U)
    assert np.abs(f3 + 3.0 * f).max() <= 1e-14 * np.abs(f3).max()
    assert np.abs(K3 + 3.0 * K).max() <= 1e-14 * np.abs(K3).max()


@pytest.mark.parametrize("M", [1, 3])
def test_closed_surface_invariants(M):
    pts, _, el = W.jittered_p2_mesh(M, 0)  # interior vertices jittered, sides plane
    faces, flags, _ = O.structured_surfaces(pts, el)
    x = pts
    u = 0.05 * np.sin(3 * x[:, [1, 2, 0]]) + 0.03 * np.cos(2 * x)  # any continuous field
    f, Kx = O.surface_pressure(pts, faces, flags, u, np.full(7, 1.3), tangent=True)
    fn = f.reshape(-1, 3)
    scale = np.abs(fn).max()
    assert np.abs(fn.sum(0)).max() <= 1e-13 * scale * len(fn)  # int n da = 0
    xc = x + u
    assert np.abs(np.cross(xc, fn).sum(0)).max() <= 1e-13 * scale * len(fn)  # int x x n da = 0
    rp, ci, v = Kx
    K = sp.csr_matrix((v, ci, rp), shape=(3 * len(pts),) * 2)
    assert abs(K - K.T).max() <= 1e-13 * abs(K).max()


def test_open_surface_tangent_is_not_symmetric():
    """The same load on one side only is not conservative: Kext - Kext^T is of the order of
    Kext (which is why the Newton solve with a follower load needs BiCGStab, not CG)."""
    pts, _, el = O.mesh(2, 2)
    faces, flags, _ = O.structured_surfaces(pts, el)
    p = np.zeros(7)
    p[4] = 1.0
    _, (rp, ci, v) = O.surface_pressure(pts, faces, flags, None, p, tangent=True)
    K = sp.csr_matrix((v, ci, rp), shape=(3 * len(pts),) * 2)
    assert abs(K - K.T).max() > 0.1 * abs(K).max()


@pytest.mark.parametrize("M", [1, 2, 3])
def test_structured_surfaces(M):
    pts, _, el = O.mesh(M, 2)
    faces, flags, owner = O.structured_surfaces(pts, el)
    assert len(faces) == 12 * M * M
    assert np.array_equal(np.bincount(flags, minlength=7)[1:], np.full(6, 2 * M * M))
    X = pts[faces]
    nrm = np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0])
    mid = X[:, :3].mean(1)
    assert np.all(np.einsum("fi,fi->f", nrm, mid - 0.5) > 0)  # outward
    side = {1: (0, 0.0), 2: (1, 0.0), 3: (2, 0.0), 4: (2, 1.0), 5: (1, 1.0), 6: (0, 1.0)}
    for k, (ax, val) in side.items():
        assert np.all(np.abs(X[flags == k][:, :, ax] - val) < 1e-12)
    for a, b, e in ((0, 1, 3), (1, 2, 4), (0, 2, 5)):  # FEDDLib P2 triangle edge order
        assert np.allclose(X[:, e], (X[:, a] + X[:, b]) / 2, atol=1e-15)
    assert all(set(faces[s]) <= set(el[owner[s]]) for s in range(len(faces)))
    # the library's own builder (host code) lists the same faces, bit for bit
    f2, fl2 = N.structured_surfaces(pts, el)
    assert np.array_equal(f2, faces) and np.array_equal(fl2, flags)
