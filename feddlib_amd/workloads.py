"""Seeded synthetic inputs for the assembly workloads (no datasets are used).

Poisson P1       unit cube M=20, f = 1
Navier-Stokes    unit cube M=40 (P2 velocity / P1 pressure), nu = 0.01, rho = 1,
                 u_d(x) = sum_{m<8} a_{m,d} sin(2 pi k_m . x + phi_m), k_m in {1,2,3}^3,
                 phi_m ~ U[0, 2 pi), a_{m,d} ~ U{-1, +1}, p = 0
Elasticity P2    unit cube M=32, interior P1 vertices jittered U(+-0.2 h) per
                 coordinate, P2 edge nodes at the midpoints, displacement
                 u = 0.05 N(0,1) smoothed by one Jacobi-Laplacian pass on the P2 graph
"""
import numpy as np

from . import native


def ns_velocity(pts, seed=0):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 4, size=(8, 3)).astype(np.float64)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=8)
    a = rng.choice([-1.0, 1.0], size=(8, 3))
    arg = 2.0 * np.pi * (pts @ k.T) + phi[None, :]  # (n, 8)
    return np.sin(arg) @ a  # (n, 3)


def ns_problem(M, seed=0):
    """P2 velocity mesh, P1 pressure mesh, velocity (n2,3) and pressure (n1,) iterates."""
    p2, _, e2 = native.mesh_structured(M, 2)
    p1, _, e1 = native.mesh_structured(M, 1)
    u = ns_velocity(p2, seed)
    p = np.zeros(p1.shape[0])
    return p2, e2, p1, e1, u, p


def jittered_p2_mesh(M, seed=0, amp=0.2, orient=True):
    """P2 cube mesh with interior vertices jittered by U(+-amp*h) and exact midpoints."""
    p2, bc2, e2 = native.mesh_structured(M, 2)
    P = 2 * M + 1
    h = 1.0 / M
    rng = np.random.default_rng(seed)
    idx = np.arange(P ** 3)
    r, s, t = idx % P, (idx // P) % P, idx // (P * P)
    vert = (r % 2 == 0) & (s % 2 == 0) & (t % 2 == 0)
    interior = (r > 0) & (r < P - 1) & (s > 0) & (s < P - 1) & (t > 0) & (t < P - 1)
    sel = vert & interior
    pts = p2.copy()
    pts[sel] += rng.uniform(-amp * h, amp * h, size=(sel.sum(), 3))
    # re-place edge nodes at the midpoints of their element edges
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    for m, (a, b) in enumerate(edges):
        pts[e2[:, 4 + m]] = 0.5 * (pts[e2[:, a]] + pts[e2[:, b]])
    e2 = np.ascontiguousarray(e2)
    if orient:
        native.correct_orientation(pts, e2)
    return pts, bc2, e2


def smoothed_displacement(pts, elems, seed=1, scale=0.05):
    """u = scale*N(0,1) per dof, then one Jacobi-Laplacian pass over the P2 node graph."""
    rng = np.random.default_rng(seed)
    n = pts.shape[0]
    u = scale * rng.standard_normal((n, 3))
    # node graph from the elements (pairs of nodes sharing an element)
    k = elems.shape[1]
    rows = np.repeat(elems, k, axis=1).ravel()
    cols = np.tile(elems, (1, k)).ravel()
    key = np.unique(rows.astype(np.int64) * n + cols)
    r, c = key // n, key % n
    off = r != c
    r, c = r[off], c[off]
    acc = u.copy()
    np.add.at(acc, r, u[c])
    deg = np.bincount(r, minlength=n).astype(np.float64)
    return acc / (1.0 + deg)[:, None]
