"""CPU reference of FE::assemblyLinElasXDim (3D), the linear-elasticity operator of
FEDDLib's LinElas problem.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
(libfeddgpu.so) never imports or calls it.

Follows feddlib/core/FE/FE_def.hpp:4701-5004 (dim 3 branch :4858-5004):
  * rule from Helper::determineDegree(dim, FEType, FEType, Grad, Grad) (FE_def.hpp:4719,
    Helper.cpp:7-136, degree 0 raised to 1): P1 -> the 1-point rule, P2 -> degree 2,
    raised to 3 (Helper.cpp:1302-1305) -> the 5-point rule; the oracle's quadrature
    tables are the same;
  * B[:, j] = x_{j+1} - x_0, Binv, |det B| (Helper::buildTransformation, :4873-4875);
  * dPhiTrans = grad phi_ref * Binv (applyBTinv, FE_def.hpp:92-105), reference gradients
    as Helper::gradPhi (Helper.cpp:442-594);
  * for every local pair (i, j) and directions (a, b): epsilonTensor of phi_i e_a and
    phi_j e_b, v_ab = sum_q w_q (2 mu eps_i : eps_j + lambda tr eps_i tr eps_j) (:4895-4950),
    times |det B|, inserted at row 3 gid(i) + a, column 3 gid(j) + b (:4953-4990).

element_matrix() forms the epsilon tensors explicitly, as the reference does.
linelas_xdim() assembles the whole mesh with the same expression in closed form,
  2 mu eps(phi_i e_a) : eps(phi_j e_b) = mu (delta_ab g_i . g_j + g_i[b] g_j[a]),
  tr eps(phi_i e_a) = g_i[a],
vectorised over elements; tests check the two against each other.  Duplicates are
summed (Tpetra fillComplete).  Returns a scipy CSR matrix.
"""
import numpy as np
import scipy.sparse as sp

import pyoracle as O


def grad_ref(fe, q):
    """Reference gradients (npe, 3) at point q."""
    x, y, z = q
    if fe == 1:
        return np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    L = 1.0 - x - y - z
    return np.array([
        [-(4 * L - 1)] * 3,
        [4 * x - 1, 0.0, 0.0],
        [0.0, 4 * y - 1, 0.0],
        [0.0, 0.0, 4 * z - 1],
        [4 * (L - x), -4 * x, -4 * x],
        [4 * y, 4 * x, 0.0],
        [-4 * y, 4 * (L - y), -4 * y],
        [-4 * z, -4 * z, 4 * (L - z)],
        [4 * z, 0.0, 4 * x],
        [0.0, 4 * z, 4 * y],
    ])


def rule(fe):
    return O.quadrature(1 if fe == 1 else 2)


def physical_gradients(fe, xs):
    """xs: (npe, 3) element nodes -> (nq, npe, 3) gradients, |det B|, weights."""
    qp, qw = rule(fe)
    B = np.stack([xs[1] - xs[0], xs[2] - xs[0], xs[3] - xs[0]], axis=1)
    Binv = np.linalg.inv(B)
    return np.array([grad_ref(fe, q) @ Binv for q in qp]), abs(np.linalg.det(B)), qw


def epsilon(g, a):
    """epsilonTensor of phi e_a for the physical gradient g (3,): sym(e_a outer g)."""
    m = np.zeros((3, 3))
    m[a, :] += 0.5 * g
    m[:, a] += 0.5 * g
    return m


def element_matrix(fe, xs, lam, mu):
    """(3 npe, 3 npe) element matrix, row 3 i + a, column 3 j + b, literal epsilon tensors."""
    G, adet, qw = physical_gradients(fe, xs)
    npe = G.shape[1]
    K = np.zeros((3 * npe, 3 * npe))
    for i in range(npe):
        for j in range(npe):
            for a in range(3):
                for b in range(3):
                    v = 0.0
                    for g, w in zip(G, qw):
                        ei, ej = epsilon(g[i], a), epsilon(g[j], b)
                        v += w * (2 * mu * np.sum(ei * ej) + lam * np.trace(ei) * np.trace(ej))
                    K[3 * i + a, 3 * j + b] = v * adet
    return K


def linelas_xdim(fe, pts, el, lam, mu, element=None):
    """element="round32": every element matrix value rounded to FP32 before the FP64 sum;
    element="scale": each element value replaced by the element matrix's largest |value|
    (test switches, as pyoracle.element_round32 / pyoracle.element_scale)."""
    pts = np.asarray(pts, float)
    el = np.asarray(el, np.int64)
    E, npe = el.shape
    qp, qw = rule(fe)
    D = np.array([grad_ref(fe, q) for q in qp])  # (nq, npe, 3)
    X = pts[el]  # (E, npe, 3)
    B = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)  # (E, 3, 3)
    Binv = np.linalg.inv(B)
    adet = np.abs(np.linalg.det(B))
    G = np.einsum("qnk,ekd->eqnd", D, Binv)  # (E, nq, npe, 3)
    gg = np.einsum("eqid,eqjd->eqij", G, G)
    I3 = np.eye(3)
    K = (mu * np.einsum("eqij,ab->eqiajb", gg, I3)
         + mu * np.einsum("eqib,eqja->eqiajb", G, G)
         + lam * np.einsum("eqia,eqjb->eqiajb", G, G))
    K = np.einsum("q,eqiajb->eiajb", qw, K) * adet[:, None, None, None, None]
    if element == "round32":
        K = K.astype(np.float32).astype(np.float64)
    elif element == "scale":  # K[e, i, a, j, b]: the largest |value| of each element
        K = np.broadcast_to(np.max(np.abs(K), axis=(1, 2, 3, 4), keepdims=True), K.shape).copy()
    dof = 3 * el[:, :, None] + np.arange(3)[None, None, :]  # (E, npe, 3)
    rows = np.broadcast_to(dof[:, :, :, None, None], K.shape).ravel()
    cols = np.broadcast_to(dof[:, None, None, :, :], K.shape).ravel()
    n = 3 * len(pts)
    return sp.csr_matrix((K.ravel(), (rows, cols)), shape=(n, n))
