"""CPU tests that pin the oracle (oracle/oracle.cpp) to facts it did not compute itself.

Every oracle function is checked against at least one of: the reference's own
compiled AceGen elements (golden vectors in tests/golden/acegen_tet10.npz and, when
built, oracle/_ref), exact rational integration, closed forms, brute force, finite
differences or invariants -- chosen so that a dropped term, a wrong sign, index or
weight fails.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def mono(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("deg,exact_to,npts", [(1, 1, 1), (2, 3, 5), (3, 3, 5), (4, 5, 15), (5, 5, 15)])
def test_quadrature_exactness(deg, exact_to, npts):
    p, w = O.quadrature(deg)
    assert len(w) == npts
    assert abs(w.sum() - 1 / 6) < 1e-15
    for a in range(exact_to + 1):
        for b in range(exact_to + 1 - a):
            for c in range(exact_to + 1 - a - b):
                q = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
                assert abs(q - mono(a, b, c)) < 1e-15, (deg, a, b, c)
    # and not exact one degree higher (the rule is what the reference says it is)
    d = exact_to + 1
    errs = [abs(np.sum(w * p[:, 0] ** a * p[:, 1] ** (d - a)) - mono(a, d - a, 0)) for a in range(d + 1)]
    assert max(errs) > 1e-10


# --------------------------------------------------------------------- basis
def _phi(fe, x):
    n = 4 if fe == 1 else 10
    out = np.zeros(n)
    v = np.zeros(1)
    for i in range(n):
        O.lib().orc_phi(fe, i, O.P(np.asarray(x, float)), O.P(v))
        out[i] = v[0]
    return out


def _gphi(fe, x):
    n = 4 if fe == 1 else 10
    out = np.zeros((n, 3))
    v = np.zeros(3)
    for i in range(n):
        O.lib().orc_grad_phi(fe, i, O.P(np.asarray(x, float)), O.P(v))
        out[i] = v
    return out


NODES2 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0], [0, .5, 0], [0, 0, .5],
                   [.5, 0, .5], [0, .5, .5]])


@pytest.mark.parametrize("fe", [1, 2])
def test_basis_kronecker_partition_gradients(fe):
    nodes = NODES2[:4] if fe == 1 else NODES2
    for k, x in enumerate(nodes):
        v = _phi(fe, x)
        assert np.allclose(v, np.eye(len(nodes))[k], atol=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.dirichlet(np.ones(4))[1:]
        assert abs(_phi(fe, x).sum() - 1) < 1e-14
        g = _gphi(fe, x)
        assert np.allclose(g.sum(0), 0, atol=1e-13)
        h = 1e-6
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            fd = (_phi(fe, x + e) - _phi(fe, x - e)) / (2 * h)
            assert np.allclose(fd, g[:, d], atol=1e-8)


def test_degree_rules():
    L = O.lib()
    # AssembleFE's "UN Grad = 2" selects the degFunc overload (Helper.cpp:7-23)
    assert L.orc_degree_func(1, 2) == 3 and L.orc_degree_func(2, 2) == 4
    # FE::assemblyLaplace uses (Grad, Grad) (FE_def.hpp:2459): P1 -> 1 (1-pt), P2 -> 2 (5-pt)
    assert L.orc_degree_pair(1, 1, 1, 1, 0) == 1 and L.orc_degree_pair(2, 2, 1, 1, 0) == 2
    # FE::assemblyRHS uses Std (FE_def.hpp:7327)
    assert L.orc_degree_type(1, 0) == 1 and L.orc_degree_type(2, 0) == 2
    assert len(O.quadrature(L.orc_degree_func(2, 2))[1]) == 15


# ---------------------------------------------------------------------- mesh
@pytest.mark.parametrize("M", [1, 2, 3])
def test_mesh_structure(M):
    p1, bc1, e1 = O.mesh(M, 1)
    p2, bc2, e2 = O.mesh(M, 2)
    n1 = M + 1
    # GID = r + s n + t n^2 with coordinates r h (MeshStructured_def.hpp:727-737)
    idx = np.arange(n1 ** 3)
    assert np.array_equal(p1 * M, np.stack([idx % n1, (idx // n1) % n1, idx // n1 ** 2], 1))
    # volume and the Kuhn 6-tet split
    x = p1[e1]
    d = np.linalg.det(np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], 2))
    assert abs(np.abs(d).sum() / 6 - 1.0) < 1e-12 and np.all(np.abs(d) > 0)
    assert np.all(np.abs(np.abs(d) - 1 / M ** 3) < 1e-14)
    # the P2 element's vertices are the P1 element's vertices (fine lattice = 2x coarse)
    assert np.allclose(p2[e2[:, :4]], p1[e1])
    # edge nodes are the midpoints of 4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(1,3) 9:(2,3)
    for k, (a, b) in enumerate([(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]):
        assert np.allclose(p2[e2[:, 4 + k]], 0.5 * (p2[e2[:, a]] + p2[e2[:, b]]))
    # every P2 lattice node is used; boundary flags mark the faces
    assert len(np.unique(e2)) == (2 * M + 1) ** 3
    on_b = np.any((p2 < 1e-12) | (p2 > 1 - 1e-12), axis=1)
    assert np.array_equal(bc2.astype(bool), on_b)


def test_correct_orientation():
    p2, _, e2 = O.mesh(2, 2)
    el = e2.copy()
    nflip = O.correct_orientation(p2, el)
    x = p2[el]
    d = np.linalg.det(np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], 2))
    assert np.all(d > 0) and nflip == 32
    for k, (a, b) in enumerate([(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]):
        assert np.allclose(p2[el[:, 4 + k]], 0.5 * (p2[el[:, a]] + p2[el[:, b]]))


# --------------------------------------------------------------------- graph
def _brute_graph(el, n, dofs=1):
    pairs = set()
    for T in el:
        for i in T:
            for j in T:
                for d1 in range(dofs):
                    for d2 in range(dofs):
                        pairs.add((dofs * i + d1, dofs * j + d2))
    rows = sorted(pairs)
    rp = np.zeros(n * dofs + 1, np.int64)
    for r, _ in rows:
        rp[r + 1] += 1
    return np.cumsum(rp), np.array([c for _, c in rows])


@pytest.mark.parametrize("M", [1, 2])
def test_graph_brute_force(M):
    p1, _, e1 = O.mesh(M, 1)
    rp, ci, _ = O.laplace(1, p1, e1)
    brp, bci = _brute_graph(e1, p1.shape[0])
    assert np.array_equal(rp, brp) and np.array_equal(ci, bci)
    p2, _, e2 = O.mesh(M, 2)
    rp, ci, _ = O.laplace_assfe(2, 3, p2, e2)
    brp, bci = _brute_graph(e2, p2.shape[0], 3)
    assert np.array_equal(rp, brp) and np.array_equal(ci, bci)


@pytest.mark.parametrize("M", [1, 2, 3, 4])
def test_graph_nnz_polynomials(M):
    p1, _, e1 = O.mesh(M, 1)
    assert len(O.laplace(1, p1, e1)[1]) == 15 * M ** 3 + 21 * M ** 2 + 9 * M + 1
    p2, _, e2 = O.mesh(M, 2)
    assert len(O.laplace(2, p2, e2)[1]) == 230 * M ** 3 + 138 * M ** 2 + 24 * M + 1


# ------------------------------------------------------------------- Laplace
def test_laplace_p1_closed_form():
    M = 3
    p1, _, e1 = O.mesh(M, 1)
    p1 = p1 + np.random.default_rng(3).uniform(-0.05, 0.05, p1.shape) / M  # distorted, still valid
    rp, ci, v = O.laplace(1, p1, e1)
    n = p1.shape[0]
    K = np.zeros((n, n))
    for T in e1:
        X = p1[T]
        B = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], 1)
        Bi = np.linalg.inv(B)
        grads = np.vstack([-Bi.sum(0), Bi])  # grad lambda_0..3 (rows)
        K[np.ix_(T, T)] += abs(np.linalg.det(B)) / 6 * grads @ grads.T
    D = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    D[rows, ci] = v
    assert np.max(np.abs(D - K)) < 1e-12 * np.max(np.abs(K))
    assert np.max(np.abs(D.sum(1))) < 1e-12  # constants are in the kernel
    # the AssembleFE path (5-point rule, "UN Grad = 2") equals the classic 1-point path:
    # the reference's own cross-check (feddlib/core/AceFemAssembly/tests/fe_test_laplace.cpp)
    rp2, ci2, v2 = O.laplace_assfe(1, 1, p1, e1)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)
    assert np.max(np.abs(v - v2)) < 1e-13 * np.max(np.abs(v))


@pytest.mark.parametrize("fe", [1, 2])
def test_rhs_closed_form(fe):
    M = 3
    pts, _, el = O.mesh(M, fe)
    b = O.rhs(fe, pts, el, 2.5)
    assert abs(b.sum() - 2.5) < 1e-12  # int f = f |Omega|
    # every element adds |T| f int phi_i; P1: |T|/4, P2: vertex -|T|/20, edge |T|/5
    vol = 1 / (6 * M ** 3)
    ref = np.zeros(pts.shape[0])
    share = [vol / 4] * 4 if fe == 1 else [-vol / 20] * 4 + [vol / 5] * 6
    for T in el:
        ref[T] += 2.5 * np.array(share)
    assert np.max(np.abs(b - ref)) < 1e-14


@pytest.mark.parametrize("fe", [1, 2])
def test_rhs_points_pinned(fe):
    """General-source assemblyRHS: the mapped points are the rule's points (P1 centroid,
    P2 the 5-point orbits), a constant source gives the constant-f vector, the sum is the
    rule applied to f over the cube (partition of unity; exact to degree 1 / 3), and the
    first moment sum_i b_i x_i equals the rule applied to f x (P1 reproduces x, P2 too)."""
    M = 3
    pts, _, el = O.mesh(M, fe)
    xq = O.rhs_quadrature_points(fe, pts, el)
    X = pts[el[:, :4]]
    if fe == 1:
        assert xq.shape[1] == 1 and np.allclose(xq[:, 0], X.mean(axis=1), atol=1e-15)
    else:
        assert xq.shape[1] == 5 and np.allclose(xq[:, 0], X.mean(axis=1), atol=1e-15)
        # the other four points: barycentric (1/2, 1/6, 1/6, 1/6) around each vertex
        orbit = np.sort(np.einsum("ek,ekd->ed", np.array([0.5, 1 / 6, 1 / 6, 1 / 6])[None].repeat(len(el), 0), X))
        cand = [np.sort(xq[:, k]) for k in range(1, 5)]
        assert any(np.allclose(c, orbit, atol=1e-14) for c in cand)
    assert np.max(np.abs(O.rhs_points(fe, pts, el, np.full(xq.shape[:2], 2.5)) - O.rhs(fe, pts, el, 2.5))) < 1e-15
    x, y, z = xq[..., 0], xq[..., 1], xq[..., 2]
    if fe == 1:
        f = 1 + 2 * x + 3 * y - z
        b = O.rhs_points(fe, pts, el, f)
        assert abs(b.sum() - 3.0) < 1e-13  # midpoint rule is exact for linear f
        vol = abs(np.linalg.det(np.stack([X[:, k] - X[:, 0] for k in (1, 2, 3)], 2))) / 6
        mid = X.mean(axis=1)
        ref = np.sum(vol * (1 + 2 * mid[:, 0] + 3 * mid[:, 1] - mid[:, 2]) * mid[:, 0])
        assert abs(b @ pts[:, 0] - ref) < 1e-13
    else:
        f = x ** 2 * y + z ** 3
        b = O.rhs_points(fe, pts, el, f)
        assert abs(b.sum() - 5 / 12) < 1e-13  # int x^2 y + z^3 over the cube, degree 3
        g = x + y
        b = O.rhs_points(fe, pts, el, g)
        assert abs(b @ pts[:, 0] - 7 / 12) < 1e-13  # int (x + y) x, degree 2
    # "Vector": component d of node i at row 3 i + d
    fv = np.stack([x, 2 * y, -z], axis=-1)
    bv = O.rhs_points(fe, pts, el, fv)
    for d, s in enumerate((1.0, 2.0, -1.0)):
        assert np.max(np.abs(bv[d::3] - s * O.rhs_points(fe, pts, el, xq[..., d]))) < 1e-15


# ------------------------------------------------------------- Navier-Stokes
def _p2_poly_basis():
    """P2 basis (FEDDLib order) as callables of barycentrics, for exact integration."""
    return None


def _exact_ns_element(X, sol, nu, rho):
    """Independent NS element (exact integration with the monomial rule over the
    reference tet, no quadrature): returns the 34x34 Newton Jacobian."""
    from itertools import product
    B = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], 1)
    Bi = np.linalg.inv(B)
    adet = abs(np.linalg.det(B))

    # polynomials as dict {(a,b,c): coeff} in reference coords x,y,z
    def P(d):
        return {k: v for k, v in d.items() if v != 0}

    def mul(p, q):
        r = {}
        for (a1, b1, c1), v1 in p.items():
            for (a2, b2, c2), v2 in q.items():
                k = (a1 + a2, b1 + b2, c1 + c2)
                r[k] = r.get(k, 0) + v1 * v2
        return r

    def add(p, q, s=1.0):
        r = dict(p)
        for k, v in q.items():
            r[k] = r.get(k, 0) + s * v
        return r

    def der(p, ax):
        r = {}
        for k, v in p.items():
            if k[ax]:
                kk = list(k)
                kk[ax] -= 1
                r[tuple(kk)] = r.get(tuple(kk), 0) + v * k[ax]
        return r

    def integ(p):
        return sum(v * mono(*k) for k, v in p.items())

    one = {(0, 0, 0): 1.0}
    x, y, z = {(1, 0, 0): 1.0}, {(0, 1, 0): 1.0}, {(0, 0, 1): 1.0}
    L0 = add(add(add(one, x, -1), y, -1), z, -1)
    lam = [L0, x, y, z]
    phi = [mul(L0, add(one, add(add(x, y), z), -2)), mul(x, add(mul(x, {(0, 0, 0): 2.0}), one, -1)),
           mul(y, add(mul(y, {(0, 0, 0): 2.0}), one, -1)), mul(z, add(mul(z, {(0, 0, 0): 2.0}), one, -1)),
           mul(mul(x, L0), {(0, 0, 0): 4.0}), mul(mul(x, y), {(0, 0, 0): 4.0}), mul(mul(y, L0), {(0, 0, 0): 4.0}),
           mul(mul(z, L0), {(0, 0, 0): 4.0}), mul(mul(x, z), {(0, 0, 0): 4.0}), mul(mul(y, z), {(0, 0, 0): 4.0})]
    # physical gradients: d/dx_d = sum_a Bi[a][d] d/dxhat_a
    gphi = [[{} for _ in range(3)] for _ in range(10)]
    for i in range(10):
        for d in range(3):
            g = {}
            for a in range(3):
                g = add(g, der(phi[i], a), Bi[a, d])
            gphi[i][d] = g
    u = sol[:30].reshape(10, 3)
    uq = [{} for _ in range(3)]
    for d in range(3):
        for k in range(10):
            uq[d] = add(uq[d], phi[k], u[k, d])
    du = [[{} for _ in range(3)] for _ in range(3)]  # du[d1][d2] = d u_d1 / d x_d2
    for d1 in range(3):
        for d2 in range(3):
            for k in range(10):
                du[d1][d2] = add(du[d1][d2], gphi[k][d2], u[k, d1])
    J = np.zeros((34, 34))
    for i, j in product(range(10), range(10)):
        a = sum(integ(mul(gphi[i][d], gphi[j][d])) for d in range(3)) * adet
        ug = {}
        for d in range(3):
            ug = add(ug, mul(uq[d], gphi[j][d]))
        nn = integ(mul(ug, phi[i])) * adet
        for d in range(3):
            J[3 * i + d, 3 * j + d] += nu * rho * a + rho * nn
        pp = mul(phi[i], phi[j])
        for d1, d2 in product(range(3), range(3)):
            J[3 * i + d1, 3 * j + d2] += rho * integ(mul(pp, du[d1][d2])) * adet
    for m, j, d in product(range(4), range(10), range(3)):
        bval = integ(mul(lam[m], gphi[j][d])) * adet
        J[30 + m, 3 * j + d] -= bval
        J[3 * j + d, 30 + m] -= bval
    return J


def _rand_element(seed):
    rng = np.random.default_rng(seed)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.3 + rng.normal(0, 0.05, (4, 3))
    X = np.vstack([V] + [(V[a] + V[b]) / 2 for a, b in [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]])
    sol = rng.normal(0, 1, 34)
    return X, sol


@pytest.mark.parametrize("seed", [0, 1])
def test_ns_element_exact_integration(seed):
    X, sol = _rand_element(seed)
    nu, rho = 0.37, 1.3
    jac, fp, rhs = O.ns_element(X, sol, nu, rho, True)
    ref = _exact_ns_element(X, sol, nu, rho)
    assert np.max(np.abs(jac - ref)) < 1e-12 * np.max(np.abs(ref))
    # block structure: A symmetric, B^T = B^T, (1,1) empty
    assert np.allclose(jac[30:, :30], jac[:30, 30:].T, rtol=0, atol=1e-15)
    assert np.all(jac[30:, 30:] == 0)


def test_ns_jacobian_is_derivative_of_residual():
    X, sol = _rand_element(5)
    nu, rho = 0.05, 1.0
    jac, fp, r0 = O.ns_element(X, sol, nu, rho, True)
    h = 1e-5
    fd = np.zeros((34, 34))
    for j in range(34):
        e = np.zeros(34)
        e[j] = h
        rp = O.ns_element(X, sol + e, nu, rho, True)[2]
        rm = O.ns_element(X, sol - e, nu, rho, True)[2]
        fd[:, j] = (rp - rm) / (2 * h)
    # residual = (A + N(u) - B - B^T)[u;p] is quadratic in u: its derivative is A + N + W - B - B^T
    assert np.max(np.abs(fd - jac)) < 1e-7 * np.max(np.abs(jac))
    # the fixed-point matrix is the Jacobian without W, and W = 0 for constant velocity
    s2 = sol.copy()
    s2[:30] = np.tile([0.3, -0.2, 0.7], 10)
    j2, f2, _ = O.ns_element(X, s2, nu, rho, True)
    assert np.max(np.abs(j2 - f2)) < 1e-14 * np.max(np.abs(j2))


def test_ns_residual_coeff_and_assembly():
    # global residual equals the fixed-point operator applied to [u;p] (coeff = 1)
    M = 2
    p2, _, e2 = O.mesh(M, 2)
    p1, _, e1 = O.mesh(M, 1)
    rng = np.random.default_rng(4)
    u = rng.normal(size=3 * p2.shape[0])
    p = rng.normal(size=p1.shape[0])
    out = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, 0.1, 1.0, True, 2, True)
    fpm = O.ns_assembly(p2, e2, e1, p1.shape[0], u, p, 0.1, 1.0, False, 1, False)
    import scipy.sparse as sp
    A = sp.csr_matrix((fpm["A"][2], fpm["A"][1], fpm["A"][0]))
    BT = sp.csr_matrix((fpm["BT"][2], fpm["BT"][1], fpm["BT"][0]), shape=(3 * p2.shape[0], p1.shape[0]))
    B = sp.csr_matrix((fpm["B"][2], fpm["B"][1], fpm["B"][0]), shape=(p1.shape[0], 3 * p2.shape[0]))
    ru = A @ u + BT @ p
    rp = B @ u
    assert np.max(np.abs(out["res_u"] - ru)) < 1e-12 * np.max(np.abs(ru))
    assert np.max(np.abs(out["res_p"] - rp)) < 1e-12 * np.max(np.abs(rp))
    # B is the transpose of B^T
    assert abs(B - BT.T).max() < 1e-15


# --------------------------------------------------------------- Neo-Hookean
def test_nh_matches_reference_acegen_golden():
    g = np.load(os.path.join(HERE, "golden", "acegen_tet10.npz"))
    for k in range(len(g["model"])):
        s, p = O.nh_skr(int(g["model"][k]), float(g["E"][k]), float(g["nu"][k]), g["xl"][k], g["ul"][k])
        sr = g["s"][k]
        pr = g["p"][k]
        assert np.max(np.abs(s - sr)) < 1e-12 * np.max(np.abs(sr)), k
        assert np.max(np.abs(p - pr)) < 1e-12 * np.max(np.abs(pr)), k


@pytest.mark.parametrize("amp,tol", [(0.01, 1e-12), (0.05, 1e-9)])
def test_nh_matches_reference_acegen_live(amp, tol):
    """Against the compiled reference.  Moderate strains agree to 1e-12; at large strains
    (displacement 0.05 on 0.3-sized elements, tangent entries ~1e6 x E) the two
    expression orders differ by conditioning-amplified rounding, hence 1e-9."""
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    rng = np.random.default_rng(21)
    for model in range(3):
        for _ in range(3):
            X, _ = _rand_element(int(rng.integers(1000)))
            U = rng.normal(0, amp, (10, 3))
            s, p = O.nh_skr(model, 700.0, 0.35, X, U)
            sr, pr = O.ref_skr(model, 700.0, 0.35, X, U)
            assert np.max(np.abs(s - sr)) < tol * np.max(np.abs(sr))
            assert np.max(np.abs(p - pr)) < tol * np.max(np.abs(pr))


@pytest.mark.parametrize("model", [0, 1, 2])
def test_nh_tangent_is_derivative_of_residual(model):
    X, _ = _rand_element(8)
    U = np.random.default_rng(9).normal(0, 0.05, (10, 3))
    s, p = O.nh_skr(model, 1000.0, 0.3, X, U)
    h = 1e-7
    fd = np.zeros((30, 30))
    for j in range(30):
        e = np.zeros(30)
        e[j] = h
        fd[:, j] = (O.nh_skr(model, 1000.0, 0.3, X, U.ravel() + e)[1] -
                    O.nh_skr(model, 1000.0, 0.3, X, U.ravel() - e)[1]) / (2 * h)
    assert np.max(np.abs(fd - s)) < 1e-6 * np.max(np.abs(s))
    assert np.max(np.abs(s - s.T)) < 1e-12 * np.max(np.abs(s))  # hyperelastic: symmetric


def test_nh_small_strain_limit_is_linear_elasticity():
    """At u = 0 every law's tangent is the small-strain operator int 2 mu eps:eps + lam div div,
    computed here independently (exact: the integrand is constant on an affine tet)."""
    X, _ = _rand_element(13)
    E, nu = 1000.0, 0.3
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    V = X[:4]
    B = np.stack([V[1] - V[0], V[2] - V[0], V[3] - V[0]], 1)
    Bi = np.linalg.inv(B)
    adet = abs(np.linalg.det(B))
    # P2 gradients are linear: integrate exactly with the degree-2 capable 5-point rule
    pq, wq = O.quadrature(2)
    K = np.zeros((30, 30))
    for q in range(len(wq)):
        g = _gphi(2, pq[q]) @ Bi  # physical gradients (10,3)
        for i in range(10):
            for j in range(10):
                for a in range(3):
                    for b in range(3):
                        K[3 * i + a, 3 * j + b] += wq[q] * adet * (
                            mu * (g[i] @ g[j]) * (a == b) + mu * g[i][b] * g[j][a] + lam * g[i][a] * g[j][b])
    for model in range(3):
        s, p = O.nh_skr(model, E, nu, X, np.zeros(30))
        # skr* carries the AceGen signed determinant: s = -sign(det_FEDD) K
        assert np.max(np.abs(-s - K)) < 1e-11 * np.max(np.abs(K)), model
        assert np.max(np.abs(p)) < 1e-12


def test_nh_frame_invariance():
    X, _ = _rand_element(17)
    U = np.random.default_rng(18).normal(0, 0.05, (10, 3))
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    # superposed rotation: x -> Q x, so u' = Q (X + u) - X
    U2 = (Q @ (X + U).T).T - X
    for model in range(3):
        s1, p1 = O.nh_skr(model, 1000.0, 0.3, X, U)
        s2, p2 = O.nh_skr(model, 1000.0, 0.3, X, U2)
        # the internal force rotates with Q
        f1 = p1.reshape(10, 3)
        f2 = p2.reshape(10, 3)
        assert np.max(np.abs(f2 - (Q @ f1.T).T)) < 1e-10 * np.max(np.abs(f1))


def test_elasticity_assembly_scatter():
    """FE::assemblyNonLinearElasticity == element-by-element scatter of -s and p."""
    M = 1
    pts, _, el = O.mesh(M, 2)
    rng = np.random.default_rng(2)
    u = rng.normal(0, 0.01, 3 * pts.shape[0])
    (rp, ci, v), r = O.elasticity(1, 1000.0, 0.3, pts, el, u)
    n = 3 * pts.shape[0]
    K = np.zeros((n, n))
    R = np.zeros(n)
    for T in el:
        dofs = (3 * T[:, None] + np.arange(3)).ravel()
        s, p = O.nh_skr(1, 1000.0, 0.3, pts[T], u.reshape(-1, 3)[T])
        K[np.ix_(dofs, dofs)] += -s
        R[dofs] += p
    D = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    D[rows, ci] = v
    assert np.max(np.abs(D - K)) < 1e-13 * np.max(np.abs(K))
    assert np.max(np.abs(r - R)) < 1e-13 * np.max(np.abs(R))
