"""Generate fedd_tensors.h: exact reference-element integrals for the GPU kernels.

The element integrals of FEDDLib's P2-P1 Navier-Stokes element
(AssembleFENavierStokes_def.hpp:150-461) and its P1/P2 Laplace elements are
polynomials of degree <= 5 on affine tets, which the reference's own
quadrature rules (Helper.cpp:1300-1517) integrate exactly.  The follower pressure
load on a 6-node surface triangle is a polynomial of degree 4 in the reference
coordinates (N_a times the cross product of two linear tangent vectors).  The kernels use the
equivalent exact reference tensors instead of quadrature loops; this script
computes them in exact rational arithmetic (monomial rule
int x^a y^b z^c = a! b! c! / (a+b+c+3)! over the unit tet) and writes them as
correctly rounded double literals.

Only two rows are needed per tensor: the kernels relabel every element so that
the row's node sits at local position 0 (vertex class) or 4 (edge class).

Run:  python gen_tensors.py > fedd_tensors.h
"""
from fractions import Fraction as Fr
from math import factorial
import itertools


def mono_int(a, b, c):
    return Fr(factorial(a) * factorial(b) * factorial(c), factorial(a + b + c + 3))


class Poly(dict):
    def __add__(self, o):
        r = Poly(self)
        for k, v in o.items():
            r[k] = r.get(k, 0) + v
        return r

    def __mul__(self, o):
        if not isinstance(o, Poly):
            return Poly({k: v * o for k, v in self.items()})
        r = Poly()
        for k1, v1 in self.items():
            for k2, v2 in o.items():
                k = (k1[0] + k2[0], k1[1] + k2[1], k1[2] + k2[2])
                r[k] = r.get(k, 0) + v1 * v2
        return r

    __rmul__ = __mul__

    def d(self, ax):
        r = Poly()
        for k, v in self.items():
            if k[ax] > 0:
                kk = list(k)
                kk[ax] -= 1
                r[tuple(kk)] = r.get(tuple(kk), 0) + v * k[ax]
        return r

    def integral(self):
        return sum((v * mono_int(*k) for k, v in self.items()), Fr(0))

    def integral_tri(self):
        """Over the unit triangle (z-free terms only): int x^a y^b = a! b! / (a+b+2)!."""
        assert all(k[2] == 0 for k in self)
        return sum((v * Fr(factorial(k[0]) * factorial(k[1]), factorial(k[0] + k[1] + 2)) for k, v in self.items()),
                   Fr(0))

    def at(self, p):
        return sum((v * p[0] ** k[0] * p[1] ** k[1] * p[2] ** k[2] for k, v in self.items()), Fr(0))


ONE = Poly({(0, 0, 0): Fr(1)})
X = Poly({(1, 0, 0): Fr(1)})
Y = Poly({(0, 1, 0): Fr(1)})
Z = Poly({(0, 0, 1): Fr(1)})
L0 = ONE + X * -1 + Y * -1 + Z * -1
P1 = [L0, X, Y, Z]  # Helper.cpp:660-675
# P2 basis in FEDDLib order (Helper.cpp:677-710)
P2 = [L0 * (ONE + X * -2 + Y * -2 + Z * -2), X * (X * 2 + ONE * -1), Y * (Y * 2 + ONE * -1),
      Z * (Z * 2 + ONE * -1), X * L0 * 4, X * Y * 4, Y * L0 * 4, Z * L0 * 4, X * Z * 4, Y * Z * 4]
VERT = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
# 6-node triangle in FEDDLib's 2D P2 order (Helper::phi, Helper.cpp:662-686): vertices (0,0), (1,0),
# (0,1), then the edges 01, 12, 02
T0 = ONE + X * -1 + Y * -1
TRI = [T0 * (T0 * 2 + ONE * -1), X * (X * 2 + ONE * -1), Y * (Y * 2 + ONE * -1), X * T0 * 4, X * Y * 4, Y * T0 * 4]


def lit(v):
    f = float(v)
    if f == 0.0:
        return "0.0"
    return repr(f)


def arr(name, vals, shape):
    flat = list(vals)
    out = ["static constexpr double %s%s = {" % (name, "".join("[%d]" % s for s in shape))]
    line = "  "
    for i, v in enumerate(flat):
        line += lit(v) + ", "
        if (i + 1) % 6 == 0:
            out.append(line.rstrip())
            line = "  "
    if line.strip():
        out.append(line.rstrip())
    out.append("};")
    return "\n".join(out)


def main():
    dP2 = [[p.d(a) for a in range(3)] for p in P2]
    dP1 = [[p.d(a) for a in range(3)] for p in P1]
    rows = {"V": 0, "E": 4}
    print("// Generated by gen_tensors.py -- exact reference-tet integrals, do not edit.")
    print("// Index conventions are documented in gen_tensors.py and fedd_ns.hip.")
    print("#pragma once")
    print("namespace fedd_tensors {")
    for cls, r in rows.items():
        # Laplacian: KS[s][j], s over (00,11,22,01,02,12) symmetric pairs; off-diagonal pairs
        # carry K^{ab}+K^{ba}.  A_rj = |det| sum_s Gs[s] KS[s][j], G_ab = sum_d Binv[a][d] Binv[b][d]
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        ks = []
        for (a, b) in pairs:
            for j in range(10):
                v = (dP2[r][a] * dP2[j][b]).integral()
                if a != b:
                    v += (dP2[r][b] * dP2[j][a]).integral()
                ks.append(v)
        print(arr("KS_" + cls, ks, (6, 10)))
        # Convection: T[a][k][j] = int phi_r phi_k d_a phi_j ; N_rj = |det| sum_{k,a} ut_{k,a} T[a][k][j]
        t = [(P2[r] * P2[k] * dP2[j][a]).integral() for a in range(3) for k in range(10) for j in range(10)]
        print(arr("T_" + cls, t, (3, 10, 10)))
        # Newton term: L[m][j] = int lambda_m phi_r phi_j
        l = [(P1[m] * P2[r] * P2[j]).integral() for m in range(4) for j in range(10)]
        print(arr("L_" + cls, l, (4, 10)))
        # Elasticity: RG[a][b][j] = int d_a phi_r d_b phi_j (reference gradients, P2)
        rg = [(dP2[r][a] * dP2[j][b]).integral() for a in range(3) for b in range(3) for j in range(10)]
        print(arr("RG_" + cls, rg, (3, 3, 10)))
        # Divergence, velocity row r: R[a][m] = int psi_m d_a phi_r
        rr = [(P1[m] * dP2[r][a]).integral() for a in range(3) for m in range(4)]
        print(arr("RV_" + cls, rr, (3, 4)))
    # Divergence, pressure row 0: RP[a][j] = int psi_0 d_a phi_j
    rp = [(P1[0] * dP2[j][a]).integral() for a in range(3) for j in range(10)]
    print(arr("RP", rp, (3, 10)))
    # P2 gradients at the vertices: D[a][k][m] = d_a phi_k (v_m); grad u is P1, so
    # d u/d xhat_a = sum_m lambda_m sum_k u_k D[a][k][m]
    dv = [dP2[k][a].at(VERT[m]) for a in range(3) for k in range(10) for m in range(4)]
    print(arr("DV", dv, (3, 10, 4)))
    # P1 Laplacian row 0: KS1[s][j] (same pair convention as KS_*)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    ks1 = []
    for (a, b) in pairs:
        for j in range(4):
            v = (dP1[0][a] * dP1[j][b]).integral()
            if a != b:
                v += (dP1[0][b] * dP1[j][a]).integral()
            ks1.append(v)
    print(arr("KS1", ks1, (6, 4)))
    # P1 elasticity row 0: RG1[a][b][j] = int d_a lambda_0 d_b lambda_j
    rg1 = [(dP1[0][a] * dP1[j][b]).integral() for a in range(3) for b in range(3) for j in range(4)]
    print(arr("RG1", rg1, (3, 3, 4)))
    # Volume integrals of the basis (loads): MV2[k] = int phi_k (P2), MV1 = int lambda (P1)
    print(arr("MV2", [p.integral() for p in P2], (10,)))
    print(arr("MV1", [p.integral() for p in P1], (4,)))
    # Follower pressure on the 6-node surface triangle (reference triangle, FEDDLib's 2D P2
    # order): PT[a][b][c] = int N_a (d_xi N_b d_eta N_c - d_xi N_c d_eta N_b), so that
    # int N_a (x_xi x x_eta) = sum_{b<c} PT[a][b][c] (x_b x x_c)
    m = [[[(TRI[a] * TRI[b].d(0) * TRI[c].d(1)).integral_tri() for c in range(6)] for b in range(6)]
         for a in range(6)]
    pt = [m[a][b][c] - m[a][c][b] for a in range(6) for b in range(6) for c in range(6)]
    print(arr("PT", pt, (6, 6, 6)))
    print("}  // namespace fedd_tensors")


if __name__ == "__main__":
    main()
