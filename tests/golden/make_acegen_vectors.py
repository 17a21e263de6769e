"""Write tests/golden/acegen_tet10.npz: input/output vectors of the reference's own
AceGen tet-10 elements skr1/skr2/skr3 (NeoHookQuadraticTets{,2,3}.f), compiled from
the reference sources by oracle/Makefile into oracle/_ref/libacegen_ref.so.

Inputs: four seeded P2 tets (vertices perturbed around a scaled unit tet, edge
nodes at midpoints, FEDDLib local order, positive FEDDLib orientation) with
seeded displacements of amplitude 0.02, E = 1000, nu = 0.3 (and one element at
E = 5, nu = 0.45).  Outputs: s(30,30) column-major and p(30), exactly as skr* returns them.
Run from the repository root:  python tests/golden/make_acegen_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]


def elements(seed=11, n=4):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * (0.2 + 0.1 * k)
        V += rng.normal(0, 0.02, (4, 3))
        X = np.vstack([V] + [(V[a] + V[b]) / 2 for a, b in EDGES])
        U = rng.normal(0, 0.02, (10, 3))
        out.append((X, U))
    return out


def main():
    if O.ref_lib() is None:
        sys.exit("oracle/_ref/libacegen_ref.so is not built (make -C oracle ref)")
    xs, us, ms, Es, nus, ss, ps = [], [], [], [], [], [], []
    for model in range(3):
        for k, (X, U) in enumerate(elements()):
            E, nu = (5.0, 0.45) if k == 3 else (1000.0, 0.3)
            s, p = O.ref_skr(model, E, nu, X, U)
            xs.append(X.ravel())
            us.append(U.ravel())
            ms.append(model)
            Es.append(E)
            nus.append(nu)
            ss.append(s)
            ps.append(p)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "acegen_tet10.npz"), xl=np.array(xs),
                        ul=np.array(us), model=np.array(ms), E=np.array(Es), nu=np.array(nus),
                        s=np.array(ss), p=np.array(ps))


if __name__ == "__main__":
    main()
