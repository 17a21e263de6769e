"""The problems of FEDDLib's unit tests whose stored solutions the assembly reproduces.

Each reference test builds the structured unit cube with Domain::buildMesh(1, "Square",
3, FEType, n=2, m, ...) on 8 ranks.  The node numbering of MeshStructured::buildMesh3D
with rank offsets is the global lattice numbering (MeshStructured_def.hpp:736-737), which
is the one-rank numbering of the (2m)^3 cube, and the EpetraExt file stores values by
global id.  Boundary flags follow MeshStructured::setStructuredMeshFlags, option 1 in 3D
(MeshStructured_def.hpp:3174-3240): inflow face x=0 -> 2, open outflow face x=1 -> 3,
the four walls y=0, y=1, z=0, z=1 (including their edges on x=0 and x=1) -> 1.
Dirichlet rows become identity rows with the prescribed value on the right-hand side
(BCBuilder::setDirichletBC); parametersSolver.xml selects the direct solver Amesos2.

  - laplaceUnitTest.cpp: m=6 (12^3 cube), f=1, u=0 on flags 1, 2, 3, infinity-norm
    tolerance 1e-11 (:170);
  - linElasUnitTest.cpp: m=4 (8^3 cube), Mu=3, Poisson ratio 0.4, volume force (0.5, 0, 0)
    (rhsX3D with parameters [t, force, degree]), u=0 on flag 1 only, tolerance 1e-11;
  - stokesUnitTest.cpp: m=3 (6^3 cube), P2-P1, viscosity 0.01, density 1, u=0 on flag 1,
    parabolic inflow 16 y (1-y) z (1-z) e_x on flag 2, do-nothing outflow on flag 3,
    relative tolerance 1e-9 (:230).
"""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_h5")

LAPLACE = {1: ("solution_laplace_3d_P1_8cores.h5", 12), 2: ("solution_laplace_3d_P2_8cores.h5", 12)}
LINELAS = ("solution_linElas_3d_P2_8cores.h5", 8)
LINELAS_FILES = {1: "solution_linElas_3d_P1_8cores.h5", 2: "solution_linElas_3d_P2_8cores.h5"}
STOKES = ("solution_stokes_velocity_3d_P2_8cores.h5", "solution_stokes_pressure_3d_P2_8cores.h5", 6)

LINELAS_MU, LINELAS_NU, LINELAS_FORCE = 3.0, 0.4, 0.5
LINELAS_E = LINELAS_MU * 2.0 * (1.0 + LINELAS_NU)  # LinElas_def.hpp:75
LINELAS_LAMBDA = LINELAS_NU * LINELAS_E / ((1.0 + LINELAS_NU) * (1.0 - 2.0 * LINELAS_NU))  # LinElas_def.hpp:76
STOKES_NU, STOKES_RHO = 0.01, 1.0


def golden(name):
    return os.path.join(GOLDEN, name)


def flags(pts, tol=1e-12):
    """Boundary flags of setStructuredMeshFlags option 1 on the unit cube (0: interior)."""
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    f = np.zeros(len(pts), np.int32)
    f[x < tol] = 2
    f[(x > 1 - tol) & (y > tol) & (y < 1 - tol) & (z > tol) & (z < 1 - tol)] = 3
    f[(z < tol) | (z > 1 - tol) | (y < tol) | (y > 1 - tol)] = 1
    return f


def node_dofs(nodes, dofs_per_node=3):
    nodes = np.asarray(nodes, np.int64)
    return np.sort((dofs_per_node * nodes[:, None] + np.arange(dofs_per_node)).ravel())


def csr(t, shape):
    """(rowptr, colind, values) -> scipy CSR."""
    return sp.csr_matrix((t[2], t[1], t[0]), shape=shape)


def with_dirichlet(A, b, rows, vals):
    """Identity rows with the prescribed values on the right-hand side."""
    A = sp.csr_matrix(A, copy=True)
    rows = np.asarray(rows, np.int64)
    keep = np.ones(A.shape[0], bool)
    keep[rows] = False
    D = sp.diags(keep.astype(float))
    E = sp.csr_matrix((np.ones(len(rows)), (rows, rows)), shape=A.shape)
    b = np.array(b, float)
    b[rows] = vals
    return (D @ A + E).tocsr(), b


def solve_with_dirichlet(A, b, rows, vals):
    Ad, bd = with_dirichlet(A, b, rows, vals)
    return spla.spsolve(Ad.tocsc(), bd)


def stokes_dirichlet(p2):
    """Velocity dof rows and values of the Stokes test (flags 1 and 2)."""
    f = flags(p2)
    nodes = np.where((f == 1) | (f == 2))[0]
    rows = node_dofs(nodes)
    vals = np.zeros(len(rows))
    inflow = np.zeros(len(p2))
    y, z = p2[:, 1], p2[:, 2]
    inflow[f == 2] = (16.0 * y * (1 - y) * z * (1 - z))[f == 2]  # inflowParabolic3D, U = H = 1
    xrows = rows % 3 == 0
    vals[xrows] = inflow[rows[xrows] // 3]
    return rows, vals


def stokes_system(A, BT, B):
    """[A BT; B 0] as one CSR matrix (velocity dofs first)."""
    return sp.bmat([[A, BT], [B, None]], format="csr")
