"""End-to-end pins of the oracle against the solutions FEDDLib's own unit tests store.

laplaceUnitTest.cpp, linElasUnitTest.cpp and stokesUnitTest.cpp solve small problems on
the structured cube and compare with EpetraExt-HDF5 files in ReferenceSolutions/.  The
same problems, assembled by the oracle (CPU restatement) and solved directly, must
reproduce those files within the reference tests' own tolerances.  The files are read
with the library's context-free HDF5 reader (fedd_hdf5_read_vector), which needs no GPU.
"""
import numpy as np
import pytest

import pyoracle as O
import reference_problems as RP
from feddlib_amd import native as N


@pytest.mark.parametrize("fname,var,n", [
    ("solution_laplace_3d_P1_8cores.h5", "solution", 13 ** 3),
    ("solution_laplace_3d_P2_8cores.h5", "solution", 25 ** 3),
    ("solution_linElas_3d_P2_8cores.h5", "solution", 3 * 17 ** 3),
    ("solution_stokes_velocity_3d_P2_8cores.h5", "velocity", 3 * 13 ** 3),
    ("solution_stokes_pressure_3d_P2_8cores.h5", "pressure", 7 ** 3),
])
def test_hdf5_reader(fname, var, n):
    v = N.hdf5_read_vector(RP.golden(fname), var)
    assert v.shape == (n,) and np.all(np.isfinite(v))
    with pytest.raises(N.FeddError, match="Requested varName: nothing not contained"):
        N.hdf5_read_vector(RP.golden(fname), "nothing")


def test_hdf5_reader_rejects_bad_files(tmp_path):
    src = open(RP.golden("solution_laplace_3d_P1_8cores.h5"), "rb").read()
    with pytest.raises(N.FeddError, match="not an HDF5 file"):
        N.hdf5_read_vector(RP.golden("MANIFEST.txt"), "solution")
    with pytest.raises(N.FeddError, match="cannot open"):
        N.hdf5_read_vector(tmp_path / "missing.h5", "solution")
    for cut in (100, 1000, 3000, len(src) - 8):  # truncated copies end in an error, never a crash
        p = tmp_path / ("cut%d.h5" % cut)
        p.write_bytes(src[:cut])
        with pytest.raises(N.FeddError):
            N.hdf5_read_vector(p, "solution")


@pytest.mark.parametrize("fe", [1, 2])
def test_oracle_laplace_reproduces_reference_solution(fe):
    fname, M = RP.LAPLACE[fe]
    pts, _, el = O.mesh(M, fe)
    n = len(pts)
    A = RP.csr(O.laplace(fe, pts, el), (n, n))
    b = O.rhs(fe, pts, el, 1.0)
    rows = np.where(RP.flags(pts) > 0)[0]
    u = RP.solve_with_dirichlet(A, b, rows, 0.0)
    g = N.hdf5_read_vector(RP.golden(fname), "solution")
    assert np.abs(u - g).max() <= 1e-11  # laplaceUnitTest.cpp:170


@pytest.mark.parametrize("model", [0, 1, 2])
def test_oracle_linelas_reproduces_reference_solution(model):
    """assemblyLinElasXDim (FE_def.hpp:4701) equals the tangent at u=0 of every element
    law here (skr1/skr2/skr3) with E = 2 mu (1 + nu); the volume force is the vector
    form of FE::assemblyRHS (FE_def.hpp:7303-7395)."""
    fname, M = RP.LINELAS
    pts, _, el = O.mesh(M, 2)
    el = np.ascontiguousarray(el)
    O.correct_orientation(pts, el)
    n = 3 * len(pts)
    K, _ = O.elasticity(model, RP.LINELAS_E, RP.LINELAS_NU, pts, el, np.zeros(n))
    b = np.zeros(n)
    b[0::3] = RP.LINELAS_FORCE * O.rhs(2, pts, el, 1.0)
    rows = RP.node_dofs(np.where(RP.flags(pts) == 1)[0])
    u = RP.solve_with_dirichlet(RP.csr(K, (n, n)), b, rows, 0.0)
    g = N.hdf5_read_vector(RP.golden(fname), "solution")
    assert np.abs(u - g).max() <= 1e-11  # linElasUnitTest.cpp tolerance


def test_oracle_stokes_reproduces_reference_solution():
    """Stokes::assemble (Stokes_def.hpp:48-90) is the Navier-Stokes Jacobian at u = 0."""
    fv, fp, M = RP.STOKES
    p2, _, e2 = O.mesh(M, 2)
    p1, _, e1 = O.mesh(M, 1)
    nv, npr = 3 * len(p2), len(p1)
    ref = O.ns_assembly(p2, e2, e1, npr, np.zeros(nv), np.zeros(npr), RP.STOKES_NU, RP.STOKES_RHO, True, 0, False)
    K = RP.stokes_system(RP.csr(ref["A"], (nv, nv)), RP.csr(ref["BT"], (nv, npr)), RP.csr(ref["B"], (npr, nv)))
    rows, vals = RP.stokes_dirichlet(p2)
    Kd, bd = RP.with_dirichlet(K, np.zeros(nv + npr), rows, vals)
    x = RP.solve_with_dirichlet(K, np.zeros(nv + npr), rows, vals)
    gv = N.hdf5_read_vector(RP.golden(fv), "velocity")
    gp = N.hdf5_read_vector(RP.golden(fp), "pressure")
    assert np.abs(x[:nv] - gv).max() <= 1e-9 * np.abs(gv).max()  # stokesUnitTest.cpp:230
    assert np.abs(x[nv:] - gp).max() <= 1e-9 * np.abs(gp).max()
    # the stored solution is a solution of this system to the reference solver's tolerance
    g = np.concatenate([gv, gp])
    assert np.linalg.norm(Kd @ g - bd) <= 1e-10 * np.linalg.norm(bd)


def test_hdf5_reader_survives_corrupted_metadata(tmp_path):
    """Single-byte corruptions of the file's metadata end in a value or an error."""
    src = bytearray(open(RP.golden("solution_laplace_3d_P1_8cores.h5"), "rb").read())
    bad = bytearray(src)
    bad[0x40:0x48] = (0x7FFFFFFF).to_bytes(8, "little")  # root object header address
    p = tmp_path / "root.h5"
    p.write_bytes(bytes(bad))
    with pytest.raises(N.FeddError, match="address outside the file"):
        N.hdf5_read_vector(p, "solution")
    rng = np.random.default_rng(0)
    p = tmp_path / "fuzz.h5"
    for trial in range(300):
        bad = bytearray(src)
        pos = int(rng.integers(0, 0x800))
        bad[pos] = int(rng.integers(0, 256))
        p.write_bytes(bytes(bad))
        try:
            v = N.hdf5_read_vector(p, "solution")
            assert v.ndim == 1
        except N.FeddError:
            pass


def test_oracle_linelas_xdim_vectorised_equals_literal():
    """The vectorised assembly of oracle/linelas.py equals the literal epsilon-tensor
    element loop of FE::assemblyLinElasXDim on single elements (P1 and P2)."""
    import linelas as LE
    for fe in (1, 2):
        pts, _, el = O.mesh(1, fe)
        for e in (0, 3, 5):
            T = el[e]
            Kl = LE.element_matrix(fe, pts[T], RP.LINELAS_LAMBDA, RP.LINELAS_MU)
            Kv = LE.linelas_xdim(fe, pts, el[e:e + 1], RP.LINELAS_LAMBDA, RP.LINELAS_MU).toarray()
            dofs = (3 * T[:, None] + np.arange(3)).ravel()
            assert np.abs(Kv[np.ix_(dofs, dofs)] - Kl).max() <= 1e-13 * np.abs(Kl).max()


def test_oracle_linelas_xdim_equals_hyperelastic_tangent_at_rest():
    """Two independent restatements: assemblyLinElasXDim (lambda, mu) and the tangent at
    u = 0 of the AceGen laws with E = 2 mu (1 + nu)."""
    import linelas as LE
    pts, _, el = O.mesh(2, 2)
    el = np.ascontiguousarray(el)
    O.correct_orientation(pts, el)
    n = 3 * len(pts)
    Kx = LE.linelas_xdim(2, pts, el, RP.LINELAS_LAMBDA, RP.LINELAS_MU)
    for model in (0, 1, 2):
        K, _ = O.elasticity(model, RP.LINELAS_E, RP.LINELAS_NU, pts, el, np.zeros(n))
        K = RP.csr(K, (n, n))
        assert abs(K - Kx).max() <= 1e-12 * abs(Kx).max()


@pytest.mark.parametrize("fe", [1, 2])
def test_oracle_linelas_xdim_reproduces_reference_solution(fe):
    import linelas as LE
    _, M = RP.LINELAS
    pts, _, el = O.mesh(M, fe)
    n = 3 * len(pts)
    K = LE.linelas_xdim(fe, pts, el, RP.LINELAS_LAMBDA, RP.LINELAS_MU)
    b = np.zeros(n)
    b[0::3] = RP.LINELAS_FORCE * O.rhs(fe, pts, el, 1.0)
    rows = RP.node_dofs(np.where(RP.flags(pts) == 1)[0])
    u = RP.solve_with_dirichlet(K, b, rows, 0.0)
    g = N.hdf5_read_vector(RP.golden(RP.LINELAS_FILES[fe]), "solution")
    assert np.abs(u - g).max() <= 1e-11  # linElasUnitTest.cpp tolerance
