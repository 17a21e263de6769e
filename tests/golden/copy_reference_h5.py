"""Copies the EpetraExt-HDF5 reference solutions of FEDDLib's unit tests that the
assembly paths here can reproduce, and records where they came from.

Source: feddlib/problems/tests/UnitTest/ReferenceSolutions/ (written by HDF5Export in
laplaceUnitTest.cpp, linElasUnitTest.cpp and stokesUnitTest.cpp, read back there with
HDF5Import and compared in the infinity norm).  The files are data (the outputs of
the reference's own tests).
Run in a checkout that has the reference at /root/reference.
"""
import hashlib
import os
import shutil

SRC = "/root/reference/feddlib/problems/tests/UnitTest/ReferenceSolutions"
FILES = [
    "solution_laplace_3d_P1_8cores.h5",        # laplaceUnitTest.cpp, dim 3, P1, m=6 on 2x2x2 ranks
    "solution_laplace_3d_P2_8cores.h5",        # laplaceUnitTest.cpp, dim 3, P2
    "solution_linElas_3d_P1_8cores.h5",        # linElasUnitTest.cpp, dim 3, P1, m=4
    "solution_linElas_3d_P2_8cores.h5",        # linElasUnitTest.cpp, dim 3, P2, m=4
    "solution_stokes_velocity_3d_P2_8cores.h5",  # stokesUnitTest.cpp, dim 3, P2-P1, m=3
    "solution_stokes_pressure_3d_P2_8cores.h5",
]

if __name__ == "__main__":
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_h5")
    os.makedirs(here, exist_ok=True)
    lines = []
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(here, f))
        h = hashlib.sha256(open(os.path.join(here, f), "rb").read()).hexdigest()
        lines.append("%s  %s  (feddlib/problems/tests/UnitTest/ReferenceSolutions/%s)" % (h, f, f))
    open(os.path.join(here, "MANIFEST.txt"), "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))
