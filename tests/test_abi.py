"""CPU tests of the C ABI: symbols, host-side mesh builders, error behaviour (no GPU)."""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from feddlib_amd import native as N
from feddlib_amd import workloads as W


def test_library_exports_every_declared_symbol():
    L = N.lib()
    declared = N.declared_symbols()
    assert len(declared) >= 25
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(N.declared_symbols()) <= set(N._SIGS)


@pytest.mark.parametrize("M", [1, 2, 3, 5])
@pytest.mark.parametrize("fe", [1, 2])
def test_mesh_structured_bit_exact(M, fe):
    a = N.mesh_structured(M, fe)
    b = O.mesh(M, fe)
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and np.array_equal(x, y)


@pytest.mark.parametrize("fe", [1, 2])
def test_correct_orientation_bit_exact(fe):
    pts, _, el = N.mesh_structured(3, fe)
    el2 = el.copy()
    assert N.correct_orientation(pts, el) == O.correct_orientation(pts, el2)
    assert np.array_equal(el, el2)


def test_mesh_builder_rejects_bad_arguments():
    pts = np.zeros((8, 3))
    el = np.zeros((6, 4), np.int32)
    rc = N.lib().fedd_mesh_structured(0, 1, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None,
                                      el.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert rc == -1


def test_no_device_fails_loudly():
    """Without a HIP device the context cannot be created and nothing falls back to the CPU."""
    if N.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(N.FeddError) as e:
        N.Context(0)
    assert e.value.code == -5


def test_workloads_are_seeded():
    p2, e2, p1, e1, u, p = W.ns_problem(2, 0)
    _, _, _, _, u2, _ = W.ns_problem(2, 0)
    assert np.array_equal(u, u2) and u.shape == (p2.shape[0], 3) and np.all(p == 0)
    assert np.max(np.abs(u)) <= 8.0
    pts, _, el = W.jittered_p2_mesh(3, 0)
    x = pts[el]
    d = np.linalg.det(np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], 2))
    assert np.all(d > 0)  # orientation-corrected, no inverted tets at +-0.2h
    # boundary vertices stay on the faces
    on_b = np.any((np.abs(pts) < 1e-14) | (np.abs(pts - 1) < 1e-14), axis=1)
    assert on_b.sum() == 7 ** 3 - 5 ** 3  # (2M+1)^3 - (2M-1)^3 boundary lattice nodes, M = 3
    u = W.smoothed_displacement(pts, el, 1)
    assert u.shape == pts.shape and np.abs(u).max() < 0.2
