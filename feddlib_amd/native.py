"""ctypes binding of libfeddgpu.so (include/feddgpu.h).

The library is the product: this module only moves host arrays across the C ABI.
It raises when the library is missing -- there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfeddgpu.so")

FEDD_OK = 0
MESH_PRIMARY, MESH_PRESSURE = 0, 1
GRAPH_SCALAR, GRAPH_VECTOR, GRAPH_NS = 0, 1, 2
BLOCK_A, BLOCK_BT, BLOCK_B, BLOCK_C = 0, 1, 2, 3
VEC_U, VEC_P, VEC_RES_U, VEC_RES_P, VEC_RHS, VEC_FEXT, VEC_X, VEC_B, VEC_X_P, VEC_B_P, VEC_FPRES = range(11)
MAX_SURFACE_FLAGS = 16
FP64, FP32_LOCAL = 0, 1
BC_RHS, BC_MINUS_VECTOR, BC_VECTOR_MINUS_BC = 0, 1, 2
NS_JACOBIAN, NS_FIXEDPOINT, NS_RHS, NS_JACOBIAN_RHS, NS_FIXEDPOINT_RHS = 0, 1, 2, 3, 4
METHOD_ROW_GATHER, METHOD_COLOUR_SCATTER = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
MLP_PARAMS = 7328


class NsParams(ctypes.Structure):
    _fields_ = [("viscosity", ctypes.c_double), ("density", ctypes.c_double),
                ("newton", ctypes.c_int), ("coeff", ctypes.c_double * 4)]


class NewtonParams(ctypes.Structure):
    _fields_ = [("neo_hooke_model", ctypes.c_int), ("E", ctypes.c_double), ("nu", ctypes.c_double),
                ("load_steps", ctypes.c_int), ("abs_tol", ctypes.c_double), ("max_newton", ctypes.c_int),
                ("lin_rtol", ctypes.c_double), ("lin_maxit", ctypes.c_int), ("tangent", ctypes.c_int),
                ("precision", ctypes.c_int), ("stall_ratio", ctypes.c_double), ("collect_samples", ctypes.c_int),
                ("preconditioner", ctypes.c_int), ("mg_reuse", ctypes.c_int), ("criterion", ctypes.c_int),
                ("rel_tol", ctypes.c_double)]


class NewtonStats(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_int), ("fallbacks", ctypes.c_int), ("rejected", ctypes.c_int),
                ("krylov_iters", ctypes.c_int64),
                ("seconds", ctypes.c_double), ("assembly_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("final_residual", ctypes.c_double), ("converged", ctypes.c_int), ("history_len", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("last_update", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NsNewtonParams(ctypes.Structure):
    _fields_ = [("viscosity", ctypes.c_double), ("density", ctypes.c_double), ("linearization", ctypes.c_int),
                ("criterion", ctypes.c_int), ("rel_tol", ctypes.c_double), ("max_its", ctypes.c_int),
                ("lin_rtol", ctypes.c_double), ("restart", ctypes.c_int), ("lin_maxit", ctypes.c_int)]


class NsNewtonStats(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_int), ("krylov_iters", ctypes.c_int64),
                ("final_residual", ctypes.c_double), ("rel_residual", ctypes.c_double),
                ("last_update", ctypes.c_double), ("last_lin_relres", ctypes.c_double),
                ("converged", ctypes.c_int), ("history_len", ctypes.c_int), ("seconds", ctypes.c_double),
                ("assembly_ms", ctypes.c_double), ("solve_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FeddError(RuntimeError):
    """Error raised by a libfeddgpu call (std::logic_error/runtime_error in FEDDLib)."""

    def __init__(self, code, msg):
        super().__init__("fedd error %d: %s" % (code, msg))
        self.code = code


_SIGS = {
    "fedd_device_count": (ctypes.c_int, [_ip]),
    "fedd_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "fedd_ctx_destroy": (None, [_vp]),
    "fedd_last_error": (ctypes.c_char_p, [_vp]),
    "fedd_synchronize": (ctypes.c_int, [_vp]),
    "fedd_mesh_structured": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _ip, _ip]),
    "fedd_mesh_correct_orientation": (ctypes.c_int, [_dp, ctypes.c_int64, _ip, ctypes.c_int64,
                                                      ctypes.c_int, _lp]),
    "fedd_mesh_upload": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64, _ip, ctypes.c_int64,
                                        ctypes.c_int]),
    "fedd_graph_build": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fedd_block_info": (ctypes.c_int, [_vp, ctypes.c_int, _lp, _lp, _lp]),
    "fedd_block_graph": (ctypes.c_int, [_vp, ctypes.c_int, _lp, _ip]),
    "fedd_block_values": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "fedd_block_device_values": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_dp)]),
    "fedd_ns_merged_info": (ctypes.c_int, [_vp, _lp, _lp]),
    "fedd_ns_merged_export": (ctypes.c_int, [_vp, _lp, _ip, _dp]),
    "fedd_write_matrix_market": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p]),
    "fedd_assembly_linear_elasticity_xdim": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "fedd_hdf5_read_vector": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_int64, _lp,
                                             ctypes.POINTER(ctypes.c_int)]),
    "fedd_colour_build": (ctypes.c_int, [_vp, _ip]),
    "fedd_colour_export": (ctypes.c_int, [_vp, _lp, _ip]),
    "fedd_vector_upload": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64]),
    "fedd_vector_download": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64]),
    "fedd_vector_device_ptr": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_dp), _lp]),
    "fedd_assembly_laplace": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fedd_assembly_rhs": (ctypes.c_int, [_vp, ctypes.c_double]),
    "fedd_rhs_quadrature_points": (ctypes.c_int, [_vp, _dp, ctypes.c_int64, _ip]),
    "fedd_assembly_rhs_points": (ctypes.c_int, [_vp, _dp, ctypes.c_int64, ctypes.c_int]),
    "fedd_assembly_navier_stokes": (ctypes.c_int, [_vp, ctypes.POINTER(NsParams), ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_int]),
    "fedd_newton_step": (ctypes.c_int, [_vp, _ip]),
    "fedd_assembly_linear_elasticity": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "fedd_assembly_nonlinear_elasticity": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_int]),
    "fedd_assembly_nonlinear_elasticity_efield": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64,
                                                                 ctypes.c_double, ctypes.c_int]),
    "fedd_dirichlet_set": (ctypes.c_int, [_vp, _ip, ctypes.c_int64]),
    "fedd_apply_dirichlet_symmetric": (ctypes.c_int, [_vp]),
    "fedd_dirichlet_values": (ctypes.c_int, [_vp, _dp, ctypes.c_int64]),
    "fedd_dirichlet_set_system": (ctypes.c_int, [_vp]),
    "fedd_dirichlet_vector": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fedd_traction_load": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp]),
    "fedd_structured_surfaces": (ctypes.c_int, [_dp, ctypes.c_int64, _ip, ctypes.c_int64, _dp, _ip, _ip, _lp]),
    "fedd_surface_set": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_int64]),
    "fedd_assembly_surface_pressure": (ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fedd_newton_pressure": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "fedd_set_preconditioner": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fedd_linear_solve": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _ip, _dp]),
    "fedd_elasticity_fp64_fixups": (ctypes.c_int, [_vp, _lp, _lp]),
    "fedd_assembly_nonlinear_elasticity_tangent": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                                  ctypes.c_int, ctypes.c_int]),
    "fedd_ns_linear_solve": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _dp]),
    "fedd_newton_navier_stokes": (ctypes.c_int, [_vp, ctypes.POINTER(NsNewtonParams), ctypes.POINTER(NsNewtonStats),
                                                 _dp, ctypes.c_int]),
    "fedd_linear_solve_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                                 _ip, _dp]),
    "fedd_newton_elasticity": (ctypes.c_int, [_vp, ctypes.POINTER(NewtonParams), ctypes.POINTER(NewtonStats),
                                              _dp, ctypes.c_int]),
    "fedd_mlp_init": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]),
    "fedd_mlp_add_samples_random": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64]),
    "fedd_mlp_add_samples_qp": (ctypes.c_int, [_vp]),
    "fedd_mlp_num_samples": (ctypes.c_int, [_vp, _lp]),
    "fedd_mlp_train": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, _fp]),
    "fedd_mlp_params": (ctypes.c_int, [_vp, _fp]),
    "fedd_mlp_set_params": (ctypes.c_int, [_vp, _fp]),
    "fedd_mlp_forward": (ctypes.c_int, [_vp, _fp, ctypes.c_int64, _fp]),
    "fedd_mlp_gradient": (ctypes.c_int, [_vp, _fp, _fp, ctypes.c_int64, _fp, _fp]),
    "fedd_mlp_norm": (ctypes.c_int, [_vp, _fp, _fp, _dp]),
    "fedd_capture_begin": (ctypes.c_int, [_vp]),
    "fedd_capture_end": (ctypes.c_int, [_vp, _lp]),
    "fedd_graph_launch": (ctypes.c_int, [_vp]),
    "fedd_stream_copy": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "fedd_event_record": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fedd_event_elapsed": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_float)]),
    "fedd_kernel_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fedd_set_reproducible": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fedd_set_assembly_method": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fedd_scatter_indices": (ctypes.c_int, [_vp, ctypes.c_int, _lp]),
    "fedd_kernel_time": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp, _lp]),
}

_lib = None


def lib():
    """Load libfeddgpu.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libfeddgpu.so is not built; run `python -m feddlib_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names of the entry points declared in include/feddgpu.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "feddgpu.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(fedd_[a-z0-9_]+)\s*\(", txt)))


def _p(a, t):
    return a.ctypes.data_as(t)


def mesh_structured(M, fe):
    """Unit-cube Kuhn mesh (MeshStructured::buildMesh3D): (pts[n,3], bc[n], elems[E,k])."""
    P = M + 1 if fe == 1 else 2 * M + 1
    n, E, k = P ** 3, 6 * M ** 3, 4 if fe == 1 else 10
    pts = np.zeros((n, 3))
    bc = np.zeros(n, np.int32)
    el = np.zeros((E, k), np.int32)
    rc = lib().fedd_mesh_structured(M, fe, _p(pts, _dp), _p(bc, _ip), _p(el, _ip))
    if rc:
        raise FeddError(rc, "fedd_mesh_structured")
    return pts, bc, el


def structured_surfaces(pts, elems, box=(0.0, 0.0, 0.0, 1.0, 1.0, 1.0)):
    """MeshStructured::buildSurfaces(3, "P2"): the boundary faces of a P2 tet mesh of the
    box (x0, y0, z0, L, W, H) as outward 6-node triangles (n, 6) and their side flags (n,)."""
    L = lib()
    pts = np.ascontiguousarray(pts, np.float64)
    elems = np.ascontiguousarray(elems, np.int32)
    b = np.ascontiguousarray(box, np.float64)
    n = ctypes.c_int64()
    rc = L.fedd_structured_surfaces(_p(pts, _dp), pts.shape[0], _p(elems, _ip), elems.shape[0], _p(b, _dp), None,
                                    None, ctypes.byref(n))
    if rc:
        raise FeddError(rc, L.fedd_last_error(None).decode())
    faces = np.zeros((n.value, 6), np.int32)
    flags = np.zeros(n.value, np.int32)
    rc = L.fedd_structured_surfaces(_p(pts, _dp), pts.shape[0], _p(elems, _ip), elems.shape[0], _p(b, _dp),
                                    _p(faces, _ip), _p(flags, _ip), ctypes.byref(n))
    if rc:
        raise FeddError(rc, L.fedd_last_error(None).decode())
    return faces, flags


def hdf5_read_vector(path, name):
    """HDF5Import::readVariablesHDF5 (HDF5Import_def.hpp:52-70): the EpetraExt multivector
    `name` of an HDF5 file, column 0, in global-id order (float64 array)."""
    L = lib()
    n, k = ctypes.c_int64(0), ctypes.c_int(0)
    rc = L.fedd_hdf5_read_vector(str(path).encode(), name.encode(), None, 0, ctypes.byref(n), ctypes.byref(k))
    if rc:
        raise FeddError(rc, L.fedd_last_error(None).decode())
    out = np.zeros(n.value)
    rc = L.fedd_hdf5_read_vector(str(path).encode(), name.encode(), _p(out, _dp), out.size, None, None)
    if rc:
        raise FeddError(rc, L.fedd_last_error(None).decode())
    return out


def correct_orientation(pts, elems):
    """Mesh::correctElementOrientation in place; returns the number of flipped elements."""
    assert elems.flags.c_contiguous and elems.dtype == np.int32
    f = ctypes.c_int64()
    rc = lib().fedd_mesh_correct_orientation(_p(np.ascontiguousarray(pts, np.float64), _dp),
                                             pts.shape[0], _p(elems, _ip), elems.shape[0],
                                             elems.shape[1], ctypes.byref(f))
    if rc:
        raise FeddError(rc, "fedd_mesh_correct_orientation")
    return f.value


def device_count():
    n = ctypes.c_int(0)
    lib().fedd_device_count(ctypes.byref(n))
    return n.value


class Context:
    """One libfeddgpu context (device buffers + HIP stream)."""

    def __init__(self, device=0):
        self._L = lib()
        h = _vp()
        rc = self._L.fedd_ctx_create(device, ctypes.byref(h))
        if rc:
            raise FeddError(rc, "fedd_ctx_create failed (no HIP device?)")
        self.h = h

    def close(self):
        if self.h:
            self._L.fedd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc:
            raise FeddError(rc, self._L.fedd_last_error(self.h).decode())

    # mesh / graph
    def mesh_upload(self, slot, pts, elems):
        pts = np.ascontiguousarray(pts, np.float64)
        elems = np.ascontiguousarray(elems, np.int32)
        self._chk(self._L.fedd_mesh_upload(self.h, slot, _p(pts, _dp), pts.shape[0], _p(elems, _ip),
                                           elems.shape[0], elems.shape[1]))

    def set_reproducible(self, on=True):
        """Bitwise-repeatable row-gather sums for graphs built afterwards (see the header)."""
        self._chk(self._L.fedd_set_reproducible(self.h, 1 if on else 0))

    def set_assembly_method(self, method):
        """METHOD_ROW_GATHER or METHOD_COLOUR_SCATTER for NS graphs built afterwards."""
        self._chk(self._L.fedd_set_assembly_method(self.h, method))

    def scatter_indices(self, block, nelem):
        """Node-level CRS positions of every element block (see fedd_scatter_indices)."""
        out = np.zeros(nelem * (100 if block == 0 else 40), np.int64)
        self._chk(self._L.fedd_scatter_indices(self.h, block, _p(out, _lp)))
        return out

    def graph_build(self, kind):
        self._chk(self._L.fedd_graph_build(self.h, kind))

    def block_info(self, block):
        r, c, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._chk(self._L.fedd_block_info(self.h, block, ctypes.byref(r), ctypes.byref(c), ctypes.byref(n)))
        return r.value, c.value, n.value

    def block_graph(self, block):
        nr, _, nnz = self.block_info(block)
        rp = np.zeros(nr + 1, np.int64)
        ci = np.zeros(max(nnz, 1), np.int32)
        self._chk(self._L.fedd_block_graph(self.h, block, _p(rp, _lp), _p(ci, _ip)))
        return rp, ci[:nnz]

    def block_values(self, block):
        _, _, nnz = self.block_info(block)
        v = np.zeros(max(nnz, 1))
        self._chk(self._L.fedd_block_values(self.h, block, _p(v, _dp)))
        return v[:nnz]

    def ns_merged(self):
        nr, nz = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self._L.fedd_ns_merged_info(self.h, ctypes.byref(nr), ctypes.byref(nz)))
        rp = np.zeros(nr.value + 1, np.int64)
        ci = np.zeros(nz.value, np.int32)
        v = np.zeros(nz.value)
        self._chk(self._L.fedd_ns_merged_export(self.h, _p(rp, _lp), _p(ci, _ip), _p(v, _dp)))
        return rp, ci, v

    def write_matrix_market(self, block, path):
        self._chk(self._L.fedd_write_matrix_market(self.h, block, str(path).encode()))

    def colour_build(self):
        n = ctypes.c_int()
        self._chk(self._L.fedd_colour_build(self.h, ctypes.byref(n)))
        return n.value

    def colour_export(self, ncolours, nelem):
        cp = np.zeros(ncolours + 1, np.int64)
        perm = np.zeros(nelem, np.int32)
        self._chk(self._L.fedd_colour_export(self.h, _p(cp, _lp), _p(perm, _ip)))
        return cp, perm

    # vectors
    def vector_upload(self, which, a):
        a = np.ascontiguousarray(a, np.float64).ravel()
        self._chk(self._L.fedd_vector_upload(self.h, which, _p(a, _dp), a.size))

    def vector_device_ptr(self, which):
        """(device address as int, length) of a context vector."""
        d, n = _dp(), ctypes.c_int64()
        self._chk(self._L.fedd_vector_device_ptr(self.h, which, ctypes.byref(d), ctypes.byref(n)))
        return ctypes.cast(d, ctypes.c_void_p).value or 0, n.value

    def vector_download(self, which, n):
        a = np.zeros(n)
        self._chk(self._L.fedd_vector_download(self.h, which, _p(a, _dp), n))
        return a

    # assembly
    def assembly_laplace(self, precision=FP64):
        self._chk(self._L.fedd_assembly_laplace(self.h, precision))

    def assembly_rhs(self, f):
        self._chk(self._L.fedd_assembly_rhs(self.h, f))

    def rhs_quadrature_points(self, nelem):
        """Physical quadrature points of FE::assemblyRHS's rule: (nelem, nq, 3)."""
        nq = ctypes.c_int()
        self._chk(self._L.fedd_rhs_quadrature_points(self.h, None, 0, ctypes.byref(nq)))
        xq = np.zeros(nelem * nq.value * 3)
        self._chk(self._L.fedd_rhs_quadrature_points(self.h, _p(xq, _dp), xq.size, ctypes.byref(nq)))
        return xq.reshape(nelem, nq.value, 3)

    def assembly_rhs_points(self, fq):
        """FE::assemblyRHS from source values at those points: fq (nelem, nq) ("Scalar")
        or (nelem, nq, 3) ("Vector"); the result is in VEC_RHS."""
        fq = np.ascontiguousarray(fq, np.float64)
        ncomp = 1 if fq.ndim == 2 else fq.shape[2]
        self._chk(self._L.fedd_assembly_rhs_points(self.h, _p(fq, _dp), fq.size, ncomp))

    def assembly_navier_stokes(self, viscosity, density, newton, mode, reassemble, precision=FP64,
                               coeff=(1.0, 1.0, 1.0, 1.0)):
        p = NsParams(viscosity, density, 1 if newton else 0, (ctypes.c_double * 4)(*coeff))
        self._chk(self._L.fedd_assembly_navier_stokes(self.h, ctypes.byref(p), mode, 1 if reassemble else 0,
                                                      precision))

    def assembly_linear_elasticity(self, E, nu, precision=FP64):
        self._chk(self._L.fedd_assembly_linear_elasticity(self.h, E, nu, precision))

    def assembly_linear_elasticity_xdim(self, lam, mu, precision=FP64):
        self._chk(self._L.fedd_assembly_linear_elasticity_xdim(self.h, lam, mu, precision))

    def assembly_nonlinear_elasticity(self, neo_hooke_model, E, nu, precision=FP64):
        self._chk(self._L.fedd_assembly_nonlinear_elasticity(self.h, neo_hooke_model, E, nu, precision))

    def elasticity_fp64_fixups(self):
        """Chunks the last FP32-local elasticity assembly re-assembled in FP64."""
        return self.elasticity_fp64_counts()[0]

    def elasticity_fp64_counts(self):
        """(chunks re-assembled in FP64, (node, element) pairs of flagged elements computed
        in FP64) of the last FP32-local elasticity assembly."""
        n, p = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self._L.fedd_elasticity_fp64_fixups(self.h, ctypes.byref(n), ctypes.byref(p)))
        return n.value, p.value

    def assembly_nonlinear_elasticity_tangent(self, neo_hooke_model, E, nu, precision=FP64, tangent=0):
        self._chk(self._L.fedd_assembly_nonlinear_elasticity_tangent(self.h, neo_hooke_model, E, nu, precision,
                                                                     tangent))

    def assembly_nonlinear_elasticity_efield(self, neo_hooke_model, E_elem, nu, precision=FP64):
        E_elem = np.ascontiguousarray(E_elem, np.float64)
        self._chk(self._L.fedd_assembly_nonlinear_elasticity_efield(self.h, neo_hooke_model, _p(E_elem, _dp),
                                                                    E_elem.size, nu, precision))

    def dirichlet_set(self, dofs):
        d = np.ascontiguousarray(dofs, np.int32)
        self._chk(self._L.fedd_dirichlet_set(self.h, _p(d, _ip), d.size))

    def apply_dirichlet_symmetric(self):
        self._chk(self._L.fedd_apply_dirichlet_symmetric(self.h))

    def dirichlet_values(self, g):
        g = np.ascontiguousarray(g, np.float64).ravel()
        self._chk(self._L.fedd_dirichlet_values(self.h, _p(g, _dp), g.size))

    def dirichlet_set_system(self):
        self._chk(self._L.fedd_dirichlet_set_system(self.h))

    def dirichlet_vector(self, mode, out, sub=-1):
        self._chk(self._L.fedd_dirichlet_vector(self.h, mode, out, sub))

    def traction_load(self, axis, value, t):
        t = np.ascontiguousarray(t, np.float64)
        self._chk(self._L.fedd_traction_load(self.h, axis, value, _p(t, _dp)))

    def surface_set(self, faces, flags):
        faces = np.ascontiguousarray(faces, np.int32)
        flags = np.ascontiguousarray(flags, np.int32)
        if faces.ndim != 2 or faces.shape[1] != 6 or flags.shape != (faces.shape[0],):
            raise ValueError("faces (n, 6) and flags (n,)")
        self._chk(self._L.fedd_surface_set(self.h, _p(faces, _ip), _p(flags, _ip), faces.shape[0]))

    def assembly_surface_pressure(self, pressure, with_tangent=False, out=VEC_FEXT):
        """pressure[flag]: FE::assembly(Nonlinear)SurfaceIntegralExternal into vector `out`
        (and Kext added to block A with with_tangent)."""
        pr = np.ascontiguousarray(pressure, np.float64)
        self._chk(self._L.fedd_assembly_surface_pressure(self.h, _p(pr, _dp), len(pr), 1 if with_tangent else 0, out))

    def newton_pressure(self, pressure):
        """Follower pressure per flag for newton_elasticity (empty: none)."""
        pr = np.ascontiguousarray(pressure, np.float64)
        self._chk(self._L.fedd_newton_pressure(self.h, _p(pr, _dp) if len(pr) else None, len(pr)))

    def set_preconditioner(self, kind):
        self._chk(self._L.fedd_set_preconditioner(self.h, kind))

    def linear_solve(self, rtol, maxit):
        it, rr = ctypes.c_int(), ctypes.c_double()
        self._chk(self._L.fedd_linear_solve(self.h, rtol, maxit, ctypes.byref(it), ctypes.byref(rr)))
        return it.value, rr.value

    def ns_linear_solve(self, rtol, restart=50, maxit=2000):
        it, rr = ctypes.c_int(), ctypes.c_double()
        self._chk(self._L.fedd_ns_linear_solve(self.h, rtol, restart, maxit, ctypes.byref(it), ctypes.byref(rr)))
        return it.value, rr.value

    def newton_navier_stokes(self, viscosity, density, linearization=1, criterion=0, rel_tol=1e-10, max_its=20,
                             lin_rtol=1e-10, restart=60, lin_maxit=5000, history_cap=256):
        P = NsNewtonParams(viscosity, density, linearization, criterion, rel_tol, max_its, lin_rtol, restart,
                           lin_maxit)
        st = NsNewtonStats()
        h = np.zeros(history_cap)
        self._chk(self._L.fedd_newton_navier_stokes(self.h, ctypes.byref(P), ctypes.byref(st), _p(h, _dp),
                                                    history_cap))
        return st.as_dict(), h[:st.history_len].copy()

    def linear_solve_noalloc(self, d_b, d_x, n, rtol, maxit):
        """fedd_linear_solve_noalloc on caller-owned device vectors (integer addresses)."""
        it, rr = ctypes.c_int(), ctypes.c_double()
        self._chk(self._L.fedd_linear_solve_noalloc(self.h, ctypes.c_void_p(d_b), ctypes.c_void_p(d_x), n, rtol,
                                                    maxit, ctypes.byref(it), ctypes.byref(rr)))
        return it.value, rr.value

    def newton_elasticity(self, model=1, E=1000.0, nu=0.3, load_steps=5, abs_tol=1e-8, max_newton=30,
                          lin_rtol=1e-10, lin_maxit=20000, tangent=0, precision=FP64, stall_ratio=0.5,
                          history_cap=512, collect_samples=False, preconditioner=0, mg_reuse=False, criterion=0,
                          rel_tol=0.0):
        P = NewtonParams(model, E, nu, load_steps, abs_tol, max_newton, lin_rtol, lin_maxit, tangent, precision,
                         stall_ratio, 1 if collect_samples else 0, preconditioner, 1 if mg_reuse else 0, criterion,
                         rel_tol)
        st = NewtonStats()
        hist = np.zeros(history_cap)
        self._chk(self._L.fedd_newton_elasticity(self.h, ctypes.byref(P), ctypes.byref(st), _p(hist, _dp),
                                                 history_cap))
        return st.as_dict(), hist[:st.history_len]

    def newton_step(self):
        s = ctypes.c_int()
        self._chk(self._L.fedd_newton_step(self.h, ctypes.byref(s)))
        return s.value

    # learned tangent surrogate
    def mlp_init(self, model=1, E=1000.0, nu=0.3, seed=0):
        self._chk(self._L.fedd_mlp_init(self.h, model, E, nu, seed))

    def mlp_add_samples_random(self, n, sigma=0.1, seed=2):
        self._chk(self._L.fedd_mlp_add_samples_random(self.h, n, sigma, seed))

    def mlp_add_samples_qp(self):
        self._chk(self._L.fedd_mlp_add_samples_qp(self.h))

    def mlp_num_samples(self):
        n = ctypes.c_int64()
        self._chk(self._L.fedd_mlp_num_samples(self.h, ctypes.byref(n)))
        return n.value

    def mlp_train(self, steps, batch=4096, lr=1e-3, seed=0, history=True):
        h = np.zeros(max(steps, 1), np.float32)
        self._chk(self._L.fedd_mlp_train(self.h, steps, batch, lr, seed, _p(h, _fp) if history else None))
        return h[:steps]

    def mlp_params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        self._chk(self._L.fedd_mlp_params(self.h, _p(p, _fp)))
        return p

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (MLP_PARAMS,)
        self._chk(self._L.fedd_mlp_set_params(self.h, _p(p, _fp)))

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], 21), np.float32)
        self._chk(self._L.fedd_mlp_forward(self.h, _p(x, _fp), x.shape[0], _p(y, _fp)))
        return y

    def mlp_gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self._L.fedd_mlp_gradient(self.h, _p(x, _fp), _p(t, _fp), x.shape[0], _p(g, _fp),
                                            _p(loss, _fp)))
        return g, float(loss[0])

    def mlp_norm(self):
        cc0 = np.zeros(21, np.float32)
        so, si = ctypes.c_float(), ctypes.c_double()
        self._chk(self._L.fedd_mlp_norm(self.h, _p(cc0, _fp), ctypes.byref(so), ctypes.byref(si)))
        return cc0, so.value, si.value

    # graph capture
    def capture_begin(self):
        self._chk(self._L.fedd_capture_begin(self.h))

    def capture_end(self):
        n = ctypes.c_int64()
        self._chk(self._L.fedd_capture_end(self.h, ctypes.byref(n)))
        return n.value

    def graph_launch(self):
        self._chk(self._L.fedd_graph_launch(self.h))

    # timing
    def synchronize(self):
        self._chk(self._L.fedd_synchronize(self.h))

    def stream_copy(self, nbytes, iters):
        """ms per copy of nbytes (read nbytes + write nbytes)."""
        ms = ctypes.c_float()
        self._chk(self._L.fedd_stream_copy(self.h, nbytes, iters, ctypes.byref(ms)))
        return ms.value

    def event_record(self, slot):
        self._chk(self._L.fedd_event_record(self.h, slot))

    def event_elapsed(self, a, b):
        ms = ctypes.c_float()
        self._chk(self._L.fedd_event_elapsed(self.h, a, b, ctypes.byref(ms)))
        return ms.value

    def kernel_timing(self, enable):
        self._chk(self._L.fedd_kernel_timing(self.h, 1 if enable else 0))

    def kernel_time(self, name):
        t, n = ctypes.c_double(), ctypes.c_int64()
        self._chk(self._L.fedd_kernel_time(self.h, name.encode(), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value
