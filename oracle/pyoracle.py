"""ctypes wrapper of oracle/liboracle.so and oracle/_ref/libacegen_ref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

_O = None
_R = None


def lib():
    global _O
    if _O is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        L = ctypes.CDLL(path)
        sig = {
            "orc_mesh_p1": (None, [ctypes.c_int, _dp, _ip, _ip]),
            "orc_mesh_p2": (None, [ctypes.c_int, _dp, _ip, _ip]),
            "orc_correct_orientation": (ctypes.c_int64, [ctypes.c_int, _dp, ctypes.c_int64, _ip]),
            "orc_crs_create": (_vp, [ctypes.c_int64]),
            "orc_crs_destroy": (None, [_vp]),
            "orc_crs_fill_complete": (None, [_vp]),
            "orc_crs_nnz": (ctypes.c_int64, [_vp]),
            "orc_crs_nrows": (ctypes.c_int64, [_vp]),
            "orc_crs_export": (None, [_vp, _lp, _ip, _dp]),
            "orc_quadrature": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
            "orc_phi": (None, [ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "orc_grad_phi": (None, [ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "orc_degree_func": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
            "orc_degree_type": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
            "orc_degree_pair": (ctypes.c_int, [ctypes.c_int] * 5),
            "orc_assembly_laplace": (_vp, [ctypes.c_int, _dp, ctypes.c_int64, _ip, ctypes.c_int64]),
            "orc_assembly_laplace_assfe": (_vp, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int64, _ip,
                                                 ctypes.c_int64]),
            "orc_assembly_rhs": (None, [ctypes.c_int, _dp, ctypes.c_int64, _ip, ctypes.c_int64,
                                        ctypes.c_double, _dp]),
            "orc_rhs_quadrature_points": (ctypes.c_int, [ctypes.c_int, _dp, _ip, ctypes.c_int64, _dp]),
            "orc_assembly_rhs_points": (None, [ctypes.c_int, _dp, ctypes.c_int64, _ip, ctypes.c_int64, _dp,
                                               ctypes.c_int, _dp]),
            "orc_ns_element": (None, [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                      _dp, _dp, _dp]),
            "orc_assembly_ns": (None, [_dp, ctypes.c_int64, _ip, _ip, ctypes.c_int64, ctypes.c_int64, _dp,
                                       _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, _dp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                       ctypes.POINTER(_vp), ctypes.POINTER(_vp), _dp, _dp]),
            "orc_time_ns_jacobian": (ctypes.c_double, [_dp, ctypes.c_int64, _ip, _ip, ctypes.c_int64,
                                                       ctypes.c_int64, _dp, _dp, ctypes.c_double,
                                                       ctypes.c_double]),
            "orc_time_ns_jacobian_mt": (ctypes.c_double, [_dp, ctypes.c_int64, _ip, _ip, ctypes.c_int64,
                                                          ctypes.c_int64, _dp, _dp, ctypes.c_double,
                                                          ctypes.c_double, ctypes.c_int]),
            "orc_nh_skr": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]),
            "orc_set_element_mode": (None, [ctypes.c_int]),
            "orc_pressure_triangle": (None, [ctypes.c_double, _dp, _dp, _dp, _dp]),
            "orc_structured_surfaces": (ctypes.c_int64, [_dp, _ip, ctypes.c_int64, _dp, _ip, _ip, _ip]),
            "orc_assembly_surface_pressure": (_vp, [_dp, ctypes.c_int64, _ip, _ip, ctypes.c_int64, _dp, _dp,
                                                    ctypes.c_int, _dp, ctypes.c_int]),
            "orc_assembly_elasticity": (_vp, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp,
                                              ctypes.c_int64, _ip, ctypes.c_int64, _dp, _dp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _O = L
    return _O


def ref_lib():
    """The reference's own AceGen elements, compiled by oracle/Makefile (None if absent)."""
    global _R
    if _R is None:
        path = os.path.join(HERE, "_ref", "libacegen_ref.so")
        if not os.path.exists(path):
            return None
        _R = ctypes.CDLL(path)
    return _R


def P(a, t=_dp):
    return a.ctypes.data_as(t)


class _element_mode:
    mode = 0

    def __enter__(self):
        lib().orc_set_element_mode(self.mode)
        return self

    def __exit__(self, *exc):
        lib().orc_set_element_mode(0)
        return False


class element_round32(_element_mode):
    """Context: every element matrix value is rounded to FP32 before it is summed in FP64
    (matrix assemblies only; the oracle's residual vectors stay FP64)."""
    mode = 1


class element_scale(_element_mode):
    """Context: matrix assemblies sum, for each entry, the largest |value| of each
    contributing element block (Navier-Stokes A, B^T, B; the whole element matrix
    otherwise): sum_e max |K_e|, the scale of the rounding error of element entries
    computed in FP32."""
    mode = 2


def mesh(M, fe):
    L = lib()
    n = (M + 1) ** 3 if fe == 1 else (2 * M + 1) ** 3
    E, k = 6 * M ** 3, (4 if fe == 1 else 10)
    pts = np.zeros((n, 3))
    bc = np.zeros(n, np.int32)
    el = np.zeros((E, k), np.int32)
    (L.orc_mesh_p1 if fe == 1 else L.orc_mesh_p2)(M, P(pts), P(bc, _ip), P(el, _ip))
    return pts, bc, el


def correct_orientation(pts, el):
    return lib().orc_correct_orientation(el.shape[1], P(np.ascontiguousarray(pts)), el.shape[0], P(el, _ip))


def crs_take(h):
    """Export and free an oracle CRS handle -> (rowptr, colind, vals)."""
    L = lib()
    if not h:
        return None
    nr = L.orc_crs_nrows(h)
    nnz = L.orc_crs_nnz(h)
    rp = np.zeros(nr + 1, np.int64)
    ci = np.zeros(max(nnz, 1), np.int32)
    v = np.zeros(max(nnz, 1))
    L.orc_crs_export(h, P(rp, _lp), P(ci, _ip), P(v))
    L.orc_crs_destroy(h)
    return rp, ci[:nnz], v[:nnz]


def laplace(fe, pts, el):
    return crs_take(lib().orc_assembly_laplace(fe, P(pts), pts.shape[0], P(el, _ip), el.shape[0]))


def laplace_assfe(fe, dofs, pts, el):
    return crs_take(lib().orc_assembly_laplace_assfe(fe, dofs, P(pts), pts.shape[0], P(el, _ip), el.shape[0]))


def rhs(fe, pts, el, f):
    b = np.zeros(pts.shape[0])
    lib().orc_assembly_rhs(fe, P(pts), pts.shape[0], P(el, _ip), el.shape[0], f, P(b))
    return b


def rhs_quadrature_points(fe, pts, el):
    """Physical quadrature points of FE::assemblyRHS's rule per element: (E, nq, 3)."""
    E = el.shape[0]
    xq = np.zeros(E * 5 * 3)
    nq = lib().orc_rhs_quadrature_points(fe, P(pts), P(el, _ip), E, P(xq))
    return xq[:E * nq * 3].reshape(E, nq, 3)


def rhs_points(fe, pts, el, fq):
    """FE::assemblyRHS from source values fq (E, nq) ("Scalar") or (E, nq, 3) ("Vector")."""
    fq = np.ascontiguousarray(fq, np.float64)
    ncomp = 1 if fq.ndim == 2 else fq.shape[2]
    b = np.zeros(ncomp * pts.shape[0])
    lib().orc_assembly_rhs_points(fe, P(pts), pts.shape[0], P(el, _ip), el.shape[0], P(fq), ncomp, P(b))
    return b


def quadrature(deg):
    p = np.zeros((24, 3))
    w = np.zeros(24)
    n = lib().orc_quadrature(deg, P(p), P(w))
    return p[:n], w[:n]


def ns_element(xl, sol, nu, rho, newton, coeff=(1, 1, 1, 1)):
    jac = np.zeros((34, 34))
    fp = np.zeros((34, 34))
    r = np.zeros(34)
    c = np.array(coeff, np.float64)
    lib().orc_ns_element(P(np.ascontiguousarray(xl, np.float64)), P(np.ascontiguousarray(sol, np.float64)),
                         nu, rho, 1 if newton else 0, P(c), P(jac), P(fp), P(r))
    return jac, fp, r


def ns_assembly(p2, e2, e1, np1, u, p, nu, rho, newton, mode, reassemble, coeff=(1, 1, 1, 1)):
    """FE::assemblyNavierStokes; returns dict of blocks (rowptr, colind, vals) or residuals."""
    L = lib()
    hs = [_vp() for _ in range(4)]
    ru = np.zeros(3 * p2.shape[0])
    rp = np.zeros(np1)
    c = np.array(coeff, np.float64)
    uu = np.ascontiguousarray(u, np.float64).ravel()
    pp = np.ascontiguousarray(p, np.float64)
    L.orc_assembly_ns(P(p2), p2.shape[0], P(e2, _ip), P(e1, _ip), e2.shape[0], np1, P(uu), P(pp), nu, rho,
                      1 if newton else 0, mode, 1 if reassemble else 0, P(c),
                      *[ctypes.byref(h) for h in hs], P(ru), P(rp))
    out = {"A": crs_take(hs[0].value), "BT": crs_take(hs[1].value), "B": crs_take(hs[2].value),
           "C": crs_take(hs[3].value)}
    if mode == 2:
        out["res_u"] = ru
        out["res_p"] = rp
    return out


def time_ns_jacobian_mt(p2, e2, e1, np1, u, p, nu, rho, nthreads):
    """Wall time of the whole element loop on `nthreads` host threads (partitions + merge)."""
    uu = np.ascontiguousarray(u, np.float64).ravel()
    return lib().orc_time_ns_jacobian_mt(P(p2), p2.shape[0], P(e2, _ip), P(e1, _ip), e2.shape[0], np1, P(uu),
                                         P(np.ascontiguousarray(p, np.float64)), nu, rho, nthreads)


def time_ns_jacobian(p2, e2, e1, np1, u, p, nu, rho, nelem):
    uu = np.ascontiguousarray(u, np.float64).ravel()
    return lib().orc_time_ns_jacobian(P(p2), p2.shape[0], P(e2, _ip), P(e1, _ip), nelem, np1, P(uu),
                                      P(np.ascontiguousarray(p, np.float64)), nu, rho)


def nh_skr(model, E, nu, xl, ul):
    s = np.zeros(900)
    p = np.zeros(30)
    lib().orc_nh_skr(model, E, nu, P(np.ascontiguousarray(xl, np.float64).ravel()),
                     P(np.ascontiguousarray(ul, np.float64).ravel()), P(s), P(p))
    return s.reshape(30, 30, order="F"), p


def ref_skr(model, E, nu, xl, ul):
    """The reference AceGen skr1/skr2/skr3 (model 0/1/2); None when _ref is not built."""
    R = ref_lib()
    if R is None:
        return None
    f = getattr(R, ["skr1", "skr2", "skr3"][model])
    v = np.zeros(1100)
    d = np.array([E, nu], np.float64)
    ul = np.ascontiguousarray(ul, np.float64).ravel().copy()
    xl = np.ascontiguousarray(xl, np.float64).ravel().copy()
    ul0 = np.zeros(30)
    s = np.zeros(900)
    p = np.zeros(30)
    ht = np.zeros(10)
    hp = np.zeros(10)
    f(P(v), P(d), P(ul), P(ul0), P(xl), P(s), P(p), P(ht), P(hp))
    return s.reshape(30, 30, order="F"), p


def pressure_triangle(p, X, u, tangent=True):
    """AceGen PressureTriangle3D6 restated: load f (6, 3) and Kext (18, 18) = -df/du."""
    f = np.zeros(18)
    K = np.zeros(324) if tangent else None
    lib().orc_pressure_triangle(p, P(np.ascontiguousarray(X, np.float64).ravel()),
                                P(np.ascontiguousarray(u, np.float64).ravel()), P(f),
                                P(K) if tangent else None)
    return f.reshape(6, 3), (K.reshape(18, 18) if tangent else None)


def structured_surfaces(pts, el, box=(0.0, 0.0, 0.0, 1.0, 1.0, 1.0)):
    """MeshStructured::buildSurfaces(3, "P2"): faces (n, 6), flags (n,), owning element (n,)."""
    pts = np.ascontiguousarray(pts, np.float64)
    el = np.ascontiguousarray(el, np.int32)
    b = np.asarray(box, np.float64)
    n = lib().orc_structured_surfaces(P(pts), P(el, _ip), el.shape[0], P(b), None, None, None)
    faces = np.zeros((n, 6), np.int32)
    flags = np.zeros(n, np.int32)
    owner = np.zeros(n, np.int32)
    lib().orc_structured_surfaces(P(pts), P(el, _ip), el.shape[0], P(b), P(faces, _ip), P(flags, _ip),
                                  P(owner, _ip))
    return faces, flags, owner


def surface_pressure(pts, faces, flags, u, pressure, tangent=False):
    """The surface-load loop of FE::assembly(Nonlinear)SurfaceIntegralExternal: load vector
    f (3 npts) and, with tangent, Kext as (rowptr, colind, vals)."""
    pts = np.ascontiguousarray(pts, np.float64)
    f = np.zeros(3 * pts.shape[0])
    pr = np.ascontiguousarray(pressure, np.float64)
    uu = None if u is None else np.ascontiguousarray(u, np.float64).ravel()
    h = lib().orc_assembly_surface_pressure(P(pts), pts.shape[0], P(np.ascontiguousarray(faces, np.int32), _ip),
                                            P(np.ascontiguousarray(flags, np.int32), _ip), len(flags),
                                            None if uu is None else P(uu), P(pr), len(pr), P(f),
                                            1 if tangent else 0)
    return f, (crs_take(h) if tangent else None)


def elasticity(model, E, nu, pts, el, u):
    r = np.zeros(3 * pts.shape[0])
    uu = np.ascontiguousarray(u, np.float64).ravel()
    h = lib().orc_assembly_elasticity(model, E, nu, P(pts), pts.shape[0], P(el, _ip), el.shape[0], P(uu), P(r))
    return crs_take(h), r
