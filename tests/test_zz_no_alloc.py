"""fedd_linear_solve_noalloc against fedd_linear_solve (block-Jacobi CG, preconditioner 0).

Both solve the Dirichlet-constrained linear-elasticity system (FE::assemblyLinElasXDim on
a P2 cube, clamped base) from x = 0 with the same kernels; the no-allocation variant keeps
its work space in the library's per-GPU scratch array and reads/writes caller-owned device
vectors (here the context's VEC_B and VEC_FEXT buffers, addressed by device pointer).
Two calls in a row, the second on a smaller mesh, so the second call runs on a scratch
array that still holds the first call's (longer) vectors.

Tolerance: the CG dot products are summed with device atomics in no fixed order, so the
two solves agree to rounding, not bit for bit.  Each must meet rtol 1e-12 on its own; the
two solutions must agree to 1e-9 relative (max norm) and the iteration counts within 2.
"""
import numpy as np
import pytest

from feddlib_amd import native as N

pytestmark = pytest.mark.gpu


def _elasticity_system(ctx, M, seed):
    pts, _, el = N.mesh_structured(M, 2)
    el = np.ascontiguousarray(el)
    N.correct_orientation(pts, el)
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    ctx.assembly_linear_elasticity_xdim(576.9230769230769, 384.61538461538464)  # E=1000, nu=0.3
    base = np.nonzero(np.abs(pts[:, 2]) < 1e-12)[0]
    dofs = np.sort(np.concatenate([3 * base + d for d in range(3)])).astype(np.int32)
    ctx.dirichlet_set(dofs)
    ctx.apply_dirichlet_symmetric()
    n = 3 * pts.shape[0]
    b = np.random.default_rng(seed).normal(size=n)
    b[dofs] = 0.0
    return n, b


def _solve_both(ctx, M, seed):
    n, b = _elasticity_system(ctx, M, seed)
    print("no-alloc solve: M=%d n=%d seed=%d" % (M, n, seed), flush=True)
    ctx.set_preconditioner(0)
    ctx.vector_upload(N.VEC_B, b)
    it_ref, rel_ref = ctx.linear_solve(1e-12, 20000)
    x_ref = ctx.vector_download(N.VEC_X, n)
    # caller-owned output buffer: the context's FEXT vector, poisoned first
    ctx.vector_upload(N.VEC_FEXT, np.full(n, np.nan))
    d_b, nb = ctx.vector_device_ptr(N.VEC_B)
    d_x, nx = ctx.vector_device_ptr(N.VEC_FEXT)
    assert nb == n and nx == n
    it, rel = ctx.linear_solve_noalloc(d_b, d_x, n, 1e-12, 20000)
    x = ctx.vector_download(N.VEC_FEXT, n)
    assert np.all(np.isfinite(x))
    assert rel <= 1e-12 and rel_ref <= 1e-12
    assert abs(it - it_ref) <= 2, (it, it_ref)
    err = np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref))
    assert err <= 1e-9, err
    # the input vector is read only
    assert np.array_equal(ctx.vector_download(N.VEC_B, n), b)
    return n


def test_noalloc_solve_matches_linear_solve_twice():
    with N.Context(0) as ctx:
        n1 = _solve_both(ctx, 4, 0)
        n2 = _solve_both(ctx, 3, 1)  # smaller system on the scratch the first call left behind
        assert n2 < n1


def test_noalloc_solve_limits():
    with N.Context(0) as ctx:
        n, b = _elasticity_system(ctx, 2, 2)
        ctx.vector_upload(N.VEC_B, b)
        ctx.vector_upload(N.VEC_FEXT, np.zeros(n))
        d_b, _ = ctx.vector_device_ptr(N.VEC_B)
        d_x, _ = ctx.vector_device_ptr(N.VEC_FEXT)
        with pytest.raises(N.FeddError, match="rows of block A"):
            ctx.linear_solve_noalloc(d_b, d_x, n - 3, 1e-10, 100)
        with pytest.raises(N.FeddError):
            ctx.linear_solve_noalloc(0, d_x, n, 1e-10, 100)
        # the context stays usable
        it, rel = ctx.linear_solve_noalloc(d_b, d_x, n, 1e-10, 1000)
        assert rel <= 1e-10
