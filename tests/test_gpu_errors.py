"""Error behaviour of the C ABI on the device path.

FEDDLib raises std::logic_error for wrong arguments and call order
(TEUCHOS_TEST_FOR_EXCEPTION, e.g. "Number Elements not the same as number assembleFE
elements." at FE_def.hpp:1140-1141); libfeddgpu returns a negative status with the
message in fedd_last_error, which the binding raises as FeddError.  Every case here is
refused on the host before any kernel runs, and the context stays usable afterwards.
Each test uses its own context so that a refused call cannot leave state behind for
other tests.
"""
import numpy as np
import pytest

import pyoracle as O
from feddlib_amd import native as N
from helpers import parity_ratio

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_STATE, ERR_LIMIT = -1, -3, -4


@pytest.fixture
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def _code(excinfo):
    return excinfo.value.code


def test_mesh_upload_rejects_bad_meshes(ctx):
    pts, _, el = N.mesh_structured(2, 1)
    with pytest.raises(N.FeddError, match="empty mesh") as e:
        ctx.mesh_upload(N.MESH_PRIMARY, pts[:0], el[:0])
    assert _code(e) == ERR_ARG
    bad = el.copy()
    bad[3, 2] = pts.shape[0]  # one past the last node
    with pytest.raises(N.FeddError, match="out of range"):
        ctx.mesh_upload(N.MESH_PRIMARY, pts, bad)
    bad[3, 2] = -1
    with pytest.raises(N.FeddError, match="out of range"):
        ctx.mesh_upload(N.MESH_PRIMARY, pts, bad)
    p2, _, e2 = N.mesh_structured(2, 2)
    with pytest.raises(N.FeddError, match="pressure mesh must be P1"):
        ctx.mesh_upload(N.MESH_PRESSURE, p2, e2)
    with pytest.raises(N.FeddError, match="nodes_per_elem"):
        ctx.mesh_upload(N.MESH_PRIMARY, pts, np.ascontiguousarray(el[:, :3]))


def test_call_order_is_enforced(ctx):
    with pytest.raises(N.FeddError, match="no primary mesh") as e:
        ctx.graph_build(N.GRAPH_SCALAR)
    assert _code(e) == ERR_STATE
    pts, _, el = N.mesh_structured(2, 1)
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    with pytest.raises(N.FeddError, match="needs a scalar graph"):
        ctx.assembly_laplace(N.FP64)
    ctx.graph_build(N.GRAPH_VECTOR)
    with pytest.raises(N.FeddError, match="needs a scalar graph"):  # wrong graph kind
        ctx.assembly_laplace(N.FP64)
    with pytest.raises(N.FeddError, match="no pressure mesh"):
        ctx.graph_build(N.GRAPH_NS)
    # the context still works after the refusals
    ctx.graph_build(N.GRAPH_SCALAR)
    ctx.assembly_laplace(N.FP64)
    rp, ci = ctx.block_graph(N.BLOCK_A)
    ref = O.laplace(1, pts, el)
    assert parity_ratio(ctx.block_values(N.BLOCK_A), ref[2], ref[0]) <= 1.0


def test_navier_stokes_preconditions(ctx):
    p2, _, e2 = N.mesh_structured(2, 2)
    p1, _, e1 = N.mesh_structured(2, 1)
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, np.ascontiguousarray(e1[:-6]))  # one hexahedron fewer
    with pytest.raises(N.FeddError, match="Number Elements not the same as number assembleFE elements.") as e:
        ctx.graph_build(N.GRAPH_NS)
    assert _code(e) == ERR_ARG
    ctx.mesh_upload(N.MESH_PRIMARY, p1, e1)
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
    with pytest.raises(N.FeddError, match="P2 velocity mesh"):
        ctx.graph_build(N.GRAPH_NS)
    ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
    ctx.graph_build(N.GRAPH_NS)
    with pytest.raises(N.FeddError, match="velocity iterate not uploaded"):
        ctx.assembly_navier_stokes(0.01, 1.0, True, N.NS_JACOBIAN, reassemble=False)
    ctx.vector_upload(N.VEC_U, np.zeros(3 * p2.shape[0]))
    with pytest.raises(N.FeddError, match="pressure iterate not uploaded"):
        ctx.assembly_navier_stokes(0.01, 1.0, True, N.NS_JACOBIAN, reassemble=False)
    ctx.vector_upload(N.VEC_P, np.zeros(p1.shape[0]))
    with pytest.raises(N.FeddError, match="unknown assembleMode"):
        ctx.assembly_navier_stokes(0.01, 1.0, True, 7, reassemble=False)
    ctx.assembly_navier_stokes(0.01, 1.0, True, N.NS_JACOBIAN, reassemble=False)  # now accepted


def test_elasticity_parameters_are_checked(ctx):
    pts, _, el = N.mesh_structured(2, 2)
    el = np.ascontiguousarray(el)
    N.correct_orientation(pts, el)
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    ctx.graph_build(N.GRAPH_VECTOR)
    with pytest.raises(N.FeddError, match="displacement not uploaded"):
        ctx.assembly_nonlinear_elasticity(1, 1000.0, 0.3)
    ctx.vector_upload(N.VEC_U, np.zeros(3 * pts.shape[0]))
    for E, nu in ((0.0, 0.3), (-5.0, 0.3), (1000.0, 0.5), (1000.0, -1.0), (float("nan"), 0.3)):
        with pytest.raises(N.FeddError, match="E must be > 0") as e:
            ctx.assembly_nonlinear_elasticity(1, E, nu)
        assert _code(e) == ERR_ARG
    with pytest.raises(N.FeddError, match="E must be > 0 in every element"):
        Ev = np.full(el.shape[0], 1000.0)
        Ev[5] = 0.0
        ctx.assembly_nonlinear_elasticity_efield(1, Ev, 0.3)
    with pytest.raises(N.FeddError, match="E field length"):
        ctx.assembly_nonlinear_elasticity_efield(1, np.full(el.shape[0] - 1, 1000.0), 0.3)
    with pytest.raises(N.FeddError, match="mu > 0"):
        ctx.assembly_linear_elasticity_xdim(1.0, 0.0)
    ctx.assembly_nonlinear_elasticity(1, 1000.0, 0.3)  # accepted after the refusals


def test_neighbour_limit(ctx):
    """A node with more than 255 neighbour nodes exceeds the one-byte column slots of the
    gather plan: FEDD_ERR_LIMIT, and the graph is not left half-built."""
    rng = np.random.default_rng(3)
    ntet = 100
    pts = np.vstack([np.zeros((1, 3)), rng.normal(size=(3 * ntet, 3))])
    el = np.hstack([np.zeros((ntet, 1), np.int32),
                    np.arange(1, 3 * ntet + 1, dtype=np.int32).reshape(ntet, 3)])
    ctx.mesh_upload(N.MESH_PRIMARY, pts, el)
    with pytest.raises(N.FeddError, match="more than 255 neighbour nodes") as e:
        ctx.graph_build(N.GRAPH_SCALAR)
    assert _code(e) == ERR_LIMIT
    with pytest.raises(N.FeddError, match="needs a scalar graph"):
        ctx.assembly_laplace(N.FP64)


def test_captured_graph_dropped_when_buffers_move(ctx):
    """A captured graph holds the device addresses of the mesh, graph and vector buffers.
    Rebuilding the graph, replacing the mesh or resizing a vector frees them, so the
    library destroys the captured graph and fedd_graph_launch refuses with
    FEDD_ERR_STATE instead of replaying kernels on freed memory."""
    from feddlib_amd import workloads as W
    p2, e2, p1, e1, u, p = W.ns_problem(2, 0)

    def setup():
        ctx.mesh_upload(N.MESH_PRIMARY, p2, e2)
        ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)
        ctx.graph_build(N.GRAPH_NS)
        ctx.vector_upload(N.VEC_U, u)
        ctx.vector_upload(N.VEC_P, p)
        ctx.assembly_navier_stokes(0.01, 1.0, True, N.NS_JACOBIAN, reassemble=False)

    def capture():
        ctx.capture_begin()
        ctx.assembly_navier_stokes(0.01, 1.0, True, N.NS_JACOBIAN, reassemble=False)
        ctx.capture_end()
        ctx.graph_launch()
        ctx.synchronize()

    setup()
    capture()
    ref = ctx.block_values(N.BLOCK_A)
    ctx.graph_build(N.GRAPH_NS)  # frees every block the graph writes
    with pytest.raises(N.FeddError, match="rebuilt after the capture") as e:
        ctx.graph_launch()
    assert _code(e) == ERR_STATE
    ctx.vector_upload(N.VEC_U, u)
    ctx.vector_upload(N.VEC_P, p)
    capture()
    ctx.mesh_upload(N.MESH_PRESSURE, p1, e1)  # a new mesh drops the graph too
    with pytest.raises(N.FeddError, match="mesh was replaced") as e:
        ctx.graph_launch()
    assert _code(e) == ERR_STATE
    setup()
    capture()
    rp, _ = ctx.block_graph(N.BLOCK_A)
    assert parity_ratio(ctx.block_values(N.BLOCK_A), ref, rp) <= 1.0
    # a same-length upload keeps the buffers (and the graph); a replay reads the new iterate
    ctx.vector_upload(N.VEC_U, 2.0 * u)
    ctx.graph_launch()
    ctx.synchronize()
    assert parity_ratio(ctx.block_values(N.BLOCK_A), ref, rp) > 1e3
