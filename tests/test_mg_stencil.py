"""The multigrid hierarchy's coarse sparsity patterns (host logic, no GPU).

The library builds every coarse level's adjacency directly as the Kuhn P1 stencil (a node
and its 14 neighbours +-e_S).  This checks that claim against the Galerkin product itself:
the pattern of P^T A P, with A the P2 node adjacency of the structured mesh and P the
midpoint prolongation on the Kuhn lattice (a node with odd coordinates on axis set S is
the midpoint of the edge e_S), repeated down the P1 levels.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from feddlib_amd import native as N

KUHN = [(0, 0, 0)] + [tuple(s * np.array(v)) for v in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1),
                                                       (0, 1, 1), (1, 1, 1)] for s in (1, -1)]


def prolongation(Pf):
    """Fine lattice Pf (odd) -> coarse lattice (Pf + 1) / 2: midpoint rule along Kuhn edges."""
    Pc = (Pf + 1) // 2
    i = np.arange(Pf ** 3)
    x = np.stack([i % Pf, (i // Pf) % Pf, i // (Pf * Pf)], 1)
    o = x & 1
    lo, hi = (x - o) >> 1, (x + o) >> 1
    cid = lambda y: y[:, 0] + Pc * (y[:, 1] + Pc * y[:, 2])
    even = o.sum(1) == 0
    rows = np.concatenate([i[even], i[~even], i[~even]])
    cols = np.concatenate([cid(lo)[even], cid(lo)[~even], cid(hi)[~even]])
    vals = np.concatenate([np.ones(even.sum()), np.full(2 * (~even).sum(), 0.5)])
    return sp.csr_matrix((vals, (rows, cols)), shape=(Pf ** 3, Pc ** 3)), Pc


def kuhn_stencil(P):
    i = np.arange(P ** 3)
    x = np.stack([i % P, (i // P) % P, i // (P * P)], 1)
    rows, cols = [], []
    for off in KUHN:
        y = x + np.array(off)
        ok = np.all((y >= 0) & (y < P), 1)
        rows.append(i[ok])
        cols.append((y[:, 0] + P * (y[:, 1] + P * y[:, 2]))[ok])
    return sp.csr_matrix((np.ones(sum(len(r) for r in rows)), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(P ** 3, P ** 3))


@pytest.mark.parametrize("M", [4, 6, 8])
def test_galerkin_pattern_is_the_kuhn_stencil(M):
    pts, _, el = N.mesh_structured(M, 2)
    n = len(pts)
    r = np.repeat(el, 10, axis=1).ravel()
    c = np.tile(el, (1, 10)).ravel()
    A = sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    A.data[:] = 1.0
    P = 2 * M + 1
    levels = 0
    while P >= 5 and (P - 1) % 2 == 0:
        Pr, P = prolongation(P)
        A = (Pr.T @ A @ Pr).tocsr()
        A.data[:] = 1.0
        S = kuhn_stencil(P)
        assert (A != S).nnz == 0, (M, P)
        levels += 1
    assert levels >= 2
