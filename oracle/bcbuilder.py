"""CPU restatement of FEDDLib's Dirichlet handling (BCBuilder), on CRS arrays.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's Dirichlet
entry points.  The product never imports it.

Follows feddlib/core/General/BCBuilder_def.hpp:
  setSystem          (:590-618)  every block row that carries Dirichlet conditions: each
                                 existing block of that row goes through setDirichletBC,
                                 with isDiagonalBlock = (blockRow == blockCol);
  setDirichletBC     (:621-654)  for every flagged node (full "Dirichlet" type: all dofs of
                                 the node) setLocalRowOne on the diagonal block, else
                                 setLocalRowZero;
  setLocalRowOne     (:657-689)  the row's stored values become 0, except the first stored
                                 column whose global id equals the row's global dof -> 1;
  setLocalRowZero    (:692-704)  the row's stored values become 0;
  setRHS             (:93-160)   rhs[dof] = g[dof];
  setBCMinusVector   (:234-300)  out[dof] = g[dof] - sub[dof];
  setVectorMinusBC   (:350-420)  out[dof] = sub[dof] - g[dof].
On one rank the local row index is the global dof (repeated map = unique map), so the row
of dof d is row d of the block.  Dirichlet dofs are given directly (the flagged nodes'
dofs), one list per block row.
"""
import numpy as np


def set_local_row_one(rowptr, colind, vals, row):
    """BCBuilder_def.hpp:657-689 on one row of a CRS block (in place)."""
    set_one = False
    for k in range(int(rowptr[row]), int(rowptr[row + 1])):
        if not set_one and int(colind[k]) == int(row):
            vals[k] = 1.0
            set_one = True
        else:
            vals[k] = 0.0


def set_local_row_zero(rowptr, colind, vals, row):
    """BCBuilder_def.hpp:692-704 on one row of a CRS block (in place)."""
    for k in range(int(rowptr[row]), int(rowptr[row + 1])):
        vals[k] = 0.0


def set_system(blocks, dirichlet_rows):
    """BCBuilder_def.hpp:590-654.  blocks[(i, j)] = (rowptr, colind, vals) for the existing
    blocks of the block matrix; dirichlet_rows[i] = constrained dofs of block row i (absent
    or empty: no conditions on that row).  Returns new value arrays, keyed like blocks."""
    out = {key: np.array(b[2], dtype=np.float64, copy=True) for key, b in blocks.items()}
    for (bi, bj), (rp, ci, _) in blocks.items():
        rows = dirichlet_rows.get(bi, ())
        for r in rows:
            if bi == bj:
                set_local_row_one(rp, ci, out[(bi, bj)], r)
            else:
                set_local_row_zero(rp, ci, out[(bi, bj)], r)
    return out


def bc_vector(mode, out, dofs, g, sub=None):
    """setRHS (mode 0), setBCMinusVector (1), setVectorMinusBC (2) at the dofs `dofs`."""
    out = np.array(out, dtype=np.float64, copy=True)
    for d in dofs:
        if mode == 0:
            out[d] = g[d]
        elif mode == 1:
            out[d] = g[d] - sub[d]
        elif mode == 2:
            out[d] = sub[d] - g[d]
        else:
            raise ValueError("mode")
    return out
