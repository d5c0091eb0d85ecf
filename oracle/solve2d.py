"""CPU restatement of ion_gym's 2-D Laplace solve
(ion_gym/physics/solver2d.py) -- TEST INFRASTRUCTURE ONLY: the checker
the engine's ``ig_solve2d`` is compared against; the product package never
imports it.

``assemble`` restates ``assemble_laplace`` (solver2d.py:31-124): the
five-point system with fixed rows the identity, planar centre -4 and unit
neighbours, cylindrical radial neighbours 1 +- h/(2r) and on the axis
centre -6 with the first off-axis node 4, a missing neighbour folded onto
the opposite one ('mirror') or the centre ('ghost_linear') at an open
edge, refused on a closed one.  ``solve`` factors it with scipy's sparse
LU as ``laplace_factor`` does (solver2d.py:127-161).  Pinned in
tests/test_solver2d_oracle.py by fields the discrete system reproduces
exactly (r^2 - 2 z^2 with the axis row, x^2 - z^2, fields linear in z
across open edges of both kinds) and against the independent r-z
assembly the benchmark bases use.
"""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.linalg import splu

EDGES = ("z0", "z1", "u0", "u1")


def assemble(fixed, h, symmetry="planar", neumann_edges=EDGES,
             edge_ghost="mirror"):
    fixed = np.asarray(fixed, bool)
    ghost = (dict.fromkeys(EDGES, edge_ghost) if isinstance(edge_ghost, str)
             else {**dict.fromkeys(EDGES, "mirror"), **edge_ghost})
    nz, nu = fixed.shape
    cyl = symmetry == "cylindrical"
    if cyl:
        neumann_edges = tuple(e for e in neumann_edges if e != "u0")
    idx = np.arange(nz * nu).reshape(nz, nu)
    fn = idx[fixed]
    rows, cols, data = [fn], [fn], [np.ones(fn.size)]
    fi, fj = np.where(~fixed)
    k = idx[fi, fj]
    if cyl:
        axis = fj == 0
        r = np.where(axis, 1.0, fj * h)
        c_up = np.where(axis, 4.0, 1.0 + h / (2.0 * r))
        c_dn = np.where(axis, 0.0, 1.0 - h / (2.0 * r))
        centre = np.where(axis, -6.0, -4.0)
    else:
        c_up = np.ones(k.size)
        c_dn = np.ones(k.size)
        centre = np.full(k.size, -4.0)
    ones = np.ones(k.size)
    rows.append(k)
    cols.append(k)
    data.append(centre)
    for coef, exists, ti, tj, mi, mj, edge in (
            (ones, fi > 0, fi - 1, fj, fi + 1, fj, "z0"),
            (ones, fi < nz - 1, fi + 1, fj, fi - 1, fj, "z1"),
            (c_dn, fj > 0, fi, fj - 1, fi, fj + 1, "u0"),
            (c_up, fj < nu - 1, fi, fj + 1, fi, fj - 1, "u1")):
        e = exists & (coef != 0.0)
        rows.append(k[e])
        cols.append(idx[ti[e], tj[e]])
        data.append(coef[e])
        m = ~exists & (coef != 0.0)
        if not m.any():
            continue
        if edge not in neumann_edges:
            raise ValueError(
                f"free node with missing neighbour on closed edge {edge}")
        rows.append(k[m])
        cols.append(idx[mi[m], mj[m]] if ghost[edge] == "mirror" else k[m])
        data.append(coef[m])
    n = nz * nu
    A = coo_matrix((np.concatenate(data),
                    (np.concatenate(rows), np.concatenate(cols))),
                   shape=(n, n)).tocsr()
    return A, fn


def solve(fixed, val, h, symmetry="planar", neumann_edges=EDGES,
          edge_ghost="mirror"):
    """The LU solution of the assembled system for one value set."""
    fixed = np.asarray(fixed, bool)
    A, fn = assemble(fixed, h, symmetry, neumann_edges, edge_ghost)
    b = np.zeros(fixed.size)
    b[fn] = np.broadcast_to(np.asarray(val, np.float64), fixed.shape)[fixed]
    return splu(A.tocsc()).solve(b).reshape(fixed.shape), A
