"""Cylindrical (r-z) finite-difference Laplace solve for the benchmark
configurations' bases.

The field solver is upstream of the flight path and not part of the
engine; this is a compact solve on the stencil ion_gym documents
(ion_gym/physics/solver2d.py:1-25): interior centre -4 with z neighbours 1
and radial neighbours 1 +- h/(2r); on the axis centre -6, z neighbours 1
and the first off-axis node 4; open edges Neumann by folding the missing
neighbour onto the opposite one.  Electrode nodes are Dirichlet.  One
sparse LU serves every right-hand side.
"""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.linalg import splu


def solve_rz(labels, values_list, h):
    """labels (nz, nr) int, 0 = free; values_list: list of dicts
    {label: volts}.  Returns one potential (nz, nr) per dict."""
    nz, nr = labels.shape
    N = nz * nr
    idx = np.arange(N).reshape(nz, nr)
    fixed = labels > 0
    rows, cols, vals = [], [], []

    def add(mask, di, dj, coef):
        ii, jj = np.nonzero(mask)
        rows.append(idx[ii, jj])
        cols.append(idx[ii + di, jj + dj])
        vals.append(np.broadcast_to(coef, ii.shape)[...] if np.ndim(coef) == 0
                    else coef[ii, jj])

    free = ~fixed
    I, J = np.meshgrid(np.arange(nz), np.arange(nr), indexing="ij")
    r = J * h
    centre = np.where(J == 0, -6.0, -4.0)
    # z neighbours with Neumann fold at the open ends
    cz_p = np.where(I == nz - 1, 0.0, 1.0) + np.where(I == 0, 1.0, 0.0)
    cz_m = np.where(I == 0, 0.0, 1.0) + np.where(I == nz - 1, 1.0, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        up = np.where(J == 0, 4.0, 1.0 + h / (2.0 * np.where(J == 0, 1, r)))
        dn = np.where(J == 0, 0.0, 1.0 - h / (2.0 * np.where(J == 0, 1, r)))
    # outer radial edge: fold the missing (j+1) onto (j-1)
    edge = J == nr - 1
    dn = np.where(edge, dn + up, dn)
    up = np.where(edge, 0.0, up)
    m = free
    rows.append(idx[m])
    cols.append(idx[m])
    vals.append(centre[m])
    for (di, dj, c) in ((1, 0, cz_p), (-1, 0, cz_m), (0, 1, up), (0, -1, dn)):
        mm = m & (c != 0.0)
        ii, jj = np.nonzero(mm)
        rows.append(idx[ii, jj])
        cols.append(idx[ii + di, jj + dj])
        vals.append(c[ii, jj])
    ii, jj = np.nonzero(fixed)
    rows.append(idx[ii, jj])
    cols.append(idx[ii, jj])
    vals.append(np.ones(len(ii)))
    A = coo_matrix((np.concatenate(vals), (np.concatenate(rows),
                                           np.concatenate(cols))),
                   shape=(N, N)).tocsc()
    lu = splu(A)
    out = []
    for values in values_list:
        b = np.zeros(N)
        lab = labels.ravel()
        for L, v in values.items():
            b[lab == L] = v
        out.append(lu.solve(b).reshape(nz, nr))
    return out
