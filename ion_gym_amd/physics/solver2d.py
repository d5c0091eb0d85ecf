"""2-D finite-difference Laplace solves (ion_gym/physics/solver2d.py).

``laplace_factor`` / ``solve_laplace`` keep the reference's signatures and
refusals and solve on the engine (``ig_solve2d``): the five-point system
of ``assemble_laplace`` (solver2d.py:31-124) -- planar interior centre -4
with unit neighbours; cylindrical (u = r) radial neighbours 1 +- h/(2r),
on the axis centre -6 with the first off-axis node 4; at an open edge the
missing neighbour's coefficient folded onto the opposite neighbour
('mirror') or the centre ('ghost_linear'); fixed nodes Dirichlet -- by
red-black SOR on the device to a tolerance far below the field's scale,
where the reference factors it with a sparse LU.

``solve_rz`` is the compact host solve of the same cylindrical system
that builds the benchmark configurations' bases on any machine.
"""
import ctypes as C

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


_KINDS = ("ghost_linear", "mirror")
_EDGES = ("z0", "z1", "u0", "u1")


def _ghost_of(edge_ghost):
    """The per-edge fold kinds, with the reference's refusals
    (solver2d.py:57-73)."""
    if isinstance(edge_ghost, str):
        if edge_ghost not in _KINDS:
            raise ValueError(f"edge_ghost must be one of {_KINDS} or a "
                             f"{{edge: kind}} mapping, got {edge_ghost!r}")
        return dict.fromkeys(_EDGES, edge_ghost)
    bad_e = [e for e in edge_ghost if e not in _EDGES]
    bad_k = [k for k in edge_ghost.values() if k not in _KINDS]
    if bad_e or bad_k:
        raise ValueError(
            f"edge_ghost mapping: unknown edge(s) {bad_e} / "
            f"kind(s) {bad_k}; edges are {_EDGES}, kinds {_KINDS}")
    out = dict.fromkeys(_EDGES, "mirror")
    out.update(edge_ghost)
    return out


def sor_omega(shape, neumann_edges=()):
    """The SOR factor for a 2-D grid: 2 / (1 + sqrt(1 - rho^2)) with rho
    the Jacobi radius of the Dirichlet problem, an axis with an open edge
    counted at twice its length (the fold mirrors the grid across it)."""
    nz, nu = shape
    ez = 2 if ("z0" in neumann_edges or "z1" in neumann_edges) else 1
    eu = 2 if ("u0" in neumann_edges or "u1" in neumann_edges) else 1
    rho = 0.5 * (np.cos(np.pi / (ez * nz + 1)) + np.cos(np.pi / (eu * nu + 1)))
    rho = min(rho, 1.0 - 1e-15)
    return 2.0 / (1.0 + np.sqrt(1.0 - rho * rho))


def laplace_factor(fixed, h, symmetry="planar",
                   neumann_edges=("z0", "z1", "u0", "u1"),
                   edge_ghost="mirror", dtype=np.float64, *, rel_tol=1e-12,
                   omega=None, max_sweeps=200000, engine=None):
    """Check the system once and return ``solve_rhs(val)`` -> phi
    (nz, nu) float64 (solver2d.py:127-161).  Each call runs red-black SOR
    on the device until a pair moves no node by ``rel_tol`` times the
    largest |val| on the fixed nodes (1 V when all are 0), from zeros.
    Over-relaxation near omega = 2 amplifies rounding: on the 1201 x 201
    funnel grid a pair's change stops falling near 2.5e-13 of the field
    scale, so the default stops at 1e-12, where the field is within 1e-9
    of the scale of the LU solution.
    ``dtype`` is accepted for the reference's signature; the device
    always solves in float64.  A free node with a missing neighbour on a
    closed edge raises ValueError, as the reference's assembly does; an
    unknown ``symmetry`` is refused (the reference would treat it as
    planar)."""
    from .. import _lib as L
    from .solver3d import SolveNotConverged, _engine
    if symmetry not in ("planar", "cylindrical"):
        raise ValueError(f"symmetry must be 'planar' or 'cylindrical', "
                         f"got {symmetry!r}")
    ghost = _ghost_of(edge_ghost)
    fixed = np.ascontiguousarray(fixed, dtype=bool)
    if fixed.ndim != 2 or min(fixed.shape) < 2:
        raise ValueError(f"fixed must be 2-D with both axes of at least 2 "
                         f"nodes, got shape {fixed.shape}")
    nz, nu = fixed.shape
    cyl = symmetry == "cylindrical"
    open_ = tuple(e for e in neumann_edges if not (cyl and e == "u0"))
    free = ~fixed
    edge_free = {"z0": free[0].any(), "z1": free[-1].any(),
                 "u0": free[:, 0].any() and not cyl,
                 "u1": free[:, -1].any()}
    for e in _EDGES:
        if edge_free[e] and e not in open_:
            raise ValueError(
                f"free node with missing neighbour on closed edge {e}")
    if omega is None:
        omega = sor_omega(fixed.shape, open_)
    mask = fixed.astype(np.uint8)
    nbits = sum(L.EDGE_BITS[e] for e in open_)
    gbits = sum(L.EDGE_BITS[e] for e in _EDGES if ghost[e] == "ghost_linear")

    def solve_rhs(val):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(val, np.float64),
                                                 fixed.shape))
        scale = float(np.abs(v[fixed]).max()) if fixed.any() else 0.0
        d = L.Solve2dDesc()
        d.nz, d.nu = nz, nu
        d.fixed = L.ptr(mask, L._u8p)
        d.val = L.ptr(v)
        d.phi0 = None
        d.h = float(h)
        d.symmetry = L.SYM_CYLINDRICAL if cyl else L.SYM_PLANAR
        d.neumann = nbits
        d.ghost_linear = gbits
        d.tol = rel_tol * (scale if scale > 0.0 else 1.0)
        d.omega = float(omega)
        d.max_sweeps = int(max_sweeps)
        phi = np.empty((nz, nu), np.float64)
        res = L.Solve3dResult()
        eng = _engine(engine)
        L.check(L.lib().ig_solve2d(eng.h, C.byref(d), L.ptr(phi),
                                   C.byref(res)), "ig_solve2d")
        if res.status != 0:
            raise SolveNotConverged(
                f"2-D SOR hit its cap of {max_sweeps} sweeps on a {nz}x{nu} "
                f"grid: final delta {res.delta:.3e} vs tol {d.tol:.3e}")
        solve_rhs.sweeps = int(res.sweeps)
        return phi

    solve_rhs.sweeps = 0
    return solve_rhs


def solve_laplace(fixed, val, h, symmetry="planar",
                  neumann_edges=("z0", "z1", "u0", "u1"),
                  edge_ghost="mirror", *, engine=None):
    """Single right-hand side (solver2d.py:164-182)."""
    return laplace_factor(fixed, h, symmetry, neumann_edges, edge_ghost,
                          engine=engine)(val)
