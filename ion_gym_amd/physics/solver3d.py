"""3-D Laplace basis solves on the engine (ion_gym/physics/solver3d.py,
ion_gym/physics/multigrid3d.py).

The reference's native solvers, with their call signatures and refusals:

* ``solve3d`` -- red-black SOR of the 7-point star with fractional legs
  (``stencil='ghost_linear'``, weight 1/theta on a short leg, or
  ``'sw'``, Shortley-Weller), even reflection at every outer face, until a
  sweep pair moves no node by ``tol`` (solver3d.py:250-293);
* ``solve3d_vmg`` -- the V-cycle multigrid that smooths with the same
  sweep at omega 1, corrects with a plain 7-point error equation solved to
  5 % on 2x-coarsened grids, projects onto verified mirror symmetries and
  stops when a whole cycle moved no node by ``tol``
  (multigrid3d.py:602-731);
* ``solve3d_mg`` -- the cascade: the mask and values coarsened with a
  rescue of thin metal, SOR from the coarsest grid up, each grid started
  from the trilinear prolongation of the one below, the fine grid solved
  by the same SOR as ``solve3d`` (multigrid3d.py:84-230);
* ``solve_bases`` / ``solve_bases_mg`` -- one solve per electrode at
  ``v_basis`` with every other electrode grounded
  (solver3d.py:296-320, multigrid3d.py:233-302).

Every sweep, residual, transfer and reduction runs on the GPU
(``ig_solve3d``); the host only sets the problem up and reads back the
field.  A solve that cannot meet its tolerance raises
``SolveNotConverged``, as the reference does.  Leg fractions come in as an
array (theta, (6, nx, ny, nz), order -x, +x, -y, +y, -z, +z); computing
them from analytic geometry scenes is upstream of this engine.
"""
import ctypes as C

import numpy as np

from .. import _lib as L

_ENGINE = None


def _engine(engine=None):
    global _ENGINE
    if engine is not None:
        return engine
    if _ENGINE is None:
        from ..engine import Engine
        _ENGINE = Engine(0)
    return _ENGINE


class SolveNotConverged(RuntimeError):
    """The iteration hit its limit (or stagnated) without meeting tol
    (solver3d.py:233-247)."""


def optimal_omega(shape, spacing=None):
    """Young's optimal SOR factor for the 7-point Laplacian on a
    rectangular grid; axes of fewer than 2 nodes carry no coupling and are
    left out (solver3d.py:201-226)."""
    if spacing is None:
        dims = [(int(n), 1.0) for n in shape if int(n) >= 2]
    else:
        dims = [(int(n), 1.0 / (float(s) * float(s)))
                for n, s in zip(shape, spacing) if int(n) >= 2]
    if not dims:
        return 1.0
    num = sum(wi * np.cos(np.pi / (ni + 1)) for ni, wi in dims)
    rho = num / sum(wi for _, wi in dims)
    rho = min(rho, 1.0 - 1e-15)
    return 2.0 / (1.0 + np.sqrt(1.0 - rho * rho))


def _check_stencil(stencil):
    if stencil not in L.STENCILS:
        raise ValueError(
            f"unknown stencil {stencil!r}; known stencils are "
            f"'ghost_linear' (linear ghost-value fractional surface) and "
            f"'sw' (Shortley-Weller).")


def _run(fixed, val, theta, phi0, stencil, tol, *, method, omega=0.0,
         max_sweeps=0, max_cycles=0, nu1=0, nu2=0, sym_axes=0, min_coarse=0,
         max_levels=0, level_tol_factor=0.0, engine=None):
    fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
    if fixed.ndim != 3:
        raise ValueError(f"fixed must be 3-D, got shape {fixed.shape}")
    shape = fixed.shape
    val = np.ascontiguousarray(np.broadcast_to(val, shape), dtype=np.float64)
    th = None
    if theta is not None:
        th = np.ascontiguousarray(theta, dtype=np.float32)
        if th.shape != (6,) + shape:
            raise ValueError(f"theta must be (6, {shape[0]}, {shape[1]}, "
                             f"{shape[2]}), got {th.shape}")
    p0 = None
    if phi0 is not None:
        p0 = np.ascontiguousarray(np.broadcast_to(phi0, shape),
                                  dtype=np.float64)
    d = L.Solve3dDesc()
    for a in range(3):
        d.n[a] = shape[a]
    d.fixed = L.ptr(fixed, L._u8p)
    d.val = L.ptr(val)
    d.theta = L.ptr(th, L._fp)
    d.phi0 = L.ptr(p0)
    d.method = method
    d.stencil = L.STENCILS[stencil]
    d.tol = float(tol)
    d.omega = float(omega)
    d.max_sweeps = int(max_sweeps)
    d.max_cycles = int(max_cycles)
    d.nu1, d.nu2 = int(nu1), int(nu2)
    d.sym_axes = int(sym_axes)
    d.min_coarse = int(min_coarse)
    d.max_levels = int(max_levels)
    d.level_tol_factor = float(level_tol_factor)
    phi = np.empty(shape, np.float64)
    res = L.Solve3dResult()
    eng = _engine(engine)
    L.check(L.lib().ig_solve3d(eng.h, C.byref(d), L.ptr(phi), C.byref(res)),
            "ig_solve3d")
    return phi, int(res.sweeps), float(res.delta), int(res.status), \
        int(res.level)


def _sor_cap(shape, max_sweeps, d, tol):
    nx, ny, nz = shape
    return SolveNotConverged(
        f"SOR hit its cap of {max_sweeps} sweeps on a {nx}x{ny}x{nz} "
        f"grid without reaching tol: final delta {d:.3e} vs tol "
        f"{tol:.3e} ({d / max(tol, 1e-30):.0f}x too loose). SOR stalls "
        f"on long thin domains; solve this geometry with "
        f"solve_bases_mg, or raise max_sweeps if you have reason to "
        f"think it will actually converge.")


def solve3d(fixed, val, theta=None, mirror=(False, False, False), tol=1e-6,
            max_sweeps=60000, omega=None, phi0=None, verbose=False,
            stencil="ghost_linear", engine=None):
    """Red-black SOR to max |update| < tol (solver3d.py:250-293).  Returns
    (phi, sweep pairs, final delta).  ``mirror`` only matters where the
    leg fractions were computed: every outer face reflects evenly
    (solver3d.py:139-171)."""
    _check_stencil(stencil)
    shape = np.shape(fixed)
    if omega is None:
        omega = optimal_omega(shape)
    phi, sw, d, st, _ = _run(fixed, val, theta, phi0, stencil, tol,
                             method=L.SOLVE_SOR, omega=omega,
                             max_sweeps=max_sweeps, engine=engine)
    if st != 0:
        raise _sor_cap(shape, max_sweeps, d, tol)
    if verbose:
        print(f"  SOR: {sw} sweeps, max|d| {d:.3e}")
    return phi, sw, d


def symmetric_axes(fixed, val, theta=None):
    """Axes along which the problem is exactly mirror-invariant (fixed and
    val equal their flips; only without leg fractions), as solve3d_vmg
    verifies before projecting (multigrid3d.py:643-650)."""
    if theta is not None:
        return []
    fixed = np.asarray(fixed, bool)
    val = np.broadcast_to(np.asarray(val, np.float64), fixed.shape)
    out = []
    for a in range(3):
        if fixed.shape[a] < 3:
            continue
        if (np.array_equal(fixed, np.flip(fixed, axis=a))
                and np.array_equal(val, np.flip(val, axis=a))):
            out.append(a)
    return out


def solve3d_mg(fixed, val, theta=None, mirror=(False, False, False),
               tol=1e-6, max_sweeps=60000, verbose=False,
               stencil="ghost_linear", min_coarse=9, max_levels=8,
               level_tol_factor=30.0, engine=None):
    """Cascadic multigrid (multigrid3d.py:175-230): (fixed, val) coarsened
    while every axis exceeds ``min_coarse`` (at most ``max_levels`` grids),
    the coarsest solved by SOR from zeros, each finer grid from the
    prolongated coarser solution to ``tol * level_tol_factor``, the fine
    grid with its fractions to ``tol``, so the result is the plain SOR
    fixed point.  Returns (phi, fine sweep pairs, final delta); raises
    SolveNotConverged, naming the grid, when any grid's SOR runs out of
    ``max_sweeps``."""
    _check_stencil(stencil)
    shape = np.shape(fixed)
    phi, sw, d, st, lv = _run(fixed, val, theta, None, stencil, tol,
                              method=L.SOLVE_CASCADE,
                              omega=optimal_omega(shape),
                              max_sweeps=max_sweeps, min_coarse=min_coarse,
                              max_levels=max_levels,
                              level_tol_factor=level_tol_factor,
                              engine=engine)
    if st != 0:
        g = tuple(shape)
        for _ in range(lv):
            g = tuple((n + 1) // 2 for n in g)
        lt = tol if lv == 0 else tol * level_tol_factor
        raise _sor_cap(g, max_sweeps, d, lt)
    if verbose:
        print(f"[mg] fine: {sw} sweeps, max|d| {d:.3e}")
    return phi, sw, d


def solve3d_vmg(fixed, val, theta=None, mirror=(False, False, False),
                tol=1e-6, max_cycles=60, nu1=3, nu2=3, verbose=False,
                stencil="ghost_linear", phi0=None, accept_stagnation=False,
                engine=None):
    """V-cycle multigrid with the validated fine sweep
    (multigrid3d.py:602-731).  Returns (phi, fine sweep pairs, final cycle
    delta); raises SolveNotConverged when the cycle change stops falling
    (unless ``accept_stagnation``) or max_cycles run out."""
    _check_stencil(stencil)
    fixed = np.asarray(fixed, bool)
    sym = symmetric_axes(fixed, val, theta)
    phi, sw, d, st, _ = _run(fixed, val, theta, phi0, stencil, tol,
                             method=L.SOLVE_VCYCLE, max_cycles=max_cycles,
                             nu1=nu1, nu2=nu2,
                             sym_axes=sum(1 << a for a in sym), engine=engine)
    if verbose:
        print(f"[vmg] {sw} fine sweeps, cycle delta {d:.3e}")
    if st == 0:
        return phi, sw, d
    if st == 2:
        if accept_stagnation:
            return phi, sw, d
        raise SolveNotConverged(
            f"V-cycle STAGNATED: delta {d:.3e} V stopped improving before "
            f"reaching tol {tol:.1e}. A stalled iterate is not a converged "
            f"field. Loosen tol toward the achieved delta, or pass "
            f"accept_stagnation=True to knowingly take the iterate.")
    raise SolveNotConverged(
        f"V-cycle exhausted max_cycles={max_cycles} at delta {d:.3e} V "
        f"without reaching tol {tol:.1e}. An iteration-capped iterate is "
        f"not a converged field; raise max_cycles or loosen tol.")


def _bases(masks, solve, v_basis, only, verbose):
    masks = dict(masks)
    if not masks:
        return {}
    metal = np.zeros(np.shape(next(iter(masks.values()))), bool)
    for m in masks.values():
        metal |= np.asarray(m, bool)
    out = {}
    items = sorted(masks.items())
    for j, (i, m) in enumerate(items, start=1):
        if only is not None and i not in only:
            continue    # every electrode is a boundary of every basis
        val = np.zeros(metal.shape)
        val[np.asarray(m, bool)] = v_basis
        phi, sw, d = solve(metal, val)
        if verbose:
            print(f"  electrode {j} of {len(items)} (basis {i}): {sw} "
                  f"sweeps, delta {d:.2e} V", flush=True)
        out[i] = phi
    return out


def _no_scene(scene):
    if scene is not None:
        raise NotImplementedError(
            "leg fractions from a geometry scene are computed upstream of "
            "the engine: pass theta=(6, nx, ny, nz) float32 fractions")


def solve_bases(masks, scene=None, mirror=(False, False, False), v_basis=1e4,
                tol=1e-6, verbose=False, stencil="ghost_linear", omega=None,
                only=None, theta=None, engine=None):
    """{index: mask} -> {index: phi} with electrode i at v_basis and the
    rest of the metal at 0, each by SOR (solver3d.py:296-320)."""
    _no_scene(scene)
    return _bases(masks, lambda f, v: solve3d(
        f, v, theta=theta, mirror=mirror, tol=tol, stencil=stencil,
        omega=omega, engine=engine), v_basis, only, verbose)


def solve_bases_mg(masks, scene=None, mirror=(False, False, False),
                   v_basis=1e4, tol=1e-6, verbose=False,
                   stencil="ghost_linear", theta=None, only=None,
                   engine=None):
    """The same bases by the V-cycle (multigrid3d.py:233-302)."""
    _no_scene(scene)
    return _bases(masks, lambda f, v: solve3d_vmg(
        f, v, theta=theta, mirror=mirror, tol=tol, stencil=stencil,
        engine=engine), v_basis, only, verbose)
