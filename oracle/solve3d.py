"""ctypes binding of the oracle's 3-D Laplace solver
(ig_oracle_solve3d.c).

TEST INFRASTRUCTURE ONLY: the checker the solver tests compare the
engine's ``ig_solve3d`` against; the product package never imports it.
Restates ion_gym/physics/solver3d.py:131-293 and
ion_gym/physics/multigrid3d.py:84-731 (see the C file for the map).
"""
import ctypes as C

import numpy as np

from .oracle import lib as _oracle_lib

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_i64 = C.c_int64
_L = None


def lib():
    global _L
    if _L is None:
        L = _oracle_lib()
        g3 = [_i64, _i64, _i64]
        L.or_sweep3.argtypes = [_dp, _u8p, _fp, *g3, C.c_double, C.c_int,
                                C.c_int]
        L.or_sweep3.restype = C.c_double
        L.or_solve3d.argtypes = [_dp, _u8p, _dp, _fp, *g3, C.c_double, _i64,
                                 C.c_double, C.c_int, C.POINTER(_i64),
                                 C.POINTER(C.c_double)]
        L.or_solve3d.restype = C.c_int
        L.or_residual3.argtypes = [_dp, _u8p, _fp, *g3, C.c_int, _dp]
        L.or_residual3.restype = None
        L.or_sweep_rhs.argtypes = [_dp, _u8p, _dp, *g3, C.c_double, C.c_int]
        L.or_sweep_rhs.restype = C.c_double
        L.or_residual_rhs.argtypes = [_dp, _u8p, _dp, *g3, _dp]
        L.or_residual_rhs.restype = None
        L.or_restrict_fw.argtypes = [_dp, *g3, _dp, *g3]
        L.or_restrict_fw.restype = None
        L.or_prolong_add.argtypes = [_dp, *g3, _dp, _u8p, *g3]
        L.or_prolong_add.restype = None
        L.or_coarsen_any.argtypes = [_u8p, *g3, _u8p, *g3]
        L.or_coarsen_any.restype = None
        L.or_optimal_omega.argtypes = g3
        L.or_optimal_omega.restype = C.c_double
        L.or_solve_err.argtypes = [_u8p, _dp, *g3, C.c_double, C.c_int,
                                   C.c_int, C.c_int, _i64, _dp]
        L.or_solve_err.restype = None
        L.or_solve3d_vmg.argtypes = [_dp, _u8p, _dp, _fp, *g3, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.POINTER(_i64),
                                     C.POINTER(C.c_double)]
        L.or_solve3d_vmg.restype = C.c_int
        L.or_coarsen_mask.argtypes = [_u8p, _dp, *g3, _u8p, _dp, *g3]
        L.or_coarsen_mask.restype = None
        L.or_prolongate.argtypes = [_dp, *g3, _dp, *g3]
        L.or_prolongate.restype = None
        L.or_solve3d_mg.argtypes = [_dp, _u8p, _dp, _fp, *g3, C.c_double,
                                    _i64, C.c_double, C.c_int, C.c_int,
                                    C.c_int, C.c_double, C.POINTER(_i64),
                                    C.POINTER(C.c_double),
                                    C.POINTER(C.c_int)]
        L.or_solve3d_mg.restype = C.c_int
        _L = L
    return _L


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def _prep(fixed, val, theta):
    fixed = np.ascontiguousarray(fixed, np.uint8)
    val = np.ascontiguousarray(np.broadcast_to(val, fixed.shape), np.float64)
    th = None if theta is None else np.ascontiguousarray(theta, np.float32)
    return fixed, val, th


def sweep(phi, fixed, theta, omega, colour, stencil="ghost_linear"):
    """One colour sweep in place; returns its max |update|."""
    assert phi.dtype == np.float64 and phi.flags.c_contiguous
    fixed = np.ascontiguousarray(fixed, np.uint8)
    th = None if theta is None else np.ascontiguousarray(theta, np.float32)
    return lib().or_sweep3(_p(phi), _p(fixed, _u8p), _p(th, _fp), *phi.shape,
                           float(omega), int(colour),
                           int(stencil == "ghost_linear"))


def solve3d(fixed, val, theta=None, tol=1e-6, max_sweeps=60000, omega=None,
            phi0=None, stencil="ghost_linear"):
    """(phi, sweep pairs, delta, status 0 converged / 1 cap reached)."""
    fixed, val, th = _prep(fixed, val, theta)
    phi = (np.zeros(fixed.shape) if phi0 is None
           else np.array(phi0, np.float64, order="C"))
    if omega is None:
        omega = optimal_omega(fixed.shape)
    sw, d = _i64(), C.c_double()
    st = lib().or_solve3d(_p(phi), _p(fixed, _u8p), _p(val), _p(th, _fp),
                          *fixed.shape, float(tol), int(max_sweeps),
                          float(omega), int(stencil == "ghost_linear"),
                          C.byref(sw), C.byref(d))
    return phi, sw.value, d.value, st


def residual(phi, fixed, theta=None, stencil="ghost_linear"):
    out = np.empty_like(phi)
    fixed = np.ascontiguousarray(fixed, np.uint8)
    th = None if theta is None else np.ascontiguousarray(theta, np.float32)
    lib().or_residual3(_p(np.ascontiguousarray(phi)), _p(fixed, _u8p),
                       _p(th, _fp), *phi.shape, int(stencil == "ghost_linear"),
                       _p(out))
    return out


def restrict_fw(fine):
    fine = np.ascontiguousarray(fine, np.float64)
    cs = tuple((n + 1) // 2 for n in fine.shape)
    out = np.empty(cs)
    lib().or_restrict_fw(_p(fine), *fine.shape, _p(out), *cs)
    return out


def prolong_add(coarse, out, fixed):
    coarse = np.ascontiguousarray(coarse, np.float64)
    fixed = np.ascontiguousarray(fixed, np.uint8)
    lib().or_prolong_add(_p(coarse), *coarse.shape, _p(out), _p(fixed, _u8p),
                         *out.shape)
    return out


def coarsen_any(fixed):
    fixed = np.ascontiguousarray(fixed, np.uint8)
    cs = tuple((n + 1) // 2 for n in fixed.shape)
    out = np.empty(cs, np.uint8)
    lib().or_coarsen_any(_p(fixed, _u8p), *fixed.shape, _p(out, _u8p), *cs)
    return out.astype(bool)


def optimal_omega(shape):
    return lib().or_optimal_omega(*[int(n) for n in shape])


def solve3d_vmg(fixed, val, theta=None, tol=1e-6, max_cycles=60, nu1=3,
                nu2=3, stencil="ghost_linear", phi0=None, sym_axes=()):
    """(phi, fine sweep pairs, cycle delta, status 0 converged / 1 cycles
    exhausted / 2 stagnated)."""
    fixed, val, th = _prep(fixed, val, theta)
    phi = (np.zeros(fixed.shape) if phi0 is None
           else np.array(phi0, np.float64, order="C"))
    sw, d = _i64(), C.c_double()
    st = lib().or_solve3d_vmg(_p(phi), _p(fixed, _u8p), _p(val), _p(th, _fp),
                              *fixed.shape, float(tol), int(max_cycles),
                              int(nu1), int(nu2),
                              int(stencil == "ghost_linear"),
                              sum(1 << a for a in sym_axes), C.byref(sw),
                              C.byref(d))
    return phi, sw.value, d.value, st


def coarsen_mask(fixed, val):
    """(fixed, val) of the 2x-coarsened grid (multigrid3d._coarsen_mask)."""
    fixed = np.ascontiguousarray(fixed, np.uint8)
    val = np.ascontiguousarray(np.broadcast_to(val, fixed.shape), np.float64)
    cs = tuple((n + 1) // 2 for n in fixed.shape)
    fc = np.empty(cs, np.uint8)
    vc = np.empty(cs)
    lib().or_coarsen_mask(_p(fixed, _u8p), _p(val), *fixed.shape,
                          _p(fc, _u8p), _p(vc), *cs)
    return fc.astype(bool), vc


def prolongate(coarse, shape):
    """Trilinear prolongation onto a fine grid (multigrid3d._prolongate)."""
    coarse = np.ascontiguousarray(coarse, np.float64)
    out = np.empty(tuple(int(n) for n in shape))
    lib().or_prolongate(_p(coarse), *coarse.shape, _p(out), *out.shape)
    return out


def solve3d_mg(fixed, val, theta=None, tol=1e-6, max_sweeps=60000,
               omega=None, stencil="ghost_linear", min_coarse=9,
               max_levels=8, level_tol_factor=30.0):
    """(phi, fine sweep pairs, delta, status 0 converged / 1 a grid's SOR
    ran out of sweeps, the grid: 0 the fine one, l the l-th coarsening)."""
    fixed, val, th = _prep(fixed, val, theta)
    phi = np.zeros(fixed.shape)
    if omega is None:
        omega = optimal_omega(fixed.shape)
    sw, d, lv = _i64(), C.c_double(), C.c_int()
    st = lib().or_solve3d_mg(_p(phi), _p(fixed, _u8p), _p(val), _p(th, _fp),
                             *fixed.shape, float(tol), int(max_sweeps),
                             float(omega), int(stencil == "ghost_linear"),
                             int(min_coarse), int(max_levels),
                             float(level_tol_factor), C.byref(sw),
                             C.byref(d), C.byref(lv))
    return phi, sw.value, d.value, st, lv.value
