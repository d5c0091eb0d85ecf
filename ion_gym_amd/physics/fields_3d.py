"""3-D field from a potential: E = -grad(phi) with the electrode-aware
one-sided differences at metal surfaces.

Restates ion_gym/physics/tracer3d.py:54-117 (build_field_aware_3d and its
_fa3d_correct pass): np.gradient with the scalar node spacing, then at
each metal node with vacuum on exactly one side along an axis that axis
derivative becomes the one-sided difference into the vacuum; vacuum nodes
are untouched.  The per-node loop of the reference is written here as
whole-array selections with the same expressions, so the values are the
same.
"""
import numpy as np


def _axis_correct(d, phi, ele, h, axis):
    n = phi.shape[axis]
    vac = ~ele
    vp = np.zeros_like(ele)
    vm = np.zeros_like(ele)
    sl = [slice(None)] * 3

    def S(a, b):
        s = list(sl)
        s[axis] = slice(a, b)
        return tuple(s)

    vp[S(0, n - 1)] = vac[S(1, n)]
    vm[S(1, n)] = vac[S(0, n - 1)]
    fwd = ele & vp & ~vm
    bwd = ele & vm & ~vp
    df = np.zeros_like(phi)
    db = np.zeros_like(phi)
    df[S(0, n - 1)] = (phi[S(1, n)] - phi[S(0, n - 1)]) / h
    db[S(1, n)] = (phi[S(1, n)] - phi[S(0, n - 1)]) / h
    return np.where(fwd, df, np.where(bwd, db, d))


def build_field_aware_3d(phi, ele, h_mm, electrode_aware=True):
    """(Ex, Ey, Ez) in V/mm on the node grid (tracer3d.py:93-117)."""
    phi = np.ascontiguousarray(phi, np.float64)
    ele = np.asarray(ele).astype(bool)
    d = np.gradient(phi, h_mm)
    dx, dy, dz = (np.ascontiguousarray(v) for v in d)
    if electrode_aware:
        dx = _axis_correct(dx, phi, ele, float(h_mm), 0)
        dy = _axis_correct(dy, phi, ele, float(h_mm), 1)
        dz = _axis_correct(dz, phi, ele, float(h_mm), 2)
    return -dx, -dy, -dz
