"""The per-ion envelope: mass and seed per ion, the summary contract and
the birth-splat outcome.  Restates ion_gym/physics/ion_envelope.py:26-223
and the birth refusals of ion_gym/physics/raster2d.py:615-646 and
ion_gym/physics/birth_domain.py:20-38 (same names, same messages' facts,
same fates)."""
from collections import namedtuple

import numpy as np

PerIon = namedtuple("PerIon", ["mz", "seed"])
REQUIRED_SUMMARY_KEYS = ("kind", "tof", "mz", "x_end", "y_end", "z_end")
UNKNOWN_FATE = -1
FATE_NAME = {0: "impact / exit", 1: "boundary exit", 2: "timeout",
             3: "bounding plane", 4: "transporter max_passes",
             5: "station impact plane", 6: "station detect",
             UNKNOWN_FATE: "unknown (no fate recorded)"}
DOMAIN_EDGE_TOL = 1e-9


class BirthInMetal(ValueError):
    """An ion whose birth lies inside metal (raster2d.py:615-621)."""


class BirthOutsideDomain(BirthInMetal):
    """An ion whose birth lies outside the solved domain."""


def fate_name(kind):
    return FATE_NAME.get(int(kind), f"unknown fate {int(kind)}")


def mz_of(mz_list, n_ions, i):
    """Mass of ion i: contiguous blocks, mz_list[(i // n_ions) % n_mz]
    (ion_envelope.py:82-92 with sim_build.mz_of)."""
    return float(mz_list[(int(i) // max(int(n_ions), 1)) % len(mz_list)])


def per_ion(mz_list, n_ions, i, seed_base):
    return PerIon(mz=mz_of(mz_list, n_ions, i), seed=int(seed_base) + int(i))


def make_summary(*, kind, tof, mz, x_end, y_end, z_end, **extras):
    """Summary with the six required keys (ion_envelope.py:162-181)."""
    req = dict(kind=kind, tof=tof, mz=mz, x_end=x_end, y_end=y_end,
               z_end=z_end)
    for k, v in req.items():
        if v is None:
            raise ValueError(f"make_summary: required key '{k}' is None")
    s = dict(kind=int(kind), tof=float(tof), mz=float(mz),
             x_end=float(x_end), y_end=float(y_end), z_end=float(z_end))
    s.update(extras)
    return s


def birth_splat(err, *, col_names, birth, mz, label):
    """One-row record at birth and a fate-0 (in metal) or fate-1 (outside
    the domain) summary (ion_envelope.py:184-223)."""
    x, y, z, vx, vy, vz, tob = (float(v) for v in birth[:7])
    outside = isinstance(err, BirthOutsideDomain)
    print(f"[flight] WARNING: {label} was born "
          + ("outside the solved domain and leaves it at birth"
             if outside else "inside metal and splats at birth")
          + f" -- {err}", flush=True)
    traj = np.full((1, len(col_names)), np.nan)
    for name, v in zip(("t", "x", "y", "z", "vx", "vy", "vz"),
                       (tob, x, y, z, vx, vy, vz)):
        if name in col_names:
            traj[0, list(col_names).index(name)] = v
    flag = (dict(born_outside_domain=True) if outside
            else dict(born_in_metal=True))
    summary = make_summary(kind=1 if outside else 0, tof=tob, mz=mz,
                           x_end=x, y_end=y, z_end=z, n_col=None,
                           note=str(err), **flag)
    return traj, summary


def rz_birth_refusal(ele, h, x, r, i):
    """The r-z birth checks (build_rz.py:679-690): outside nodes 0..n-1
    (birth_domain.py:27-38), then the nearest-node metal rule
    (raster2d.py:624-646).  Returns None or the exception to record."""
    gx, gr = x / h, r / h
    nz, nr = ele.shape
    for g, n, ax in ((gx, nz, "x"), (gr, nr, "r")):
        if not (-DOMAIN_EDGE_TOL <= g <= n - 1 + DOMAIN_EDGE_TOL):
            return BirthOutsideDomain(
                f"ion {i} birth at (x={x:.3f}, r={r:.3f}) mm lies outside "
                f"the solved domain: {ax} is at node {g:.3f}, and the domain "
                f"spans nodes 0 to {n - 1} (pitch {h} mm)")
    gi, gj = int(np.floor(gx + 0.5)), int(np.floor(gr + 0.5))
    if 0 <= gi < nz and 0 <= gj < nr and ele[gi, gj] > 0:
        return BirthInMetal(
            f"ion {i} birth at (x={x:.3f}, r={r:.3f}) mm lies inside the "
            f"effective metal of electrode {int(ele[gi, gj])} (nearest-node "
            f"rule, node pitch {h} mm)")
    return None


def xyz_birth_refusal(metal, h, x, y, z, i):
    """The 3-D birth checks (tracer3d.py:192-230): outside nodes 0..n-1 of
    any axis in node units x / h (birth_domain.py:27-38), then the
    all-eight-corners metal rule of the kernels' impact test
    (tracer3d.py:174-189).  Returns None or the exception to record."""
    m = np.asarray(metal)
    nx, ny, nz = m.shape
    inv_h = 1.0 / float(h)
    g = (x * inv_h, y * inv_h, z * inv_h)
    for gv, n, ax in zip(g, (nx, ny, nz), "xyz"):
        if not (-DOMAIN_EDGE_TOL <= gv <= n - 1 + DOMAIN_EDGE_TOL):
            return BirthOutsideDomain(
                f"ion {i} birth at (x={x:.3f}, y={y:.3f}, z={z:.3f}) mm lies "
                f"outside the solved domain: {ax} is at node {gv:.3f}, and "
                f"the domain spans nodes 0 to {n - 1} (pitch {h} mm)")
    if all(0.0 <= gv <= n - 1 for gv, n in zip(g, (nx, ny, nz))):
        a, b, c = (min(int(gv), n - 2) for gv, n in zip(g, (nx, ny, nz)))
        if np.all(m[a:a + 2, b:b + 2, c:c + 2]):
            return BirthInMetal(
                f"ion {i} birth at (x={x:.3f}, y={y:.3f}, z={z:.3f}) mm lies "
                f"inside the effective metal (all-eight-corner voxel rule, "
                f"node pitch {h} mm)")
    return None


def planar_birth_refusal(ele, h, x, y, i):
    """The planar birth checks (build_planar.py:1588-1605): outside nodes
    0..n-1 of x or y, then the nearest-node metal rule on the planar
    labels.  Returns None or the exception to record."""
    gx, gy = x / h, y / h
    nx, ny = ele.shape
    for g, n, ax in ((gx, nx, "x"), (gy, ny, "y")):
        if not (-DOMAIN_EDGE_TOL <= g <= n - 1 + DOMAIN_EDGE_TOL):
            return BirthOutsideDomain(
                f"ion {i} birth at ({x:.3f}, {y:.3f}) mm lies outside the "
                f"solved domain: {ax} is at node {g:.3f}, and the domain "
                f"spans nodes 0 to {n - 1} (pitch {h} mm)")
    gi, gj = int(np.floor(gx + 0.5)), int(np.floor(gy + 0.5))
    if 0 <= gi < nx and 0 <= gj < ny and ele[gi, gj] > 0:
        return BirthInMetal(
            f"ion {i} birth at ({x:.3f}, {y:.3f}) mm lies inside the "
            f"effective metal of electrode {int(ele[gi, gj])} (nearest-node "
            f"rule, node pitch {h} mm)")
    return None


def call_rng_preserved(fn, *args, **kw):
    """ion_gym/physics/numba_rng.py:20-28 saves and restores numba's
    per-thread generator around a kernel call.  The engine keeps no
    global generator (every ion's draws come from its own counter-based
    stream), so there is nothing to save: the call is made as is."""
    return fn(*args, **kw)
