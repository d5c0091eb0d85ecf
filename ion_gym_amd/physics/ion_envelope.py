"""The per-ion envelope: mass and seed per ion, the run's seed policy, the
summary contract and the birth-splat outcome.  Restates
ion_gym/physics/ion_envelope.py:26-223 and the birth refusals of
ion_gym/physics/raster2d.py:615-646 and ion_gym/physics/birth_domain.py:20-38
(same names, same facts in the messages, same fates)."""
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
BASE_CHANNELS = ("t", "x", "y", "z", "vx", "vy", "vz")


class BirthInMetal(ValueError):
    """An ion whose birth lies inside metal (raster2d.py:615-621)."""


class BirthOutsideDomain(BirthInMetal):
    """An ion whose birth lies outside the solved domain."""


def fate_name(kind):
    return FATE_NAME.get(int(kind), f"unknown fate {int(kind)}")


def resolve_run_seed(spec):
    """The run's seed (ion_envelope.py:56-79): source.seed when stated;
    for source.seed = None one 31-bit draw of OS entropy, made once per
    source, kept on it and printed, so births and every per-ion stream of
    the run share it and the run can be repeated."""
    from ..births import resolve_run_seed as _source_seed
    return _source_seed(spec.source)


def run_seed_base(spec, verbose=True):
    """The base of the per-ion seeds (ion_envelope.py:26-50): a stated
    source.seed is the fixed (common random numbers) policy, announced when
    ``verbose``; None is the random policy of resolve_run_seed."""
    raw = getattr(spec.source, "seed", 0)
    if raw is None:
        return resolve_run_seed(spec)
    if verbose:
        print(f"[seed] policy FIXED (CRN): source.seed = {int(raw)}; reruns "
              f"are bit-identical -- set source.seed: null for fresh entropy "
              f"per run")
    return int(raw)


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
    """The outcome of a refused birth (ion_envelope.py:184-223): the other
    ions fly on, this one ends where and when it was born.  In metal it
    splats (fate 0, ``born_in_metal``); outside the solved domain it has
    left it (fate 1, ``born_outside_domain``).  The record is one row with
    the birth state in the base channels and NaN in every optional one;
    n_col is None (the ion never flew).  A warning naming the ion and the
    reason is printed."""
    state = dict(zip(BASE_CHANNELS[1:], (float(v) for v in birth[:6])))
    state["t"] = float(birth[6])
    outside = isinstance(err, BirthOutsideDomain)
    what = ("outside the solved domain and leaves it at birth" if outside
            else "inside metal and splats at birth")
    print(f"[flight] WARNING: {label} was born {what} -- {err}", flush=True)
    row = [state.get(c, np.nan) for c in col_names]
    traj = np.asarray([row], np.float64)
    flag = "born_outside_domain" if outside else "born_in_metal"
    summary = make_summary(kind=1 if outside else 0, tof=state["t"], mz=mz,
                           x_end=state["x"], y_end=state["y"],
                           z_end=state["z"], n_col=None, note=str(err),
                           **{flag: True})
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
