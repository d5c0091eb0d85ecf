"""Solved per-electrode bases on disk, keyed on the geometry alone
(ion_gym/io/basis_cache.py), in fa_cache entries.

Laplace is linear in the electrode voltages, so a new tune re-weights the
same bases: the key describes what the solver sees and nothing it does
not.  Two layouts, ion_gym's two:

* r-z and planar (basis_cache.py:257-277): arrays ``b{idx}`` float64, one
  per electrode basis, and ``ele`` uint8 labels (0 vacuum, i electrode i),
  under ``cache_key_dict(spec)``: the deck's geometry with every voltage
  and appearance field of the electrodes dropped (dc, rf_groups, color,
  name), the DC-ladder placement held at None (dc_group, dc_index,
  dc_weight), each electrode's basis index made explicit (a reorder
  changes the basis-electrode map), plus the layout's kind and format and
  the key of the solver code.  A load returns the labels as int16 and
  treats a boolean ``ele`` (an older format) as a miss, said out loud.
* 3-D (build_stl3d.py:870-935): arrays ``basis_{idx}`` and ``labels``
  uint8 under a caller's key dict; ``solve_bases_cached`` keys the device
  solver's output on the masks' content, the solve settings and the
  solver code.

The solver-code key hashes this package's solver sources, so a change to
the solver misses every entry it could have changed.
"""
import hashlib
from dataclasses import asdict, is_dataclass
from pathlib import Path

import numpy as np

from . import fa_cache

CACHE_KIND = "planar_bases_v1"
CACHE_FORMAT = 2
CACHE_KIND_3D = "device_bases_3d_v1"
MAX_LABEL = 255                 # ``ele`` / ``labels`` are stored in 8 bits

_VOLTAGE_FIELDS = ("dc", "rf_groups", "color", "name")
_LADDER_FIELDS = ("dc_group", "dc_index", "dc_weight")

_PKG = Path(__file__).resolve().parent.parent
SOLVE_SOURCES = ("physics/solver2d.py", "physics/solver3d.py",
                 "csrc/ig_solve.hip", "csrc/ig_solve.h")


def solve_code_key():
    """16-hex hash of the solver sources (file name and bytes each)."""
    h = hashlib.sha256()
    for rel in SOLVE_SOURCES:
        h.update(rel.encode())
        h.update((_PKG / rel).read_bytes())
    return h.hexdigest()[:16]


def _as_dict(obj):
    return asdict(obj) if is_dataclass(obj) else dict(obj)


def _symmetry(sym):
    if is_dataclass(sym):
        return asdict(sym)
    return sym


def geometry_key_dict(spec):
    """The voltage-free description of what the solver sees."""
    g = spec.geometry
    electrodes = []
    for pos, el in enumerate(g.electrodes, start=1):
        d = {k: v for k, v in _as_dict(el).items()
             if k not in _VOLTAGE_FIELDS}
        d.update({k: None for k in _LADDER_FIELDS})
        d["_basis"] = d.get("basis") if d.get("basis") is not None else pos
        electrodes.append(d)
    return {"kind": CACHE_KIND, "_fmt": CACHE_FORMAT,
            "width_mm": g.width_mm, "height_mm": g.height_mm,
            "depth_mm": g.depth_mm or 0.0, "mm_per_gu": g.mm_per_gu,
            "symmetry": _symmetry(g.symmetry),
            "builder": getattr(spec, "builder", ""),
            "electrodes": electrodes}


def cache_key_dict(spec):
    """The geometry and the code that solves it."""
    return dict(geometry_key_dict(spec), _solve_code=solve_code_key())


def key(spec):
    return fa_cache.spec_key(cache_key_dict(spec))


def describe_spec(spec):
    """One line naming an entry: route, grid, pitch, electrode count."""
    g = spec.geometry
    dims = [d for d in (g.width_mm, g.height_mm, g.depth_mm or None)
            if d is not None]
    grid = "x".join(str(int(round(d / g.mm_per_gu)) + 1) for d in dims)
    route = getattr(spec, "route", None) or \
        ("3-D" if g.depth_mm else "2-D/rz")
    return (f"{route} | {grid} @ {g.mm_per_gu:g} mm/gu | "
            f"{len(g.electrodes)} electrodes")


def _labels_u8(ele, who):
    ele = np.ascontiguousarray(ele)
    top = int(ele.max()) if ele.size else 0
    if top > MAX_LABEL or (ele.size and int(ele.min()) < 0):
        raise ValueError(f"{who}: electrode label {top} does not fit the "
                         f"8-bit label array (at most {MAX_LABEL} "
                         f"electrodes)")
    return ele.astype(np.uint8)


def _spec_json(spec):
    try:
        return asdict(spec)
    except TypeError:
        return None


def store(spec, bases, ele, root=None):
    """Publish {idx: (n0, n1) basis} and the labels under the spec's key;
    returns the key."""
    arrays = {f"b{int(k)}": np.ascontiguousarray(v, np.float64)
              for k, v in bases.items()}
    arrays["ele"] = _labels_u8(ele, "basis_cache.store")
    k = fa_cache.store(cache_key_dict(spec), arrays, root=root)
    fa_cache.annotate(k, getattr(spec, "name", ""),
                      label=describe_spec(spec), spec_json=_spec_json(spec),
                      root=root)
    return k


def load(spec, root=None):
    """({idx: basis float64}, ele int16) or (None, None) on a miss; the
    arrays are copied out of the memmap (writable)."""
    arrays, _meta = fa_cache.load(cache_key_dict(spec), root=root)
    if arrays is None:
        return None, None
    ele = np.ascontiguousarray(arrays["ele"])
    if ele.dtype == bool:
        print("basis_cache: ignoring a pre-format-2 (bool) `ele` entry -- "
              "labels would collapse to electrode #1; re-solving.")
        return None, None
    bases = {int(n[1:]): np.array(a, np.float64)
             for n, a in arrays.items() if n.startswith("b")}
    fa_cache.annotate(cache_key_dict(spec), getattr(spec, "name", ""),
                      root=root)
    return bases, ele.astype(np.int16)


# ------------------------------------------------------------------ 3-D
def store_3d(key_dict, bases, labels, root=None, name="", label=None):
    """Publish 3-D bases as ``basis_{idx}`` plus uint8 ``labels`` under a
    key dict; returns the key."""
    arrays = {f"basis_{int(k)}": np.ascontiguousarray(v)
              for k, v in bases.items()}
    arrays["labels"] = _labels_u8(labels, "basis_cache.store_3d")
    k = fa_cache.store(key_dict, arrays, root=root)
    fa_cache.annotate(k, name, label=label, root=root)
    return k


def load_3d(key_dict, root=None, mmap=True):
    """({idx: basis}, labels) from a 3-D entry, or (None, None)."""
    arrays, _meta = fa_cache.load(key_dict, root=root, mmap=mmap)
    if arrays is None:
        return None, None
    bases = {int(n.split("_", 1)[1]): a for n, a in arrays.items()
             if n.startswith("basis_")}
    return bases, arrays["labels"]


def _digest(a):
    a = np.ascontiguousarray(a)
    h = hashlib.sha256(f"{a.dtype.str}{a.shape}".encode())
    h.update(a.reshape(-1).view(np.uint8).tobytes())
    return h.hexdigest()[:16]


def device_key_dict(masks, *, method, mirror, v_basis, tol, stencil,
                    theta=None):
    """Key of a device 3-D basis solve: each mask's content, the leg
    fractions' content, the solve settings and the solver code."""
    return {"kind": CACHE_KIND_3D, "_fmt": CACHE_FORMAT,
            "shape": list(np.shape(next(iter(masks.values())))),
            "masks": {str(int(i)): _digest(np.asarray(m, bool))
                      for i, m in sorted(masks.items())},
            "theta": None if theta is None else
            _digest(np.asarray(theta, np.float32)),
            "method": method, "mirror": [bool(m) for m in mirror],
            "v_basis": float(v_basis), "tol": float(tol),
            "stencil": stencil, "_solve_code": solve_code_key()}


def solve_bases_cached(masks, *, method="mg", mirror=(False, False, False),
                       v_basis=1e4, tol=1e-6, stencil="ghost_linear",
                       theta=None, root=None, engine=None, name="",
                       verbose=False):
    """{idx: mask} -> ({idx: basis}, hit): the device solver's bases for
    these masks (``method`` 'sor', 'mg' cascade or 'vmg' V-cycle, as
    solver3d.solve_bases / solve_bases_mg), from the cache when an entry
    for the same masks, settings and solver code exists, else solved on
    the device and published.  Hit bases are read-only memmaps."""
    from ..physics import solver3d as S
    if not masks:
        raise ValueError("solve_bases_cached: no electrode masks")
    kd = device_key_dict(masks, method=method, mirror=mirror,
                         v_basis=v_basis, tol=tol, stencil=stencil,
                         theta=theta)
    try:
        bases, _labels = load_3d(kd, root=root)
    except (IOError, OSError, ValueError) as e:
        print(f"basis_cache: entry {fa_cache.spec_key(kd)} unreadable ({e}); "
              f"re-solving")
        bases = None
    if bases is not None:
        return bases, True
    if method == "sor":
        bases = S.solve_bases(masks, mirror=mirror, v_basis=v_basis, tol=tol,
                              verbose=verbose, stencil=stencil, theta=theta,
                              engine=engine)
    elif method in ("mg", "vmg"):
        solve = S.solve_bases_mg if method == "vmg" else _cascade_bases
        bases = solve(masks, mirror=mirror, v_basis=v_basis, tol=tol,
                      verbose=verbose, stencil=stencil, theta=theta,
                      engine=engine)
    else:
        raise ValueError(f"solve_bases_cached: method {method!r} is not one "
                         f"of 'sor', 'mg', 'vmg'")
    labels = np.zeros(np.shape(next(iter(masks.values()))), np.int64)
    for i, m in masks.items():
        labels[np.asarray(m, bool)] = int(i)
    store_3d(kd, bases, labels, root=root, name=name,
             label=f"3-D {'x'.join(map(str, labels.shape))} | {method} | "
                   f"{len(masks)} electrodes")
    return bases, False


def _cascade_bases(masks, *, mirror, v_basis, tol, verbose, stencil, theta,
                   engine):
    from ..physics import solver3d as S
    return S._bases(masks, lambda f, v: S.solve3d_mg(
        f, v, theta=theta, mirror=mirror, tol=tol, stencil=stencil,
        engine=engine), v_basis, None, verbose)
