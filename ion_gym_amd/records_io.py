"""Output formats of a flight: the trajectory file and the results bank.

The engine's fly functions return the reference's ``(traj, summary)``
contract; this module writes and reads the two files ion_gym keeps results
in, in its formats, so a run flown here opens wherever ion_gym's own do:

* the trajectory file (ion_gym/io/field_io.py:484-650, ``.traj.npz``,
  format 2): one member ``t<i>`` per saved ion (float64, or float32 with
  the time column kept at float64 in ``time<i>``), ``label``, ``mz_da``,
  ``mz_source``, ``summary_cols`` and the (n, 7) ``summary`` table (ion,
  fate, tof, x/y/z_end, mz), the column names, and a JSON ``_meta`` frame
  stored as uint8 bytes;
* the results bank (ion_gym/io/results_bank.py:43-121): named per-ion
  arrays plus a ``__meta_json__`` string member that must carry the
  operating point; an existing bank is refused unless overwritten; writes
  are atomic.

Nothing is pickled, and both load with ``allow_pickle=False``.
"""
import json
import math
import os
import re
import tempfile
import time

import numpy as np

TRAJ_FORMAT = 2
TRAJ_DTYPES = ("float64", "float32")
SUMMARY_COLS = ["ion", "fate", "tof", "x_end", "y_end", "z_end", "mz_da"]
_META = "_meta"
BANK_META_KEY = "__meta_json__"
BANK_SCHEMA = 1


class TrajectoryFormatError(ValueError):
    """A file that is not a readable trajectory file."""


class BankError(RuntimeError):
    """A results-bank contract violation."""


def _atomic_savez(path, **members):
    d = os.path.dirname(os.path.abspath(path)) or "."
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(prefix=".traj.", suffix=".npz", dir=d)
    ok = False
    try:
        with os.fdopen(fd, "wb") as f:
            np.savez_compressed(f, **members)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
        ok = True
    finally:
        if not ok and os.path.exists(tmp):
            os.remove(tmp)


def _value(summary, key):
    v = (summary or {}).get(key)
    if v is None or isinstance(v, bool):
        return None
    try:
        f = float(v)
    except (TypeError, ValueError):
        return None
    return f if math.isfinite(f) else None


def _fate(summary):
    k = (summary or {}).get("kind")
    try:
        return int(k)
    except (TypeError, ValueError):
        return -1


def _jsonable(v):
    if isinstance(v, dict):
        return {str(k): _jsonable(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, float) and not math.isfinite(v):
        return None
    return v


def trajectory_members(results, columns, *, label="simulation",
                       dtype="float64", mz_of=None, geometry_key="",
                       provenance=None, spec_name=None):
    """(members, meta) of a trajectory file for results given as (index,
    traj, summary) tuples or objects with .index/.traj/.summary; an ion
    without a trajectory is skipped and counted.  An ion's mass is its
    summary's mz, else mz_of(index), else NaN ("unknown")."""
    if dtype not in TRAJ_DTYPES:
        raise ValueError(f"trajectory precision {dtype!r}: choose one of "
                         f"{TRAJ_DTYPES}")
    cols = list(columns)
    tcol = cols.index("t") if "t" in cols else None
    if dtype == "float32" and tcol is None:
        raise ValueError("a float32 trajectory file keeps its time column at "
                         "float64, but these columns name no 't'")
    results = list(results)
    members, index, summaries, masses, sources = {}, [], [], [], []
    for r in results:
        if isinstance(r, (tuple, list)):
            i, traj = r[0], r[1]
            summ = r[2] if len(r) > 2 else {}
        else:
            i, traj, summ = r.index, getattr(r, "traj", None), \
                getattr(r, "summary", None)
        if traj is None:
            continue
        i, summ = int(i), summ or {}
        traj = np.asarray(traj)
        if traj.ndim != 2 or traj.shape[1] != len(cols):
            raise ValueError(f"ion {i}: trajectory shape {traj.shape} does "
                             f"not match the {len(cols)} columns {cols}")
        members[f"t{i}"] = np.ascontiguousarray(traj, dtype)
        if dtype == "float32":
            members[f"time{i}"] = np.ascontiguousarray(traj[:, tcol],
                                                       np.float64)
        mz = _value(summ, "mz")
        if mz is not None:
            m, src = mz, "summary"
        elif mz_of is not None:
            try:
                m, src = float(mz_of(i)), "spec"
            except (ValueError, TypeError, IndexError) as e:
                m, src = float("nan"), f"unknown ({type(e).__name__})"
        else:
            m, src = float("nan"), "unknown (no mass)"
        index.append(i)
        summaries.append(_jsonable(summ))
        masses.append(m)
        sources.append(src)
    meta = dict(kind="ion_gym_trajectories", traj_format=TRAJ_FORMAT,
                precision=dtype, columns=cols,
                frame="Convention A: axial coordinate horizontal (z axial)",
                units=dict(position="mm", time="us", velocity="mm/us"),
                geometry_key=str(geometry_key), spec_name=spec_name or label,
                label=label, ion_index=index, summaries=summaries,
                mz_da=["unknown" if math.isnan(m) else m for m in masses],
                mz_source=sources, n_skipped=len(results) - len(index),
                provenance=_jsonable(provenance or {}))
    rows = [[float(i), float(_fate(s)),
             *(np.nan if _value(s, k) is None else _value(s, k)
               for k in ("tof", "x_end", "y_end", "z_end")), m]
            for i, s, m in zip(index, summaries, masses)]
    plain = {"label": np.array([str(label)]),
             "mz_da": np.array(masses, float),
             "mz_source": np.array(sources),
             "summary_cols": np.array(SUMMARY_COLS),
             "columns": np.array([str(c) for c in cols]),
             "summary": (np.array(rows, float) if rows
                         else np.zeros((0, len(SUMMARY_COLS))))}
    blob = np.frombuffer(json.dumps(meta, sort_keys=True).encode(), np.uint8)
    return {**members, **plain, _META: blob}, meta


def save_trajectories(results, columns, path, **kw):
    """Write a trajectory file atomically; returns the path."""
    members, _ = trajectory_members(results, columns, **kw)
    _atomic_savez(path, **members)
    return path


def read_trajectory_meta(source):
    """The ``_meta`` frame of a trajectory file; refuses anything else."""
    with np.load(source, allow_pickle=False) as z:
        if _META not in z.files:
            raise TrajectoryFormatError("no embedded _meta: not a trajectory "
                                        "file")
        meta = json.loads(bytes(np.asarray(z[_META], np.uint8)).decode())
    if meta.get("kind") != "ion_gym_trajectories":
        raise TrajectoryFormatError(f"_meta.kind {meta.get('kind')!r} != "
                                    f"'ion_gym_trajectories'")
    return meta


def load_trajectories(source):
    """({ion index: (n, k) float64 trajectory}, meta); a float32 file's
    time column is restored from its float64 copy."""
    meta = read_trajectory_meta(source)
    if meta.get("traj_format") not in (1, TRAJ_FORMAT):
        raise TrajectoryFormatError(f"traj_format {meta.get('traj_format')} "
                                    f"is not readable")
    cols = meta.get("columns") or []
    tcol = cols.index("t") if "t" in cols else None
    trajs = {}
    with np.load(source, allow_pickle=False) as z:
        for i in meta.get("ion_index", []):
            name = f"t{int(i)}"
            if name not in z.files:
                raise TrajectoryFormatError(f"{name} listed in _meta but "
                                            f"missing from the file")
            arr = np.array(z[name], dtype=np.float64)
            if f"time{int(i)}" in z.files and tcol is not None:
                arr[:, tcol] = z[f"time{int(i)}"]
            trajs[int(i)] = arr
    return trajs, meta


def bank_run(path, *, meta, arrays, overwrite=False, seed=None):
    """Write one results bank (results_bank.py:43-121); returns its path.
    ``meta`` must be JSON-serialisable and carry 'operating_point'."""
    if "operating_point" not in meta:
        raise BankError("bank_run: meta lacks 'operating_point' -- a banked "
                        "number without its operating point is not a result")
    if not arrays:
        raise BankError("bank_run: no arrays -- an empty bank is not a result")
    if BANK_META_KEY in arrays:
        raise BankError(f"bank_run: array name {BANK_META_KEY!r} is reserved")
    p = str(path)
    if not p.endswith(".npz"):
        p += ".npz"
    if os.path.exists(p) and not overwrite:
        raise BankError(f"bank_run: {p!r} exists -- pass overwrite=True or "
                        f"choose a new name")
    full = dict(meta)
    if seed is not None:
        full["seed"], full["seed_mode"] = int(seed), "fixed"
    elif "seed" not in full:
        full["seed"], full["seed_mode"] = None, "not recorded"
    full["schema"] = BANK_SCHEMA
    full["created_utc"] = time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime())
    try:
        meta_json = json.dumps(full)
    except TypeError as e:
        raise BankError(f"bank_run: meta is not JSON-serializable ({e})") \
            from e
    payload = {BANK_META_KEY: np.array(meta_json)}
    for name, arr in arrays.items():
        payload[name] = np.asarray(arr)
    tmp = p + ".tmp"
    try:
        with open(tmp, "wb") as fh:
            np.savez_compressed(fh, **payload)
        os.replace(tmp, p)
    except BaseException:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise
    return p


def load_bank(path):
    """(meta, {name: array}) of a bank; refuses a file without the meta."""
    p = str(path)
    with np.load(p, allow_pickle=False) as z:
        if BANK_META_KEY not in z.files:
            raise BankError(f"load_bank: {p!r} has no {BANK_META_KEY!r} "
                            f"block -- not a results bank")
        meta = json.loads(str(z[BANK_META_KEY][()]))
        arrays = {k: z[k] for k in z.files if k != BANK_META_KEY}
    return meta, arrays


_SAFE = re.compile(r"[^A-Za-z0-9._-]+")


def bank_name(label):
    """A file-system-safe bank name for a label."""
    return _SAFE.sub("_", str(label)).strip("_") or "bank"
