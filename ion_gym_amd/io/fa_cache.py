"""Content-keyed disk entries of solved field arrays, in ion_gym's on-disk
format (ion_gym/io/fa_cache.py), so an entry either side writes the other
can serve.

An entry is a directory ``<root>/<key>/`` holding one ``<name>.npy`` per
array and a ``meta.json`` of the form

    {"key": <16 hex>, "version": 2, "spec": <the JSON spec keyed>,
     "arrays": {<name>: {"shape": [...], "dtype": "<numpy str>",
                         "sha_head": <16 hex>}}}

``key`` is the first 16 hex digits of the SHA-256 of the canonical JSON of
``{"v": CACHE_VERSION, "spec": spec}`` (keys sorted, no spaces, floats
rounded to 10 decimals so representation noise keeps the key,
fa_cache.py:64-89); ``sha_head`` the first 16 hex digits of the SHA-256 of
the array's first 4096 raw bytes (:93-94).  Arrays come back as read-only
memmaps: a hit costs page-cache reads, not a deserialisation.  A new entry
is written into a hidden sibling directory and renamed into place when
complete, so a reader never sees half an entry; a writer that loses the
rename to another publishes nothing and removes its copy (:113-158).
A load refuses, as IOError, an array whose file is missing, unreadable
(a truncated ``.npy``), of another shape or dtype than the manifest, or
whose head hash differs (:161-201).
"""
import hashlib
import json
import os
import tempfile
import time
from pathlib import Path

import numpy as np

CACHE_VERSION = 2
KEY_HEX = 16
HEAD_BYTES = 4096


def cache_root():
    """The cache directory: ``$ION_GYM_CACHE`` when set, else
    ``~/.ion_gym_cache``; read at every call (paths.py:251-261)."""
    env = os.environ.get("ION_GYM_CACHE")
    return Path(env).expanduser() if env else Path.home() / ".ion_gym_cache"


def _root(root):
    return Path(cache_root() if root is None else root).expanduser()


def __getattr__(name):
    if name == "DEFAULT_ROOT":
        return str(cache_root())
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


# ------------------------------------------------------------------ keys
def _plain(x):
    """JSON-ready copy of x: numpy scalars and arrays to Python values,
    floats rounded to 10 decimals, dicts with sorted keys."""
    if isinstance(x, dict):
        return {k: _plain(x[k]) for k in sorted(x)}
    if isinstance(x, (list, tuple)):
        return [_plain(v) for v in x]
    if isinstance(x, np.ndarray):
        return _plain(x.tolist())
    if isinstance(x, (float, np.floating)):
        return round(float(x), 10)
    if isinstance(x, np.integer):
        return int(x)
    return x


def canonical(obj):
    """The text a key hashes."""
    return json.dumps(_plain(obj), sort_keys=True, separators=(",", ":"))


def spec_key(spec):
    """16-hex key of a JSON-able spec dict."""
    text = canonical({"v": CACHE_VERSION, "spec": spec})
    return hashlib.sha256(text.encode()).hexdigest()[:KEY_HEX]


def _as_key(spec_or_key):
    if isinstance(spec_or_key, str) and len(spec_or_key) == KEY_HEX:
        return spec_or_key
    return spec_key(spec_or_key)


def sha_head(a):
    """16-hex hash of an array's first 4096 bytes."""
    raw = np.ascontiguousarray(a).reshape(-1).view(np.uint8)[:HEAD_BYTES]
    return hashlib.sha256(raw.tobytes()).hexdigest()[:KEY_HEX]


_sha_head = sha_head


# ------------------------------------------------------------------ entries
def _read_meta(entry):
    """The entry's meta.json as a dict, or None when absent/unreadable."""
    try:
        with open(entry / "meta.json", encoding="utf-8") as f:
            meta = json.load(f)
    except (OSError, ValueError):
        return None
    return meta if isinstance(meta, dict) else None


def _complete(entry):
    meta = _read_meta(entry)
    if meta is None or not isinstance(meta.get("arrays"), dict):
        return False
    return all((entry / f"{n}.npy").exists() for n in meta["arrays"])


def _write_entry(tmp, key, spec, arrays):
    manifest = {}
    for name, a in arrays.items():
        a = np.ascontiguousarray(a)
        np.save(tmp / f"{name}.npy", a)
        manifest[name] = {"shape": list(a.shape), "dtype": str(a.dtype),
                          "sha_head": sha_head(a)}
    with open(tmp / "meta.json", "w", encoding="utf-8") as f:
        json.dump({"key": key, "version": CACHE_VERSION, "spec": spec,
                   "arrays": manifest}, f, indent=1)
        f.flush()
        os.fsync(f.fileno())


def store(spec, arrays, root=None):
    """Publish {name: array} under the spec's key (atomically); returns the
    key.  A complete entry already there is kept; an entry whose manifest
    names a missing array is replaced."""
    base = _root(root)
    key = spec_key(spec)
    dest = base / key
    if (dest / "meta.json").exists():
        if _complete(dest):
            return key
        failed = remove_tree(dest)
        if failed:
            raise CachePurgeIncomplete(0, 0, failed)
        print(f"fa_cache: entry {key} was INCOMPLETE (arrays listed in "
              f"meta.json missing); replacing it with this solve")
    base.mkdir(parents=True, exist_ok=True)
    tmp = Path(tempfile.mkdtemp(prefix=f".{key}.", dir=base))
    published = False
    try:
        _write_entry(tmp, key, spec, arrays)
        try:
            os.replace(tmp, dest)
            published = True
        except OSError:
            # another writer published first: theirs is served
            if not (dest / "meta.json").exists():
                raise
    finally:
        if not published and tmp.is_dir():
            failed = remove_tree(tmp)
            if failed:
                print(f"fa_cache: could not remove the unused temp copy "
                      f"{tmp} ({len(failed)} path(s) failed); clear_all "
                      f"removes it")
    return key


def _open_array(entry, key, name, info, mmap):
    path = entry / f"{name}.npy"
    if not path.exists():
        raise IOError(f"cache {key}: missing array {name}")
    try:
        a = np.load(path, mmap_mode="r" if mmap else None)
    except (ValueError, OSError, EOFError) as e:
        raise IOError(f"cache {key}: {name} unreadable ({e}) -- truncated "
                      f"or corrupt entry")
    if list(a.shape) != list(info.get("shape", [])) or \
            str(a.dtype) != info.get("dtype"):
        raise IOError(f"cache {key}: {name} shape/dtype mismatch (got "
                      f"{a.shape}/{a.dtype}, meta {info.get('shape')}/"
                      f"{info.get('dtype')}) -- truncated or corrupt entry")
    want = info.get("sha_head")
    if want is None or sha_head(a) != want:
        raise IOError(f"cache {key}: {name} head hash "
                      f"{'missing' if want is None else 'mismatch'} -- "
                      f"corrupt entry; remove it and re-solve")
    return a


def load(spec_or_key, root=None, mmap=True, names=None):
    """(arrays, meta) of an entry, or (None, None) on a miss (no entry, or
    another key or version in its meta.json).  Arrays are read-only
    memmaps unless ``mmap`` is False."""
    key = _as_key(spec_or_key)
    entry = _root(root) / key
    if not (entry / "meta.json").exists():
        return None, None
    with open(entry / "meta.json", encoding="utf-8") as f:
        meta = json.load(f)
    if meta.get("key") != key or meta.get("version") != CACHE_VERSION:
        return None, None
    listed = meta.get("arrays") or {}
    out = {n: _open_array(entry, key, n, listed.get(n, {}), mmap)
           for n in (names or list(listed))}
    return out, meta


def verify(spec_or_key, root=None):
    """True when the entry exists and every array reads back whole with
    its manifest's shape, dtype and head hash."""
    try:
        arrays, _meta = load(spec_or_key, root=root, mmap=False)
    except (IOError, OSError, ValueError):
        return False
    return arrays is not None


# ------------------------------------------------------------------ provenance
def annotate(spec_or_key, name, label=None, root=None, spec_json=None):
    """Record in the entry's provenance.json (a list, never meta.json) that
    the spec ``name`` produced or resolved it: one record per (name,
    label), its time refreshed on a repeat.  Descriptive only: a missing
    entry or an empty name is a no-op, an unreadable sidecar is rebuilt,
    an unwritable one is reported and ignored (fa_cache.py:213-271)."""
    name = (name or "").strip()
    entry = _root(root) / _as_key(spec_or_key)
    if not name or not entry.is_dir():
        return
    side = entry / "provenance.json"
    records = []
    if side.exists():
        try:
            records = json.loads(side.read_text(encoding="utf-8"))
            if not isinstance(records, list):
                raise ValueError(f"expected a list, got "
                                 f"{type(records).__name__}")
        except (OSError, ValueError) as e:
            print(f"fa_cache: provenance sidecar unreadable for {entry.name} "
                  f"({e}) -- rebuilding it (descriptive data only).")
            records = []
    now = time.strftime("%Y-%m-%d %H:%M")
    rec = next((r for r in records if isinstance(r, dict)
                and r.get("name") == name and r.get("label") == label), None)
    if rec is None:
        rec = {"name": name, "label": label}
        records.append(rec)
    rec["when"] = now
    if spec_json is not None:
        rec["spec_json"] = spec_json
    try:
        fd, tmp = tempfile.mkstemp(dir=entry, prefix=".prov.")
        with os.fdopen(fd, "w") as f:
            json.dump(records, f, indent=1)
        os.replace(tmp, side)
    except OSError as e:
        print(f"fa_cache: could not write provenance for {entry.name} ({e}); "
              f"the entry itself is unaffected")


# ------------------------------------------------------------------ inventory
def _tree_bytes(path):
    """(bytes, files that could not be sized) under path."""
    size = unsized = 0
    for dirpath, _dirs, files in os.walk(path):
        for f in files:
            try:
                size += os.path.getsize(os.path.join(dirpath, f))
            except OSError:
                unsized += 1
    return size, unsized


def entries(root=None):
    """One dict per entry directory (key, bytes, unsized_files, arrays,
    mtime, spec_name, shape, produced_by, descriptor, has_spec_json,
    problems); a malformed entry is listed with the reason, never
    raised (fa_cache.py:274-346)."""
    base = _root(root)
    if not base.is_dir():
        return []
    out = []
    for entry in sorted(p for p in base.iterdir()
                        if p.is_dir() and not p.name.startswith(".")):
        size, unsized = _tree_bytes(entry)
        row = {"key": entry.name, "bytes": size, "unsized_files": unsized,
               "arrays": sum(1 for _d, _s, fs in os.walk(entry)
                             for f in fs if f.endswith(".npy")),
               "mtime": entry.stat().st_mtime, "spec_name": None,
               "shape": None, "produced_by": [], "problems": []}
        side = entry / "provenance.json"
        if side.exists():
            try:
                recs = json.loads(side.read_text(encoding="utf-8"))
                recs = [r for r in recs if isinstance(r, dict)] \
                    if isinstance(recs, list) else []
                row["produced_by"] = [r["name"] for r in recs
                                      if r.get("name")]
                labels = [r["label"] for r in recs if r.get("label")]
                row["descriptor"] = labels[-1] if labels else None
                row["has_spec_json"] = any(r.get("spec_json") for r in recs)
            except (OSError, ValueError) as e:
                row["problems"].append(f"provenance.json unreadable "
                                       f"({type(e).__name__}: {e})")
        if (entry / "meta.json").exists():
            meta = _read_meta(entry)
            if meta is None:
                row["problems"].append("meta.json unreadable")
            else:
                spec = meta.get("spec")
                row["spec_name"] = (spec.get("name")
                                    if isinstance(spec, dict) else None)
                arrs = meta.get("arrays") or {}
                if isinstance(arrs, dict) and arrs:
                    first = next(iter(arrs.values()))
                    row["shape"] = (first.get("shape")
                                    if isinstance(first, dict) else None)
        out.append(row)
    return out


class CachePurgeIncomplete(OSError):
    """A removal that left paths behind: carries what was removed
    (n_removed, bytes_freed) and {path: reason} for each path left."""

    def __init__(self, n_removed, bytes_freed, failures):
        self.n_removed = n_removed
        self.bytes_freed = bytes_freed
        self.failures = dict(failures)
        shown = "; ".join(f"{p} ({w})" for p, w in
                          list(self.failures.items())[:5])
        extra = len(self.failures) - 5
        super().__init__(
            f"cache purge incomplete: {n_removed} entries removed "
            f"({bytes_freed / 1e6:.1f} MB); could not delete "
            f"{len(self.failures)} path(s): {shown}"
            + (f" (+{extra} more)" if extra > 0 else ""))


def _why(e):
    return f"{type(e).__name__}: {e.strerror or e}"


def remove_tree(path):
    """Delete a directory tree bottom-up; {path: reason} for every path
    that could not be deleted ({} when all is gone or nothing was
    there)."""
    path = Path(path)
    failed = {}
    if not os.path.lexists(path):
        return failed

    def blocked(p):
        prefix = str(p) + os.sep
        return any(k.startswith(prefix) for k in failed)

    for dirpath, dirs, files in os.walk(path, topdown=False):
        for name in files + [d for d in dirs
                             if os.path.islink(os.path.join(dirpath, d))]:
            p = os.path.join(dirpath, name)
            try:
                os.remove(p)
            except OSError as e:
                failed[p] = _why(e)
        for d in dirs:
            p = os.path.join(dirpath, d)
            if os.path.isdir(p) and not os.path.islink(p) and p not in failed:
                try:
                    os.rmdir(p)
                except OSError as e:
                    if not blocked(p):
                        failed[p] = _why(e)
    try:
        os.rmdir(path)
    except OSError as e:
        if not blocked(path):
            failed[str(path)] = _why(e)
    return failed


def remove(key, root=None):
    """Delete one entry: (removed, bytes freed); (False, 0) when absent.
    Raises CachePurgeIncomplete when paths are left."""
    entry = _root(root) / key
    if not entry.is_dir():
        return False, 0
    freed, unsized = _tree_bytes(entry)
    failed = remove_tree(entry)
    if failed:
        raise CachePurgeIncomplete(0, 0, failed)
    if unsized:
        print(f"fa_cache: {unsized} file(s) in {key} could not be sized; "
              f"{freed / 1e6:.1f} MB freed is a lower bound")
    return True, freed


def clear_all(root=None):
    """Delete every entry directory under root (loose files are kept):
    (entries removed, bytes freed).  Removes all it can, then raises
    CachePurgeIncomplete naming what was left."""
    base = _root(root)
    if not base.is_dir():
        return 0, 0
    n = freed = unsized = 0
    failed = {}
    for entry in base.iterdir():
        if entry.is_symlink() or not entry.is_dir():
            continue
        size, n_uns = _tree_bytes(entry)
        left = remove_tree(entry)
        if left:
            failed.update(left)
            continue
        n += 1
        freed += size
        unsized += n_uns
    if failed:
        raise CachePurgeIncomplete(n, freed, failed)
    if unsized:
        print(f"fa_cache: {unsized} file(s) could not be sized; "
              f"{freed / 1e6:.1f} MB freed is a lower bound")
    return n, freed
