"""Ion births on the device: generate_births
(ion_gym/physics/sim_build.py:82-283) for a SourceSpec, drawn by the
engine from the same numpy PCG64 stream, in the same order, as the
reference's per-ion Python loop.

The engine resolves where each ion's draws start without a sequential
pass (ig_births / ig_ensemble_births_at in include/ion_gym_amd.h); this
module builds the source descriptor the reference's way (the Gaussian
layout refusals, the zero-direction warning, the seed), reads births
files on the host, and passes numpy's own standard-normal ziggurat tables
so that device normals are numpy's.
"""
import csv
import ctypes as C
import math
import os
import secrets
import warnings

import numpy as np

from . import _lib as L
from .physics.flight_spec import FWHM_PER_SIGMA, SOURCE_DISTRIBUTIONS

GAUSSIAN_BIRTH_TRIES = 1000     # ion_gym/physics/sim_build.py:78
_DIST = {"point": L.BIRTH_POINT, "disc": L.BIRTH_DISC, "line": L.BIRTH_LINE,
         "grid": L.BIRTH_GRID, "box": L.BIRTH_BOX,
         "gaussian": L.BIRTH_GAUSSIAN}
_M64 = (1 << 64) - 1
_M128 = (1 << 128) - 1
_PCG_MULT = 0x2360ED051FC65DA44385DF649FCCF645
_ZIG = None


def _pcg_out(s):
    hi, lo = s >> 64, s & _M64
    rot = hi >> 58
    v = hi ^ lo
    return ((v >> rot) | (v << ((64 - rot) & 63))) & _M64


def ziggurat_tables():
    """numpy's standard-normal ziggurat tables as (k uint64[256], w
    float64[256], f float64[256]).

    numpy's Generator draws a standard normal from one 64-bit word w: layer
    i = w & 0xff, sign bit 8, magnitude m = bits 9..60; x = m * w_i is
    returned at once when m < k_i.  Both tables are read off numpy itself:
    a PCG64 state is chosen whose next output is a given word (the output
    function XSL-RR is inverted for a state whose top six bits are zero),
    and the state numpy leaves behind shows whether a second word was
    taken.  m = 1 returns w_i exactly (the second word, also chosen, makes
    the slow-path test of layer 1 pass); k_i is the least m for which a
    second word is taken, found by bisection.  The layer densities are
    f_i = exp(-x_i^2 / 2) at the layer edges x_i = w_i 2^52 (f_0 = 1); they
    enter only the slow-path comparison, whose outcome a last-bit
    difference would change with probability ~1e-16."""
    global _ZIG
    if _ZIG is not None:
        return _ZIG
    inv = pow(_PCG_MULT, -1, 1 << 128)
    bg = np.random.PCG64(0)
    gen = np.random.Generator(bg)

    def run(word, second=None):
        h = 0x1234567
        post = (h << 64) | (h ^ word)
        if second is None:
            inc = 0xDA3E39CB94B95BDB
        else:
            # the increment that makes the following output `second`
            nxt = (post * _PCG_MULT) & _M128
            h2 = 0x7654321
            if ((h2 ^ second) & 1) == (nxt & 1):
                h2 ^= 1
            inc = (((h2 << 64) | (h2 ^ second)) - nxt) & _M128
        pre = ((post - inc) * inv) & _M128
        bg.state = {"bit_generator": "PCG64",
                    "state": {"state": pre, "inc": inc},
                    "has_uint32": 0, "uinteger": 0}
        assert _pcg_out(post) == word
        x = gen.standard_normal()
        return x, bg.state["state"]["state"] == post

    def word(i, m):
        return (i & 0xFF) | ((m & 0xFFFFFFFFFFFFF) << 9)

    k = np.zeros(256, np.uint64)
    w = np.zeros(256)
    for i in range(256):
        w[i] = run(word(i, 1), second=0)[0]
        if not run(word(i, 0))[1]:
            k[i] = 0
            continue
        lo, hi = 0, 1 << 52
        if run(word(i, hi - 1))[1]:
            k[i] = hi
            continue
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if run(word(i, mid))[1]:
                lo = mid
            else:
                hi = mid
        k[i] = hi
    x = w * 2.0 ** 52
    f = np.array([1.0] + [math.exp(-0.5 * v * v) for v in x[1:]])
    _ZIG = (k, w, f)
    return _ZIG


def resolve_run_seed(source):
    """The run seed (ion_gym/physics/ion_envelope.py:56-79): the declared
    seed, or for seed None one 31-bit draw of OS entropy per source,
    cached on it and printed so the run can be repeated."""
    if source.seed is not None:
        return int(source.seed)
    drawn = getattr(source, "_drawn_seed", None)
    if drawn is None:
        drawn = secrets.randbits(31)
        source._drawn_seed = drawn
        print(f"[seed] random mode: drew {drawn} for births and collisions "
              f"-- set source.seed: {drawn} to reproduce this run exactly",
              flush=True)
    return int(drawn)


def _check_dist(dist):
    if dist not in _DIST:
        raise ValueError(
            f"generate_births: source.distribution {dist!r} is not one this "
            f"generator places (known: {list(SOURCE_DISTRIBUTIONS)}; 'file' "
            f"needs source.births_file)")


def births_desc(source, seed=None):
    """(ig_births_desc, keep-alive) for a non-file source; raises what
    the reference raises before its first draw."""
    s = source
    _check_dist(s.distribution)
    note = s.zero_direction_note()
    if note:
        warnings.warn(note, stacklevel=3)
    if not s.mz_list:
        raise ValueError("source.mz_list is empty: no mass to fly")
    n_per = int(s.n_ions)
    if n_per < 1:
        raise ValueError(f"source.n_ions = {s.n_ions!r}: a whole number >= "
                         f"1 of ions per mass")
    d = L.BirthsDesc()
    d.distribution = _DIST[s.distribution]
    d.axis = {"x": 0, "y": 1, "z": 2}[s.axis]
    d.n_per = n_per
    mz = np.ascontiguousarray(np.asarray(s.mz_list, float))
    d.n_mz = len(mz)
    d.mz_list = mz.ctypes.data_as(L._dp)
    for k, v in enumerate((s.x0_mm, s.y0_mm, s.z0_mm)):
        d.p0[k] = float(v)
    d.r_mm = float(s.r_mm)
    d.len_mm = float(s.len_mm)
    box = (list(s.box_mm) + [0.0, 0.0, 0.0])[:3]
    for k in range(3):
        d.box_mm[k] = float(box[k])
    if s.distribution == "gaussian":
        sig, trc = s.gaussian_layout()
        for k in range(3):
            d.gauss_sigma[k] = float(sig[k])
            d.gauss_trunc[k] = float(trc[k])
    d.gauss_tries = GAUSSIAN_BIRTH_TRIES
    d.has_ke = int(s.ke_hi > 0 or s.ke_lo > 0)
    d.ke_lo = float(s.ke_lo)
    d.ke_hi = float(s.ke_hi)
    dv = np.array(s.direction, float)
    nrm = float(np.linalg.norm(dv))
    if nrm > 0.0:
        u = dv / nrm
        d.has_dir = 1
        for k in range(3):
            d.dir_unit[k] = float(u[k])
    d.temperature_k = float(s.temperature_k)
    dvf = list(s.dv_fwhm_ms or [])
    if any(x for x in dvf):
        if len(dvf) != 3:
            raise ValueError(
                f"source: dv_fwhm_ms must be a 3-vector of per-axis velocity "
                f"FWHM in m/s (x, y, z); got {s.dv_fwhm_ms!r} with "
                f"{len(dvf)} entries")
        if any(x < 0 for x in dvf):
            raise ValueError(f"source: dv_fwhm_ms entries are FWHM widths "
                             f"and cannot be negative; got {s.dv_fwhm_ms!r}")
        sig3 = np.array(dvf, float) * 1e-3 / FWHM_PER_SIGMA
        d.has_dv = 1
        for k in range(3):
            d.dv_sigma[k] = float(sig3[k])
    vd = list(getattr(s, "v_drift_mm_us", None) or [])
    if any(x for x in vd):
        if len(vd) != 3:
            raise ValueError(
                f"source: v_drift_mm_us must be a 3-vector of per-axis drift "
                f"velocity in mm/us; got {vd!r} with {len(vd)} entries")
        d.has_drift = 1
        for k in range(3):
            d.v_drift[k] = float(vd[k])
    d.tob_span_us = float(s.tob_span_us)
    st = np.random.PCG64(resolve_run_seed(s) if seed is None
                         else int(seed)).state["state"]
    d.pcg_state[0], d.pcg_state[1] = st["state"] >> 64, st["state"] & _M64
    d.pcg_inc[0], d.pcg_inc[1] = st["inc"] >> 64, st["inc"] & _M64
    zk, zw, zf = ziggurat_tables()
    d.zig_k = zk.ctypes.data_as(L._u64p)
    d.zig_w = zw.ctypes.data_as(L._dp)
    d.zig_f = zf.ctypes.data_as(L._dp)
    return d, (mz, zk, zw, zf)


def read_births_file(source, deck_path=None):
    """(N, 7) births from source.births_file (sim_build.py:88-137): a
    relative path is read beside the deck when one is named; columns x, y,
    z, vx, vy, vz are required and tob defaults to 0; the row count must
    be n_ions x the number of masses."""
    bf = source.births_file
    if not bf:
        raise ValueError("source: distribution 'file' needs births_file")
    if not os.path.isabs(bf) and deck_path:
        cand = os.path.join(os.path.dirname(os.path.abspath(deck_path)), bf)
        if not os.path.exists(cand):
            raise FileNotFoundError(
                f"births_file {bf!r} is relative, so it is read beside the "
                f"deck: {cand} does not exist")
        bf = cand
    need = ("x", "y", "z", "vx", "vy", "vz")
    rows = []
    with open(bf, encoding="utf-8") as fh:
        rd = csv.DictReader(fh)
        have = set(rd.fieldnames or ())
        missing = [c for c in need if c not in have]
        if missing:
            raise ValueError(f"births_file {bf!r}: missing column(s) "
                             f"{missing}; header has {sorted(have)}")
        for r in rd:
            rows.append([float(r.get(k, 0.0)) for k in (*need, "tob")])
    out = np.array(rows, float).reshape(-1, 7)
    want = int(source.n_ions) * max(1, len(source.mz_list))
    if len(out) != want:
        raise ValueError(
            f"births_file {bf!r} has {len(out)} rows but the source declares "
            f"n_ions={source.n_ions} x {max(1, len(source.mz_list))} masses "
            f"= {want}; masses are assigned by contiguous block")
    return out


def generate_births(source, engine, seed=None, deck_path=None):
    """(N, 7) = [x, y, z, vx, vy, vz, tob] (mm, mm/us, us) for the source,
    N = n_ions x len(mz_list), drawn on the engine's device (a births file
    is read on the host)."""
    if source.distribution == "file":
        return read_births_file(source, deck_path)
    d, keep = births_desc(source, seed)
    n = int(d.n_per) * int(d.n_mz)
    out = np.empty((n, 7))
    rc = L.lib().ig_births(engine.h, C.byref(d), out.ctypes.data_as(L._dp))
    if rc == L.E_BIRTH_TRIES:
        msg = L.lib().ig_last_error().decode(errors="replace")
        raise RuntimeError(f"generate_births: {msg}")
    L.check(rc, "ig_births")
    del keep
    return out


def ion_attributes(source, sigma_m2=None, seed=None):
    """The per-ion arrays of a source other than its births: mass (Da) by
    contiguous block, acc = q e / (m amu) (ion_gym/physics/build_rz.py:
    703-713), the per-ion seed base + i (ion_envelope.py:82-92) and the
    hard-sphere cross-section (a number, or a function of the mass)."""
    from .physics.constants import E_CHG, KG_AMU
    n_per = int(source.n_ions)
    mz = np.asarray(source.mz_list, float)
    m = np.repeat(mz, n_per)
    base = resolve_run_seed(source) if seed is None else int(seed)
    out = dict(m_amu=m, acc=int(source.charge) * E_CHG / (m * KG_AMU),
               seed=base + np.arange(len(m), dtype=np.int64))
    if sigma_m2 is not None:
        out["sigma_m2"] = (np.asarray([sigma_m2(v) for v in m], float)
                           if callable(sigma_m2)
                           else np.full(len(m), float(sigma_m2)))
    return out


def births_into(engine, ensemble, offset, source, seed=None):
    """Draw the source's births on the device straight into ensemble
    slots [offset, offset + N), which must be loaded already (per-ion
    masses, charges and seeds; births may be passed as None)."""
    if source.distribution == "file":
        raise ValueError("births_into: a births file is read on the host; "
                         "load its rows with the ensemble instead")
    d, keep = births_desc(source, seed)
    rc = L.lib().ig_ensemble_births_at(engine.h, ensemble.id, int(offset),
                                       C.byref(d))
    if rc == L.E_BIRTH_TRIES:
        msg = L.lib().ig_last_error().decode(errors="replace")
        raise RuntimeError(f"generate_births: {msg}")
    L.check(rc, "ig_ensemble_births_at")
    del keep


def refuse_births(engine, fields, ensemble, offset=0, n=None):
    """The fly functions' birth refusals on the device for the unflown
    ions of slots [offset, offset + n) (ig_ensemble_refuse_births): outside
    the domain -> fate 1, inside metal -> fate 0 with the electrode label;
    such ions get face -2 (refused at birth) and are never flown.  Returns
    how many were refused."""
    n = ensemble.n - int(offset) if n is None else int(n)
    out = C.c_int64()
    L.check(L.lib().ig_ensemble_refuse_births(engine.h, ensemble.id,
                                              fields.id, int(offset), n,
                                              C.byref(out)),
            "ig_ensemble_refuse_births")
    return out.value
