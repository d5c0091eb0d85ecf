"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker that tests/, ``smoke()`` and the
benchmark's CPU-baseline leg compare the GPU engine against.  The product
package ``ion_gym_amd`` never imports this module.

The oracle restates the reference's numba flight kernels in strict IEEE C
(ion_gym/physics/tracer_rz.py:113-446, tracer3d.py:280-575,
collision3d.py:265-374, sds.py:134-169); see ig_oracle.h for the map.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

RNG_MT = 0
RNG_PHILOX = 1

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class RZFields(C.Structure):
    _fields_ = [("nx", C.c_int64), ("nu", C.c_int64),
                ("EzA", _dp), ("EuA", _dp), ("EzK", _dp), ("EuK", _dp),
                ("EzG", _dp), ("EuG", _dp), ("tau_gate", C.c_double),
                ("ele", _dp), ("label", _u8p), ("u0", C.c_double),
                ("mm", C.c_double), ("K", C.c_int32), ("ch_kind", _i64p),
                ("ch_om", _dp), ("ch_ph", _dp), ("ch_duty", _dp),
                ("tab_t", _dp), ("tab_v", _dp), ("tab_off", _i64p)]


class XYZFields(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("EAx", _dp), ("EAy", _dp), ("EAz", _dp),
                ("ExK", _dp), ("EyK", _dp), ("EzK", _dp), ("ele", _u8p),
                ("h_mm", C.c_double), ("K", C.c_int32), ("ch_kind", _i64p),
                ("ch_om", _dp), ("ch_ph", _dp), ("ch_amp", _dp),
                ("ch_off", _dp), ("ch_duty", _dp), ("tab_t", _dp),
                ("tab_v", _dp), ("tab_off", _i64p)]


class PlanarFields(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64),
                ("ExA", _dp), ("EyA", _dp), ("ExK", _dp), ("EyK", _dp),
                ("ele", _dp), ("mm", C.c_double), ("K", C.c_int32),
                ("ch_kind", _i64p), ("ch_om", _dp), ("ch_ph", _dp),
                ("ch_duty", _dp), ("tab_t", _dp), ("tab_v", _dp),
                ("tab_off", _i64p)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("t_max", C.c_double),
                ("max_steps", C.c_int64), ("rec_every", C.c_int64),
                ("rec_rows", C.c_int64), ("rec_cols", C.c_int64),
                ("flags", C.c_int32 * 13),
                ("T_k", C.c_double), ("P_pa", C.c_double),
                ("c_star", C.c_double), ("c_bar", C.c_double),
                ("sig1d", C.c_double), ("m_gas", C.c_double),
                ("gas", C.c_int32), ("vg", C.c_double * 3),
                ("sds_stats", _dp), ("bnd_on", C.c_int32 * 6),
                ("bnd_val", C.c_double * 6), ("n_pl", C.c_int32),
                ("pl_col", _i64p), ("pl_val", _dp), ("pl_sgn", _dp),
                ("pl_w", _dp), ("pl_kind", _i64p), ("rng", C.c_int32),
                ("rng_key", C.c_uint64), ("tp_col", C.c_int32),
                ("tp_accept", C.c_double), ("tp_emit", C.c_double),
                ("tp_sgn", C.c_double), ("tp_maxp", C.c_int64)]


class Ions(C.Structure):
    _fields_ = [("n", C.c_int64)] + [
        (k, _dp) for k in ("x", "y", "z", "vx", "vy", "vz", "tob", "m_ion",
                           "acc", "sigma")] + [("seed", _i64p), ("sds", _dp)]


class Out(C.Structure):
    _fields_ = [(k, _dp) for k in ("x", "y", "z", "vx", "vy", "vz", "t")] + [
        ("kind", _i32p), ("face", _i32p), ("label", _i32p),
        ("ncol", _i64p), ("steps", _i64p), ("nrec", _i64p), ("rec", _dp)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_rz_fly.argtypes = [C.POINTER(RZFields), C.POINTER(Params),
                                C.POINTER(Ions), C.POINTER(Out), C.c_int]
        L.or_xyz_fly.argtypes = [C.POINTER(XYZFields), C.POINTER(Params),
                                 C.POINTER(Ions), C.POINTER(Out), C.c_int]
        L.or_planar_fly.argtypes = [C.POINTER(PlanarFields), C.POINTER(Params),
                                 C.POINTER(Ions), C.POINTER(Out), C.c_int]
        L.or_bilin.restype = C.c_double
        L.or_bilin.argtypes = [_dp, C.c_double, C.c_double, C.c_int64,
                               C.c_int64]
        L.or_tri.restype = C.c_double
        L.or_tri.argtypes = [_dp, C.c_double, C.c_double, C.c_double,
                             C.c_int64, C.c_int64, C.c_int64]
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32)]
        L.or_mt_draws.argtypes = [C.c_uint32, C.c_int, C.c_int64, _dp]
        L.or_collide_run.argtypes = [C.c_uint32, C.c_int64, C.c_double,
                                     C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, _dp]
        L.or_c_bar_rel.restype = C.c_double
        L.or_c_bar_rel.argtypes = [C.c_double, C.c_double, C.c_double]
        L.or_icdf_steps_at.restype = C.c_double
        L.or_icdf_steps_at.argtypes = [_dp, C.c_double, C.c_double]
        L.or_wave_rz.restype = C.c_double
        L.or_wave_rz.argtypes = [C.c_int64, C.c_double, C.c_double,
                                 C.c_double, _dp, _dp, C.c_int64, C.c_int64,
                                 C.c_double]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def bilin(F, gx, gu):
    F = np.ascontiguousarray(F, np.float64)
    return lib().or_bilin(_p(F), gx, gu, F.shape[0], F.shape[1])


def tri(F, x, y, z):
    F = np.ascontiguousarray(F, np.float64)
    return lib().or_tri(_p(F), x, y, z, *F.shape)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return tuple(int(v) for v in o)


def mt_draws(seed, n, normal=False):
    out = np.empty(n)
    lib().or_mt_draws(seed, 1 if normal else 0, n, _p(out))
    return out


def collide_run(seed, n, v0, m_ion, m_gas, sig1d):
    out = np.empty((n, 3))
    lib().or_collide_run(seed, n, v0[0], v0[1], v0[2], m_ion, m_gas, sig1d,
                         _p(out))
    return out


def c_bar_rel(speed, c_star, c_bar):
    return lib().or_c_bar_rel(speed, c_star, c_bar)


def icdf_steps_at(stats, log_mr, u):
    s = np.ascontiguousarray(stats, np.float64)
    return lib().or_icdf_steps_at(_p(s), log_mr, u)


def wave_rz(kind, om, ph, duty, t, tab_t=None, tab_v=None, o0=0, o1=0):
    tt = np.ascontiguousarray(tab_t if tab_t is not None else np.zeros(1))
    tv = np.ascontiguousarray(tab_v if tab_v is not None else np.zeros(1))
    return lib().or_wave_rz(kind, om, ph, duty, _p(tt), _p(tv), o0, o1, t)


class _Keep:
    """Holds the numpy arrays a ctypes struct points into."""

    def __init__(self):
        self.a = []

    def __call__(self, arr, dtype=np.float64):
        if arr is None:
            return None
        arr = np.ascontiguousarray(arr, dtype)
        self.a.append(arr)
        return arr


def _params(keep, fl, route_cols):
    p = Params()
    p.dt = fl["dt"]
    p.t_max = fl["t_max"]
    p.max_steps = int(fl.get("max_steps", 0))
    p.rec_every = int(fl.get("rec_every", 1))
    p.rec_rows = int(fl.get("rec_rows", 0))
    p.rec_cols = int(route_cols)
    for i, v in enumerate(fl.get("flags", [0] * 13)):
        p.flags[i] = int(bool(v))
    gas = fl.get("gas", {})
    p.gas = int(gas.get("model", 0))
    for k in ("T_k", "P_pa", "c_star", "c_bar", "sig1d", "m_gas"):
        setattr(p, k, float(gas.get(k, 0.0)))
    for i, v in enumerate(gas.get("flow", (0.0, 0.0, 0.0))):
        p.vg[i] = v
    st = keep(gas.get("sds_stats"))
    p.sds_stats = _p(st)
    for i in range(6):
        p.bnd_on[i] = int(bool(fl.get("bnd_on", [0] * 6)[i]))
        p.bnd_val[i] = float(fl.get("bnd_val", [0.0] * 6)[i])
    tp = fl.get("transporter")
    p.tp_col = -1
    if tp is not None:
        p.tp_col = {"x": 0, "y": 1, "z": 2}[str(tp["axis"])]
        p.tp_accept = float(tp["accept_mm"])
        p.tp_emit = float(tp["emit_mm"])
        p.tp_sgn = float(tp.get("direction", -1))
        p.tp_maxp = int(tp.get("max_passes", 10 ** 9))
    planes = fl.get("planes", [])
    p.n_pl = len(planes)
    if planes:
        p.pl_col = _p(keep([pl[0] for pl in planes], np.int64), _i64p)
        p.pl_val = _p(keep([pl[1] for pl in planes]))
        p.pl_sgn = _p(keep([pl[2] for pl in planes]))
        p.pl_w = _p(keep([list(pl[3]) for pl in planes]))
        p.pl_kind = _p(keep([pl[4] for pl in planes], np.int64), _i64p)
    p.rng = int(fl.get("rng", RNG_PHILOX))
    p.rng_key = int(fl.get("rng_key", 0))
    return p


def _ions(keep, ions):
    """Ion arrays under the engine's names (m_amu, sigma_m2, ...)."""
    io = Ions()
    io.n = len(ions["x"])
    names = {"m_ion": "m_amu", "sigma": "sigma_m2"}
    for k in ("x", "y", "z", "vx", "vy", "vz", "tob", "m_ion", "acc",
              "sigma"):
        v = ions.get(names.get(k, k))
        setattr(io, k, _p(keep(v if v is not None else np.zeros(io.n))))
    io.seed = _p(keep(ions["seed"], np.int64), _i64p)
    io.sds = _p(keep(ions.get("sds")))
    return io


def _out(n, rec_rows, rec_cols):
    res = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz", "t")}
    res.update(kind=np.zeros(n, np.int32), face=np.zeros(n, np.int32),
               label=np.zeros(n, np.int32), ncol=np.zeros(n, np.int64),
               steps=np.zeros(n, np.int64), nrec=np.zeros(n, np.int64))
    res["rec"] = (np.zeros((n, rec_rows, rec_cols)) if rec_rows > 0
                  else None)
    o = Out()
    for k in ("x", "y", "z", "vx", "vy", "vz", "t"):
        setattr(o, k, _p(res[k]))
    o.kind = _p(res["kind"], _i32p)
    o.face = _p(res["face"], _i32p)
    o.label = _p(res["label"], _i32p)
    o.ncol = _p(res["ncol"], _i64p)
    o.steps = _p(res["steps"], _i64p)
    o.nrec = _p(res["nrec"], _i64p)
    o.rec = _p(res["rec"])
    return o, res


def fly_rz(fields, flight, ions, n_threads=None):
    """Fly ions through an r-z field pack (reference layout dict: EzA, EuA,
    EzK, EuK, [EzG, EuG, tau_gate], ele (mirrored), [label], u0, mm,
    ch_kind, ch_om, ch_ph, ch_duty, [tab_t, tab_v, tab_off]).  Returns a
    dict of numpy arrays."""
    keep = _Keep()
    f = RZFields()
    EzA = keep(fields["EzA"])
    f.nx, f.nu = EzA.shape
    f.EzA = _p(EzA)
    f.EuA = _p(keep(fields["EuA"]))
    K = int(np.asarray(fields["ch_kind"]).shape[0])
    f.K = K
    f.EzK = _p(keep(fields["EzK"] if K else np.zeros(1)))
    f.EuK = _p(keep(fields["EuK"] if K else np.zeros(1)))
    tau = float(fields.get("tau_gate", -1.0))
    f.tau_gate = tau
    if fields.get("EzG") is not None:
        f.EzG = _p(keep(fields["EzG"]))
        f.EuG = _p(keep(fields["EuG"]))
    f.ele = _p(keep(fields["ele"]))
    if fields.get("label") is not None:
        f.label = _p(keep(fields["label"], np.uint8), _u8p)
    f.u0 = float(fields["u0"])
    f.mm = float(fields["mm"])
    f.ch_kind = _p(keep(fields["ch_kind"] if K else np.zeros(1), np.int64),
                   _i64p)
    for k in ("ch_om", "ch_ph", "ch_duty"):
        setattr(f, k, _p(keep(fields[k] if K else np.zeros(1))))
    f.tab_t = _p(keep(fields.get("tab_t", np.zeros(1))))
    f.tab_v = _p(keep(fields.get("tab_v", np.zeros(1))))
    f.tab_off = _p(keep(fields.get("tab_off", np.zeros(K + 1)), np.int64),
                   _i64p)
    ncols = 7 + sum(bool(v) for v in flight.get("flags", [0] * 13))
    p = _params(keep, flight, ncols)
    io = _ions(keep, ions)
    o, res = _out(io.n, p.rec_rows, ncols)
    rc = lib().or_rz_fly(C.byref(f), C.byref(p), C.byref(io), C.byref(o),
                         n_threads or os.cpu_count())
    if rc != 0:
        raise RuntimeError(f"or_rz_fly failed ({rc})")
    return res


def fly_xyz(fields, flight, ions, n_threads=None):
    """Fly ions through a 3-D field pack (reference layout dict: EAx, EAy,
    EAz, ExK, EyK, EzK, ele (bool), h_mm, ch_kind, ch_om, ch_ph, ch_amp,
    ch_off, ch_duty, [tab_t, tab_v, tab_off]).  ``flight`` dt/t_max are in
    seconds (RK4) or, with SDS gas (model 2, the _fly3d_sds flight), in
    microseconds; ion velocities in mm/us; ``acc`` carries qm."""
    keep = _Keep()
    f = XYZFields()
    EAx = keep(fields["EAx"])
    f.nx, f.ny, f.nz = EAx.shape
    f.EAx = _p(EAx)
    f.EAy = _p(keep(fields["EAy"]))
    f.EAz = _p(keep(fields["EAz"]))
    K = int(np.asarray(fields["ch_kind"]).shape[0])
    f.K = K
    for k in ("ExK", "EyK", "EzK"):
        setattr(f, k, _p(keep(fields[k] if K else np.zeros(1))))
    f.ele = _p(keep(fields["ele"], np.uint8), _u8p)
    f.h_mm = float(fields["h_mm"])
    f.ch_kind = _p(keep(fields["ch_kind"] if K else np.zeros(1), np.int64),
                   _i64p)
    for k in ("ch_om", "ch_ph", "ch_amp", "ch_off", "ch_duty"):
        setattr(f, k, _p(keep(fields[k] if K else np.zeros(1))))
    f.tab_t = _p(keep(fields.get("tab_t", np.zeros(1))))
    f.tab_v = _p(keep(fields.get("tab_v", np.zeros(1))))
    f.tab_off = _p(keep(fields.get("tab_off", np.zeros(K + 1)), np.int64),
                   _i64p)
    cols = 11 if flight.get("transporter") is not None else 10
    p = _params(keep, flight, cols)
    io = _ions(keep, ions)
    o, res = _out(io.n, p.rec_rows, cols)
    rc = lib().or_xyz_fly(C.byref(f), C.byref(p), C.byref(io), C.byref(o),
                          n_threads or os.cpu_count())
    if rc != 0:
        raise RuntimeError(f"or_xyz_fly failed ({rc})")
    return res


PLANAR_FLAGS = ("speed", "ke_ev", "ke_x", "ke_y", "ke_z", "e_field",
                "e_axial", "e_radial", "e_x", "e_y", "e_z", "radius", "n_col")


def fly_planar(fields, flight, ions, n_threads=None):
    """Fly ions through a planar field pack (reference layout dict: ExA,
    EyA, ExK, EyK (V/m), ele, mm, ch_kind, ch_om, ch_ph, ch_duty, [tab_t,
    tab_v, tab_off]).  ``flight["flags"]`` are the 13 planar channel flags
    in PLANAR_FLAGS order; dt/t_max in us."""
    keep = _Keep()
    f = PlanarFields()
    ExA = keep(fields["ExA"])
    f.nx, f.ny = ExA.shape
    f.ExA = _p(ExA)
    f.EyA = _p(keep(fields["EyA"]))
    K = int(np.asarray(fields["ch_kind"]).shape[0])
    f.K = K
    f.ExK = _p(keep(fields["ExK"] if K else np.zeros(1)))
    f.EyK = _p(keep(fields["EyK"] if K else np.zeros(1)))
    f.ele = _p(keep(fields["ele"]))
    f.mm = float(fields["mm"])
    f.ch_kind = _p(keep(fields["ch_kind"] if K else np.zeros(1), np.int64),
                   _i64p)
    for k in ("ch_om", "ch_ph", "ch_duty"):
        setattr(f, k, _p(keep(fields[k] if K else np.zeros(1))))
    f.tab_t = _p(keep(fields.get("tab_t", np.zeros(1))))
    f.tab_v = _p(keep(fields.get("tab_v", np.zeros(1))))
    f.tab_off = _p(keep(fields.get("tab_off", np.zeros(K + 1)), np.int64),
                   _i64p)
    ncols = 7 + sum(bool(v) for v in flight.get("flags", [0] * 13))
    p = _params(keep, flight, ncols)
    io = _ions(keep, ions)
    o, res = _out(io.n, p.rec_rows, ncols)
    rc = lib().or_planar_fly(C.byref(f), C.byref(p), C.byref(io),
                             C.byref(o), n_threads or os.cpu_count())
    if rc != 0:
        raise RuntimeError(f"or_planar_fly failed ({rc})")
    return res
