"""ctypes binding of libion_gym_amd.so (include/ion_gym_amd.h).

The library is built in-tree by ``ion_gym_amd._build``.  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in the package.
ctypes releases the GIL for the duration of every foreign call, so worker
threads of the ensemble driver overlap with device work.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libion_gym_amd.so")

IG_OK = 0
ROUTE_RZ, ROUTE_XYZ, ROUTE_PLANAR = 0, 1, 2
PREC_F64, PREC_F32_STORE, PREC_F32 = 0, 1, 2
PRECISIONS = {"f64": PREC_F64, "f32_store": PREC_F32_STORE, "f32": PREC_F32}
FOLD_NONE, FOLD_LINKED = 0, 1
WAVE_SIN, WAVE_COS, WAVE_SQUARE, WAVE_TAB_HOLD, WAVE_TAB_LINEAR = range(5)
GAS_VACUUM, GAS_HS, GAS_SDS = 0, 1, 2
FATE_FLYING = -2
FACE_BIRTH = -2
FATE_EMPTY = -4
MAX_CHANNELS = 16
MAX_PLANES = 16
NOALLOC_MAX_IONS = 1 << 20       # IG_NOALLOC_MAX_IONS

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class FieldsDesc(C.Structure):
    _fields_ = [("route", C.c_int32), ("precision", C.c_int32),
                ("n0", C.c_int64), ("n1", C.c_int64), ("n2", C.c_int64),
                ("h_mm", C.c_double), ("u0", C.c_double),
                ("n_chan", C.c_int32), ("has_gate", C.c_int32),
                ("tau_gate", C.c_double), ("grids", _dp),
                ("chan_scale", _dp), ("metal", _u8p), ("label", _u8p),
                ("ch_kind", _i32p), ("ch_om", _dp), ("ch_ph", _dp),
                ("ch_duty", _dp), ("ch_amp", _dp), ("ch_off", _dp),
                ("tab_t", _dp), ("tab_v", _dp), ("tab_off", _i64p),
                ("n_tab", C.c_int64), ("fold_channels", C.c_int32)]


class BasesDesc(C.Structure):
    _fields_ = [("route", C.c_int32), ("n0", C.c_int64), ("n1", C.c_int64),
                ("n2", C.c_int64), ("n_bases", C.c_int32), ("phi", _dp),
                ("metal", _u8p), ("base_volts", C.c_double),
                ("coord0", _dp), ("coord1", _dp), ("h_mm", C.c_double)]


class FlightParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("t_max", C.c_double),
                ("max_steps", C.c_int64), ("rec_every", C.c_int32),
                ("rec_rows", C.c_int32), ("rec_mask", C.c_uint32),
                ("gas_model", C.c_int32), ("T_k", C.c_double),
                ("P_pa", C.c_double), ("c_star", C.c_double),
                ("c_bar", C.c_double), ("sig1d", C.c_double),
                ("m_gas", C.c_double), ("flow", C.c_double * 3),
                ("sds_stats", _dp), ("bnd_on", C.c_int32 * 6),
                ("bnd_val", C.c_double * 6), ("n_planes", C.c_int32),
                ("pl_col", C.c_int32 * MAX_PLANES),
                ("pl_val", C.c_double * MAX_PLANES),
                ("pl_sgn", C.c_double * MAX_PLANES),
                ("pl_w", (C.c_double * 4) * MAX_PLANES),
                ("pl_kind", C.c_int32 * MAX_PLANES),
                ("rng_key", C.c_uint64), ("coll_batch", C.c_int32),
                ("occupancy", C.c_int32), ("block", C.c_int32),
                ("compact", C.c_int32), ("tp_axis", C.c_int32),
                ("tp_accept_mm", C.c_double), ("tp_emit_mm", C.c_double),
                ("tp_direction", C.c_double), ("tp_max_passes", C.c_int64)]


class Tune(C.Structure):
    _fields_ = [("dt", C.c_double), ("t_max", C.c_double),
                ("rng_key", C.c_uint64),
                ("set_scale", C.c_double * (MAX_CHANNELS + 2)),
                ("ch_kind", C.c_int32 * MAX_CHANNELS),
                ("ch_om", C.c_double * MAX_CHANNELS),
                ("ch_ph", C.c_double * MAX_CHANNELS),
                ("ch_duty", C.c_double * MAX_CHANNELS),
                ("ch_amp", C.c_double * MAX_CHANNELS),
                ("ch_off", C.c_double * MAX_CHANNELS)]


class Ions(C.Structure):
    _fields_ = [("n", C.c_int64)] + [
        (k, _dp) for k in ("x", "y", "z", "vx", "vy", "vz", "tob", "m_amu",
                           "acc", "sigma_m2")] + [
        ("seed", _i64p), ("sds", _dp), ("fate_in", _i32p)]


class Results(C.Structure):
    _fields_ = [(k, _dp) for k in ("x", "y", "z", "vx", "vy", "vz",
                                   "t_us")] + [
        ("fate", _i32p), ("face", _i32p), ("label", _i32p),
        ("ncol", _i64p), ("steps", _i64p), ("nrec", _i32p),
        ("records", _dp)]


BIRTH_POINT, BIRTH_DISC, BIRTH_LINE, BIRTH_GRID, BIRTH_BOX, \
    BIRTH_GAUSSIAN = range(6)
E_ARG = -1                      # IG_E_ARG
E_BIRTH_TRIES = -6
_u64p = C.POINTER(C.c_uint64)
_d3 = C.c_double * 3
_u64x2 = C.c_uint64 * 2


class BirthsDesc(C.Structure):
    _fields_ = [("distribution", C.c_int32), ("axis", C.c_int32),
                ("n_per", C.c_int64), ("n_mz", C.c_int32),
                ("mz_list", _dp), ("p0", _d3), ("r_mm", C.c_double),
                ("len_mm", C.c_double), ("box_mm", _d3),
                ("gauss_sigma", _d3), ("gauss_trunc", _d3),
                ("gauss_tries", C.c_int32), ("has_ke", C.c_int32),
                ("ke_lo", C.c_double), ("ke_hi", C.c_double),
                ("dir_unit", _d3), ("has_dir", C.c_int32),
                ("temperature_k", C.c_double), ("has_dv", C.c_int32),
                ("dv_sigma", _d3), ("has_drift", C.c_int32),
                ("v_drift", _d3), ("tob_span_us", C.c_double),
                ("pcg_state", _u64x2), ("pcg_inc", _u64x2),
                ("zig_k", _u64p), ("zig_w", _dp), ("zig_f", _dp)]


SOLVE_SOR, SOLVE_VCYCLE, SOLVE_CASCADE = 0, 1, 2
STENCIL_GHOST_LINEAR, STENCIL_SW = 0, 1
STENCILS = {"ghost_linear": STENCIL_GHOST_LINEAR, "sw": STENCIL_SW}
_fp = C.POINTER(C.c_float)


class Solve3dDesc(C.Structure):
    _fields_ = [("n", C.c_int64 * 3), ("fixed", _u8p), ("val", _dp),
                ("theta", _fp), ("phi0", _dp), ("method", C.c_int32),
                ("stencil", C.c_int32), ("tol", C.c_double),
                ("omega", C.c_double), ("max_sweeps", C.c_int64),
                ("max_cycles", C.c_int32), ("nu1", C.c_int32),
                ("nu2", C.c_int32), ("sym_axes", C.c_int32),
                ("min_coarse", C.c_int32), ("max_levels", C.c_int32),
                ("level_tol_factor", C.c_double)]


class Solve3dResult(C.Structure):
    _fields_ = [("sweeps", C.c_int64), ("delta", C.c_double),
                ("status", C.c_int32), ("level", C.c_int32)]


class Solve2dDesc(C.Structure):
    _fields_ = [("nz", C.c_int64), ("nu", C.c_int64), ("fixed", _u8p),
                ("val", _dp), ("phi0", _dp), ("h", C.c_double),
                ("symmetry", C.c_int32), ("neumann", C.c_int32),
                ("ghost_linear", C.c_int32), ("tol", C.c_double),
                ("omega", C.c_double), ("max_sweeps", C.c_int64)]


SYM_PLANAR, SYM_CYLINDRICAL = 0, 1
EDGE_BITS = {"z0": 1, "z1": 2, "u0": 4, "u1": 8}


# every entry point of include/ion_gym_amd.h: name -> (restype, argtypes)
_VP = C.c_void_p
SIGNATURES = {
    "ig_last_error": (C.c_char_p, []),
    "ig_abi_version": (C.c_int, []),
    "ig_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ig_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_int32,
                                 C.POINTER(C.c_int32)]),
    "ig_engine_create": (C.c_int, [C.c_int, C.POINTER(_VP)]),
    "ig_engine_destroy": (C.c_int, [_VP]),
    "ig_sync": (C.c_int, [_VP]),
    "ig_fields_upload": (C.c_int, [_VP, C.POINTER(FieldsDesc),
                                   C.POINTER(C.c_int64)]),
    "ig_fields_release": (C.c_int, [_VP, C.c_int64]),
    "ig_fields_bytes": (C.c_int, [_VP, C.c_int64, C.POINTER(C.c_int64)]),
    "ig_ensemble_create": (C.c_int, [_VP, C.c_int64, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_int64)]),
    "ig_ensemble_release": (C.c_int, [_VP, C.c_int64]),
    "ig_ensemble_load": (C.c_int, [_VP, C.c_int64, C.POINTER(Ions)]),
    "ig_ensemble_read": (C.c_int, [_VP, C.c_int64, C.POINTER(Results)]),
    "ig_ensemble_load_at": (C.c_int, [_VP, C.c_int64, C.c_int64,
                                      C.POINTER(Ions), C.c_int32]),
    "ig_ensemble_read_at": (C.c_int, [_VP, C.c_int64, C.c_int64, C.c_int64,
                                      C.POINTER(Results)]),
    "ig_ensemble_set_tunes": (C.c_int, [_VP, C.c_int64, C.c_int32,
                                        C.POINTER(Tune)]),
    "ig_ensemble_live_by_tune": (C.c_int, [_VP, C.c_int64, C.c_int32,
                                           C.POINTER(C.c_int64)]),
    "ig_ensemble_live": (C.c_int, [_VP, C.c_int64, C.POINTER(C.c_int64)]),
    "ig_ensemble_steps": (C.c_int, [_VP, C.c_int64, C.POINTER(C.c_int64)]),
    "ig_fly": (C.c_int, [_VP, C.c_int64, C.POINTER(FlightParams),
                         C.c_int64]),
    "ig_fly_batch": (C.c_int, [_VP, C.c_int64, C.POINTER(FlightParams),
                               C.POINTER(Ions), C.POINTER(Results)]),
    "ig_fly_graph": (C.c_int, [_VP, C.c_int64, C.POINTER(FlightParams),
                               C.c_int64, C.c_int32]),
    "ig_fly_noalloc": (C.c_int, [_VP, C.c_int64, C.POINTER(FlightParams),
                                 C.c_int64]),
    "ig_bases_upload": (C.c_int, [_VP, C.POINTER(BasesDesc),
                                  C.POINTER(C.c_int64)]),
    "ig_bases_release": (C.c_int, [_VP, C.c_int64]),
    "ig_fields_compose": (C.c_int, [_VP, C.c_int64, C.c_int64, C.c_int32,
                                    C.c_int32, _i32p, _dp]),
    "ig_fields_layout": (C.c_int, [_VP, C.c_int64, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]),
    "ig_fields_read_set": (C.c_int, [_VP, C.c_int64, C.c_int32, C.c_int32,
                                     _dp]),
    "ig_last_fly_ms": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "ig_births": (C.c_int, [_VP, C.POINTER(BirthsDesc), _dp]),
    "ig_ensemble_births_at": (C.c_int, [_VP, C.c_int64, C.c_int64,
                                        C.POINTER(BirthsDesc)]),
    "ig_ensemble_refuse_births": (C.c_int, [_VP, C.c_int64, C.c_int64,
                                            C.c_int64, C.c_int64,
                                            C.POINTER(C.c_int64)]),
    "ig_solve3d": (C.c_int, [_VP, C.POINTER(Solve3dDesc), _dp,
                             C.POINTER(Solve3dResult)]),
    "ig_solve2d": (C.c_int, [_VP, C.POINTER(Solve2dDesc), _dp,
                             C.POINTER(Solve3dResult)]),
}


class EngineError(RuntimeError):
    """A C-ABI call returned an error status."""


_lib = None


def lib():
    """The loaded library; raises ImportError naming the fix when absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built: run `python -c \"import "
                f"__graft_entry__ as g; g.build()\"` (or python "
                f"ion_gym_amd/_build.py). The engine has no CPU fallback.")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != IG_OK:
        msg = lib().ig_last_error().decode(errors="replace")
        raise EngineError(f"{what or 'ion_gym_amd'} failed ({rc}): {msg}")


def ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)
