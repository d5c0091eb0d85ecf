"""Host-side engine: field packs, device-resident ensembles, flights.

Thin Python layer over the C ABI (include/ion_gym_amd.h).  It takes the
reference's own array layouts -- the r-z pack that
ion_gym/physics/build_rz.py:725-738 hands _fly_rec_full and the 3-D pack of
ion_gym/physics/tracer3d.py:281-290 -- converts them into the engine's
descriptor and returns numpy results.
"""
import ctypes as C

import numpy as np

from . import _lib as L

# r-z optional record channels, in the kernel's flag order
# (ion_gym/physics/build_rz.py:569-571)
RZ_CHANNELS = ("speed", "ke_ev", "ke_x", "ke_y", "ke_z", "e_field",
               "e_axial", "e_radial", "radius", "n_col", "path_mm",
               "e_axial_tint", "ke_tint")
BASE_CHANNELS = ["t", "x", "y", "z", "vx", "vy", "vz"]
XYZ_REC_COLS = ["t", "x", "y", "z", "vx", "vy", "vz", "ex", "ey", "ez"]


# the planar route's optional channels in its record order
# (ion_gym/physics/build_planar.py:1162-1214)
PLANAR_CHANNELS = ("speed", "ke_ev", "ke_x", "ke_y", "ke_z", "e_field",
                   "e_axial", "e_radial", "e_x", "e_y", "e_z", "radius",
                   "n_col")
_CHANNEL_BIT = {c: b for b, c in enumerate(RZ_CHANNELS)}
_CHANNEL_BIT.update(e_x=13, e_y=14, e_z=15)


def rec_mask(channels):
    m = 0
    for c in channels:
        if c not in _CHANNEL_BIT:
            raise ValueError(f"record channel {c!r} is not one of "
                             f"{sorted(_CHANNEL_BIT)}")
        m |= 1 << _CHANNEL_BIT[c]
    return m


def mask_channels(mask):
    return [c for b, c in enumerate(RZ_CHANNELS) if mask >> b & 1]


class _Keep(list):
    def __call__(self, a, dtype=np.float64):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype)
        self.append(a)
        return a


class Bases:
    """Per-electrode potential bases resident on the device."""

    def __init__(self, engine, bid):
        self.engine = engine
        self.id = bid

    def release(self):
        if self.id is not None:
            L.check(L.lib().ig_bases_release(self.engine.h, self.id),
                    "ig_bases_release")
            self.id = None


class FieldPack:
    """A field pack resident on the device."""

    def __init__(self, engine, fid, route, precision, shape, n_chan):
        self.engine = engine
        self.id = fid
        self.route = route
        self.precision = precision
        self.shape = shape
        self.n_chan = n_chan
        # field sets as stored (linked drive channels may be folded)
        self.n_sets = self.layout()["n_sets"]

    @property
    def nbytes(self):
        out = C.c_int64()
        L.check(L.lib().ig_fields_bytes(self.engine.h, self.id,
                                        C.byref(out)), "ig_fields_bytes")
        return out.value

    def layout(self):
        """The stored layout: field sets, drive channels (after folding
        linked channels), values per node record, channels given."""
        v = [C.c_int32() for _ in range(4)]
        L.check(L.lib().ig_fields_layout(self.engine.h, self.id,
                                         *[C.byref(x) for x in v]),
                "ig_fields_layout")
        n_sets, n_chan, stride, given = (x.value for x in v)
        return dict(n_sets=n_sets, n_chan=n_chan, stride=stride,
                    n_chan_given=given, folded=n_chan != given)

    def release(self):
        if self.id is not None and self.engine.h:
            L.check(L.lib().ig_fields_release(self.engine.h, self.id),
                    "ig_fields_release")
        self.id = None


def _fold_mode(precision, fold):
    """The pack's fold mode: linked drive channels are folded by default in
    the FP32-basis mode (the only precision that folds); ``fold=False``
    keeps every set as given (a pack to compose into)."""
    if fold is None:
        fold = precision == "f32"
    return L.FOLD_LINKED if fold else L.FOLD_NONE


class Ensemble:
    """Ion state resident on the device across ig_fly calls."""

    def __init__(self, engine, eid, capacity, rec_rows, rec_cols):
        self.engine = engine
        self.id = eid
        self.capacity = capacity
        self.rec_rows = rec_rows
        self.rec_cols = rec_cols
        self.n = 0

    def load(self, ions):
        keep = _Keep()
        io = make_ions(keep, ions)
        L.check(L.lib().ig_ensemble_load(self.engine.h, self.id,
                                         C.byref(io)), "ig_ensemble_load")
        self.n = io.n

    def read(self, records=True):
        return read_results(self.engine, self.id, self.n,
                            self.rec_rows if records else 0, self.rec_cols)

    def live(self):
        out = C.c_int64()
        L.check(L.lib().ig_ensemble_live(self.engine.h, self.id,
                                         C.byref(out)), "ig_ensemble_live")
        return out.value

    def steps(self):
        out = C.c_int64()
        L.check(L.lib().ig_ensemble_steps(self.engine.h, self.id,
                                          C.byref(out)), "ig_ensemble_steps")
        return out.value

    def release(self):
        if self.id is not None and self.engine.h:
            L.check(L.lib().ig_ensemble_release(self.engine.h, self.id),
                    "ig_ensemble_release")
        self.id = None


_BIRTH_KEYS = ("x", "y", "z", "vx", "vy", "vz", "tob")


def make_ions(keep, ions):
    """The C ABI's ig_ions for a dict of per-ion arrays.  Without any of
    the birth keys (x..vz, tob) the slots' births stay zero, to be drawn
    on the device (ion_gym_amd.births.births_into)."""
    no_births = all(ions.get(k) is None for k in _BIRTH_KEYS)
    n = int(len(ions["m_amu"] if no_births else ions["x"]))
    io = L.Ions()
    io.n = n
    for k in ("x", "y", "z", "vx", "vy", "vz", "tob", "m_amu", "acc"):
        a = ions.get(k)
        if a is None:
            if no_births and k in _BIRTH_KEYS:
                continue
            if k == "tob":
                a = np.zeros(n)
            else:
                raise ValueError(f"ions[{k!r}] is required")
        a = keep(a)
        if a.shape != (n,):
            raise ValueError(f"ions[{k!r}] has shape {a.shape}, want ({n},)")
        setattr(io, k, L.ptr(a))
    io.sigma_m2 = L.ptr(keep(ions.get("sigma_m2")))
    seed = keep(ions["seed"], np.int64)
    if seed.shape != (n,):
        raise ValueError("ions['seed'] must have one entry per ion")
    io.seed = L.ptr(seed, L._i64p)
    sds = ions.get("sds")
    if sds is not None:
        sds = keep(sds)
        if sds.shape != (n, 4):
            raise ValueError("ions['sds'] must be (n, 4)")
    io.sds = L.ptr(sds)
    fin = ions.get("fate_in")
    io.fate_in = L.ptr(keep(fin, np.int32), L._i32p) if fin is not None \
        else None
    return io


def read_results(engine, eid, n, rec_rows, rec_cols):
    res = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz",
                                    "t_us")}
    res.update(fate=np.zeros(n, np.int32), face=np.zeros(n, np.int32),
               label=np.zeros(n, np.int32), ncol=np.zeros(n, np.int64),
               steps=np.zeros(n, np.int64), nrec=np.zeros(n, np.int32))
    res["records"] = (np.zeros((n, rec_rows, rec_cols)) if rec_rows > 0
                      else None)
    o = _results_struct(res)
    L.check(L.lib().ig_ensemble_read(engine.h, eid, C.byref(o)),
            "ig_ensemble_read")
    return res


def _results_struct(res):
    o = L.Results()
    for k in ("x", "y", "z", "vx", "vy", "vz", "t_us"):
        setattr(o, k, L.ptr(res[k]))
    for k in ("fate", "face", "label", "nrec"):
        setattr(o, k, L.ptr(res[k], L._i32p))
    o.ncol = L.ptr(res["ncol"], L._i64p)
    o.steps = L.ptr(res["steps"], L._i64p)
    o.records = L.ptr(res["records"])
    return o


class Flight:
    """Flight parameters (the kernel scalars of tracer_rz.py:114-123 /
    tracer3d.py:281-290).  dt/t_max are in the route's clock: r-z us,
    3-D s."""

    def __init__(self, dt, t_max, *, max_steps=0, rec_every=1, rec_rows=0,
                 channels=(), gas=None, bounds=None, planes=(), rng_key=0,
                 coll_batch=0, occupancy=0, block=0, compact=3,
                 transporter=None):
        self.dt = float(dt)
        self.t_max = float(t_max)
        self.max_steps = int(max_steps)
        self.rec_every = int(rec_every)
        self.rec_rows = int(rec_rows)
        self.channels = list(channels)
        self.gas = dict(gas or {})
        self.bounds = bounds
        self.planes = list(planes)
        self.rng_key = int(rng_key)
        self.coll_batch = int(coll_batch)
        self.occupancy = int(occupancy)
        self.block = int(block)
        # live-ion list before each window: 0 none, 1 in slot order, 2 in
        # cell order, 3 the engine's choice per route
        self.compact = int(compact)
        # the 3-D periodic transporter (tracer3d.py:722-735): a dict with
        # axis ('x'|'y'|'z'), accept_mm, emit_mm, direction (default -1)
        # and max_passes (default 10**9), or None
        self.transporter = None if transporter is None else dict(transporter)

    def launch(self, **kw):
        """A copy with other launch settings (block, occupancy,
        coll_batch); results do not depend on them."""
        import copy
        f = copy.copy(self)
        for k, v in kw.items():
            if k not in ("block", "occupancy", "coll_batch", "compact"):
                raise ValueError(f"not a launch setting: {k}")
            setattr(f, k, int(v))
        return f

    def struct(self, keep):
        p = L.FlightParams()
        p.dt = self.dt
        p.t_max = self.t_max
        p.max_steps = self.max_steps
        p.rec_every = self.rec_every
        p.rec_rows = self.rec_rows
        p.rec_mask = rec_mask(self.channels)
        g = self.gas
        p.gas_model = int(g.get("model", L.GAS_VACUUM))
        for k in ("T_k", "P_pa", "c_star", "c_bar", "sig1d", "m_gas"):
            setattr(p, k, float(g.get(k, 0.0)))
        for i, v in enumerate(g.get("flow", (0.0, 0.0, 0.0))):
            p.flow[i] = float(v)
        st = g.get("sds_stats")
        if st is not None:
            st = keep(st)
            if st.shape != (5, 1002):
                raise ValueError("sds_stats must be (5, 1002)")
        p.sds_stats = L.ptr(st)
        if self.bounds is not None:
            on, val = self.bounds
            for i in range(6):
                p.bnd_on[i] = int(bool(on[i]))
                p.bnd_val[i] = float(val[i])
        if len(self.planes) > L.MAX_PLANES:
            raise ValueError(f"at most {L.MAX_PLANES} planes")
        p.n_planes = len(self.planes)
        for i, (col, val, sgn, win, kind) in enumerate(self.planes):
            p.pl_col[i] = int(col)
            p.pl_val[i] = float(val)
            p.pl_sgn[i] = float(sgn)
            for c in range(4):
                p.pl_w[i][c] = float(win[c])
            p.pl_kind[i] = int(kind)
        p.rng_key = self.rng_key & 0xFFFFFFFFFFFFFFFF
        p.coll_batch = self.coll_batch
        p.occupancy = self.occupancy
        p.block = self.block
        p.compact = int(self.compact)
        tp = self.transporter
        if tp is not None:
            p.tp_axis = 1 + {"x": 0, "y": 1, "z": 2}[str(tp["axis"])]
            p.tp_accept_mm = float(tp["accept_mm"])
            p.tp_emit_mm = float(tp["emit_mm"])
            p.tp_direction = float(tp.get("direction", -1))
            p.tp_max_passes = int(tp.get("max_passes", 10 ** 9))
        return p


class Engine:
    """One device plus one HIP stream (ig_engine)."""

    def __init__(self, device=0):
        self.h = None
        h = C.c_void_p()
        L.check(L.lib().ig_engine_create(int(device), C.byref(h)),
                "ig_engine_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            L.lib().ig_engine_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter teardown
            pass

    def sync(self):
        L.check(L.lib().ig_sync(self.h), "ig_sync")

    # ------------------------------------------------------------ fields
    def upload_rz(self, pack, precision="f64", chan_scale=None, fold=None):
        """Upload an r-z pack in the reference's layout: EzA, EuA
        (nx, nu) V/m; EzK, EuK (K, nx, nu); optional EzG, EuG, tau_gate;
        ele (nx, nu) mirrored metal; optional label (nx, nu); u0; mm;
        ch_kind, ch_om, ch_ph, ch_duty (ion_gym/physics/build_rz.py:
        725-738)."""
        keep = _Keep()
        EzA = np.asarray(pack["EzA"], np.float64)
        nx, nu = EzA.shape
        K = int(len(pack["ch_kind"]))
        tau = float(pack.get("tau_gate", -1.0))
        has_gate = (tau >= 0.0 and pack.get("EzG") is not None)
        sets = [(EzA, pack["EuA"])]
        for k in range(K):
            sets.append((pack["EzK"][k], pack["EuK"][k]))
        if has_gate:
            sets.append((pack["EzG"], pack["EuG"]))
        grids = np.empty((len(sets), 2, nx, nu), np.float64)
        for s, (a, b) in enumerate(sets):
            grids[s, 0] = a
            grids[s, 1] = b
        d = L.FieldsDesc()
        d.route = L.ROUTE_RZ
        d.precision = L.PRECISIONS[precision]
        d.n0, d.n1, d.n2 = nx, nu, 1
        d.h_mm = float(pack["mm"])
        d.u0 = float(pack["u0"])
        d.n_chan = K
        d.has_gate = int(has_gate)
        d.tau_gate = tau if has_gate else -1.0
        d.grids = L.ptr(keep(grids))
        d.chan_scale = L.ptr(keep(chan_scale))
        d.metal = L.ptr(keep(np.asarray(pack["ele"]) > 0.5, np.uint8),
                        L._u8p)
        if pack.get("label") is not None:
            d.label = L.ptr(keep(pack["label"], np.uint8), L._u8p)
        self._waves(d, keep, pack, K, affine=False)
        d.fold_channels = _fold_mode(precision, fold)
        fid = C.c_int64()
        L.check(L.lib().ig_fields_upload(self.h, C.byref(d), C.byref(fid)),
                "ig_fields_upload")
        return FieldPack(self, fid.value, L.ROUTE_RZ, precision, (nx, nu),
                         K)

    def upload_planar(self, pack, precision="f64", chan_scale=None,
                      fold=None):
        """Upload a planar pack in the reference's layout: ExA, EyA
        (nx, ny) V/m; ExK, EyK (K, nx, ny); ele (nx, ny) metal; mm;
        ch_kind, ch_om, ch_ph, ch_duty; optional tab_t, tab_v, tab_off
        (ion_gym/physics/build_planar.py:1111-1118)."""
        keep = _Keep()
        ExA = np.asarray(pack["ExA"], np.float64)
        nx, ny = ExA.shape
        K = int(len(pack["ch_kind"]))
        grids = np.empty((1 + K, 2, nx, ny), np.float64)
        grids[0, 0], grids[0, 1] = ExA, pack["EyA"]
        for k in range(K):
            grids[1 + k, 0] = pack["ExK"][k]
            grids[1 + k, 1] = pack["EyK"][k]
        d = L.FieldsDesc()
        d.route = L.ROUTE_PLANAR
        d.precision = L.PRECISIONS[precision]
        d.n0, d.n1, d.n2 = nx, ny, 1
        d.h_mm = float(pack["mm"])
        d.u0 = 0.0
        d.n_chan = K
        d.has_gate = 0
        d.tau_gate = -1.0
        d.grids = L.ptr(keep(grids))
        del grids
        d.chan_scale = L.ptr(keep(chan_scale))
        d.metal = L.ptr(keep(np.asarray(pack["ele"]) > 0.5, np.uint8),
                        L._u8p)
        self._waves(d, keep, pack, K, affine=False)
        d.fold_channels = _fold_mode(precision, fold)
        fid = C.c_int64()
        L.check(L.lib().ig_fields_upload(self.h, C.byref(d), C.byref(fid)),
                "ig_fields_upload")
        return FieldPack(self, fid.value, L.ROUTE_PLANAR, precision,
                         (nx, ny), K)

    def upload_xyz(self, pack, precision="f64", chan_scale=None, fold=None):
        """Upload a 3-D pack in the reference's layout: EAx, EAy, EAz
        (nx, ny, nz) V/mm; ExK, EyK, EzK (K, nx, ny, nz); ele (bool); h_mm;
        ch_kind, ch_om, ch_ph, ch_amp, ch_off, ch_duty; optional tab_t,
        tab_v, tab_off (ion_gym/physics/tracer3d.py:281-290)."""
        keep = _Keep()
        EAx = np.asarray(pack["EAx"], np.float64)
        nx, ny, nz = EAx.shape
        K = int(len(pack["ch_kind"]))
        grids = np.empty((1 + K, 3, nx, ny, nz), np.float64)
        grids[0, 0], grids[0, 1], grids[0, 2] = EAx, pack["EAy"], pack["EAz"]
        for k in range(K):
            grids[1 + k, 0] = pack["ExK"][k]
            grids[1 + k, 1] = pack["EyK"][k]
            grids[1 + k, 2] = pack["EzK"][k]
        d = L.FieldsDesc()
        d.route = L.ROUTE_XYZ
        d.precision = L.PRECISIONS[precision]
        d.n0, d.n1, d.n2 = nx, ny, nz
        d.h_mm = float(pack["h_mm"])
        d.u0 = 0.0
        d.n_chan = K
        d.has_gate = 0
        d.tau_gate = -1.0
        d.grids = L.ptr(keep(grids))
        del grids
        d.chan_scale = L.ptr(keep(chan_scale))
        d.metal = L.ptr(keep(np.asarray(pack["ele"]).astype(bool),
                             np.uint8), L._u8p)
        self._waves(d, keep, pack, K, affine=True)
        d.fold_channels = _fold_mode(precision, fold)
        fid = C.c_int64()
        L.check(L.lib().ig_fields_upload(self.h, C.byref(d), C.byref(fid)),
                "ig_fields_upload")
        return FieldPack(self, fid.value, L.ROUTE_XYZ, precision,
                         (nx, ny, nz), K)

    @staticmethod
    def _waves(d, keep, pack, K, affine):
        if K > L.MAX_CHANNELS:
            raise ValueError(f"at most {L.MAX_CHANNELS} drive channels")
        if K == 0:
            return
        d.ch_kind = L.ptr(keep(pack["ch_kind"], np.int32), L._i32p)
        d.ch_om = L.ptr(keep(pack["ch_om"]))
        d.ch_ph = L.ptr(keep(pack["ch_ph"]))
        d.ch_duty = L.ptr(keep(pack.get("ch_duty", np.full(K, 0.5))))
        if affine:
            d.ch_amp = L.ptr(keep(pack.get("ch_amp", np.ones(K))))
            d.ch_off = L.ptr(keep(pack.get("ch_off", np.zeros(K))))
        tab_t = pack.get("tab_t")
        if tab_t is not None and len(tab_t) > 0:
            d.tab_t = L.ptr(keep(tab_t))
            d.tab_v = L.ptr(keep(pack["tab_v"]))
            d.tab_off = L.ptr(keep(pack["tab_off"], np.int64), L._i64p)
            d.n_tab = len(tab_t)

    # --------------------------------------------------------- ensembles
    def ensemble(self, capacity, rec_rows=0, rec_cols=0):
        eid = C.c_int64()
        L.check(L.lib().ig_ensemble_create(self.h, int(capacity),
                                           int(rec_rows), int(rec_cols),
                                           C.byref(eid)),
                "ig_ensemble_create")
        return Ensemble(self, eid.value, int(capacity), int(rec_rows),
                        int(rec_cols))

    def fly(self, fields, flight, ens):
        """Advance every live ion of ``ens`` (asynchronous)."""
        keep = _Keep()
        p = flight.struct(keep)
        L.check(L.lib().ig_fly(self.h, fields.id, C.byref(p), ens.id),
                "ig_fly")

    def fly_noalloc(self, fields, flight, ens):
        """fly() without device allocation (ig_fly_noalloc): the window's
        scratch is the device's fixed region; at most NOALLOC_MAX_IONS ions
        per ensemble, and calls overlapping on one device race."""
        keep = _Keep()
        p = flight.struct(keep)
        L.check(L.lib().ig_fly_noalloc(self.h, fields.id, C.byref(p),
                                       ens.id), "ig_fly_noalloc")

    # ----------------------------------------------- field composition
    def upload_bases_rz(self, phi, ele, Z_mm, U_mm, base_volts=1e4):
        """Per-electrode r-z potentials for composition on the device:
        ``phi`` (n_bases, nz, nr) at ``base_volts`` each (the reference's
        cache convention), ``ele`` (nz, nr) labels or mask (nonzero =
        metal), ``Z_mm``/``U_mm`` the half-grid coordinates
        (ionbench.build_field_aware's inputs)."""
        keep = _Keep()
        phi = keep(np.asarray(phi, np.float64))
        nb, nz, nr = phi.shape
        Ue = np.concatenate([-np.asarray(U_mm)[:0:-1], np.asarray(U_mm)])
        d = L.BasesDesc()
        d.route = L.ROUTE_RZ
        d.n0, d.n1, d.n2 = nz, nr, 1
        d.n_bases = nb
        d.phi = L.ptr(phi)
        d.metal = L.ptr(keep(np.asarray(ele) != 0, np.uint8), L._u8p)
        d.base_volts = float(base_volts)
        d.coord0 = L.ptr(keep(np.asarray(Z_mm, np.float64)))
        d.coord1 = L.ptr(keep(np.asarray(Ue, np.float64)))
        bid = C.c_int64()
        L.check(L.lib().ig_bases_upload(self.h, C.byref(d), C.byref(bid)),
                "ig_bases_upload")
        return Bases(self, bid.value)

    def upload_bases_3d(self, phi, ele, h_mm):
        """Per-electrode 3-D unit potentials (n_bases, nx, ny, nz) with the
        node metal mask and spacing (build_field_aware_3d's inputs)."""
        keep = _Keep()
        phi = keep(np.asarray(phi, np.float64))
        nb, nx, ny, nz = phi.shape
        d = L.BasesDesc()
        d.route = L.ROUTE_XYZ
        d.n0, d.n1, d.n2 = nx, ny, nz
        d.n_bases = nb
        d.phi = L.ptr(phi)
        d.metal = L.ptr(keep(np.asarray(ele) != 0, np.uint8), L._u8p)
        d.base_volts = 1.0
        d.h_mm = float(h_mm)
        bid = C.c_int64()
        L.check(L.lib().ig_bases_upload(self.h, C.byref(d), C.byref(bid)),
                "ig_bases_upload")
        return Bases(self, bid.value)

    def compose(self, bases, fields, set_index, terms):
        """Set ``set_index`` of ``fields`` := -grad(sum of w * phi_b /
        base_volts over ``terms`` [(b, w), ...] in order), on the device.
        ``fields`` keeps its sets as given: an FP32-basis pack to compose
        into is uploaded with ``fold=False``."""
        keep = _Keep()
        bi = keep(np.array([t[0] for t in terms], np.int32), np.int32)
        w = keep(np.array([t[1] for t in terms], np.float64))
        L.check(L.lib().ig_fields_compose(
            self.h, bases.id, fields.id, int(set_index), len(terms),
            L.ptr(bi, L._i32p), L.ptr(w)), "ig_fields_compose")

    def read_set(self, fields, set_index, comp):
        """One component grid of one stored set of a resident pack
        (doubles; a folded pack's sets as stored, ``fields.layout()``)."""
        shape = tuple(int(v) for v in fields.shape)
        out = np.empty(shape, np.float64)
        L.check(L.lib().ig_fields_read_set(self.h, fields.id, int(set_index),
                                           int(comp), L.ptr(out)),
                "ig_fields_read_set")
        return out

    def fly_graph(self, fields, flight, ens, n_windows):
        """Advance ``ens`` by ``n_windows`` windows of ``flight.max_steps``
        steps as one replay of a captured HIP graph (captured on the first
        call with these arguments; asynchronous)."""
        keep = _Keep()
        p = flight.struct(keep)
        L.check(L.lib().ig_fly_graph(self.h, fields.id, C.byref(p), ens.id,
                                     int(n_windows)), "ig_fly_graph")

    def last_fly_ms(self):
        out = C.c_double()
        L.check(L.lib().ig_last_fly_ms(self.h, C.byref(out)),
                "ig_last_fly_ms")
        return out.value

    def fly_batch(self, fields, flight, ions):
        """Load, fly to fate and read one batch of ions (ig_fly_batch)."""
        keep = _Keep()
        p = flight.struct(keep)
        io = make_ions(keep, ions)
        n = io.n
        if fields.route in (L.ROUTE_RZ, L.ROUTE_PLANAR):
            cols = 7 + len(set(flight.channels))
        else:
            cols = 11 if flight.transporter is not None else 10
        rows = flight.rec_rows
        res = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz",
                                        "t_us")}
        res.update(fate=np.zeros(n, np.int32), face=np.zeros(n, np.int32),
                   label=np.zeros(n, np.int32), ncol=np.zeros(n, np.int64),
                   steps=np.zeros(n, np.int64), nrec=np.zeros(n, np.int32))
        res["records"] = np.zeros((n, rows, cols)) if rows > 0 else None
        o = _results_struct(res)
        L.check(L.lib().ig_fly_batch(self.h, fields.id, C.byref(p),
                                     C.byref(io), C.byref(o)),
                "ig_fly_batch")
        return res


def device_count():
    n = C.c_int()
    L.check(L.lib().ig_device_count(C.byref(n)), "ig_device_count")
    return n.value


def device_info(device=0):
    """(name, compute units) of a device."""
    buf = C.create_string_buffer(256)
    cu = C.c_int32()
    L.check(L.lib().ig_device_info(int(device), buf, 256, C.byref(cu)),
            "ig_device_info")
    return buf.value.decode(errors="replace"), int(cu.value)


def read_slice(engine, ens, off, n):
    """Final state of ions [off, off + n) of an ensemble (no records)."""
    res = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz", "t_us")}
    res.update(fate=np.zeros(n, np.int32), face=np.zeros(n, np.int32),
               label=np.zeros(n, np.int32), ncol=np.zeros(n, np.int64),
               steps=np.zeros(n, np.int64), nrec=np.zeros(n, np.int32))
    res["records"] = None
    o = _results_struct(res)
    L.check(L.lib().ig_ensemble_read_at(engine.h, ens.id, int(off), int(n),
                                        o), "ig_ensemble_read_at")
    return res
