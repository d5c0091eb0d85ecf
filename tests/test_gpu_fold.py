"""Linked drive channels folded at upload (IG_FOLD_LINKED, the FP32-basis
mode's default): the stored layouts of the configurations, and flights on
a folded pack against the same flights on the pack with every set kept,
with the FP64 flight as the reference.  Folding moves the field by the
FP32 rounding only, so a folded flight's error against FP64 stays within
the unfolded FP32 flight's own error (twice it, plus a floor of a few
FP32 ulps).  Tunes on a folded pack: a request's drive (its own TW
amplitude and speed) flies folded; a tune that breaks a linked group is
refused, and composing into a folded pack is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POS = ("x", "y", "z")


def _layout(fp):
    lay = fp.layout()
    return lay["n_sets"], lay["n_chan"], lay["n_chan_given"]


def _rel(a, b, scale):
    return np.abs(a - b) / np.maximum(np.abs(b), scale)


def _err(run, ref, scale):
    same = run["fate"] == ref["fate"]
    return max(float(np.max(_rel(run[k][same], ref[k][same], scale)))
               for k in POS), float(same.mean())


def _fly(engine, upload, pack, ions, flight, prec, fold, tunes=None,
         tune_of=None):
    from ion_gym_amd import _lib as L
    fp = upload(pack, prec, fold=fold)
    ens = engine.ensemble(len(ions["x"]))
    if tunes is not None:
        arr = (L.Tune * len(tunes))()
        from ion_gym_amd.serve import _fill_tune
        for t, d in enumerate(tunes):
            _fill_tune(arr[t], d)
        L.check(L.lib().ig_ensemble_set_tunes(engine.h, ens.id, len(tunes),
                                              arr))
    if tune_of is None:
        ens.load(ions)
    else:
        from ion_gym_amd.engine import _Keep, make_ions
        for t in range(len(tunes)):
            sel = np.flatnonzero(tune_of == t)
            keep = _Keep()
            io = make_ions(keep, {k: v[sel] for k, v in ions.items()})
            L.check(L.lib().ig_ensemble_load_at(engine.h, ens.id,
                                                int(sel[0]), io, t))
        ens.n = len(ions["x"])
    engine.fly(fp, flight, ens)
    out = ens.read(records=False)
    lay = _layout(fp)
    ens.release()
    fp.release()
    return out, lay


def test_stored_layouts(engine):
    """The funnel's opposite-phase RF pair and the quadrupole's rod pairs
    fold to one drive set each; the SLIM board's eight TW phases to four
    (phases pi apart) plus a constant set for their offsets, and its two
    RF guard phases to one: 11 sets -> 7.  FP64 and FP32-storage packs,
    and fold=False, keep every set."""
    from ion_gym_amd import configs
    f = configs.funnel_rz(n_ions=8, coarse=4)["model"].pack()
    for prec, fold, want in (("f32", None, (2, 1, 2)),
                             ("f64", None, (3, 2, 2)),
                             ("f32_store", None, (3, 2, 2)),
                             ("f32", False, (3, 2, 2))):
        fp = engine.upload_rz(f, prec, fold=fold)
        assert _layout(fp) == want, (prec, fold)
        fp.release()
    q = configs.quad_3d(n_ions=8, coarse=8)["pack"]
    fp = engine.upload_xyz(q, "f32")
    assert _layout(fp) == (2, 1, 2)
    fp.release()
    s = configs.slim_3d(coarse=8, workers=4)["pack"]
    fp = engine.upload_xyz(s, "f32")
    assert _layout(fp) == (7, 6, 10)
    fp.release()


def test_funnel_folded_flight(engine):
    """Funnel r-z, vacuum, 3000 RF-driven steps of 2000 ions."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.funnel_rz(n_ions=2000, seed=5, coarse=2)
    pack = c["model"].pack()
    fl = Flight(c["dt"], c["t_max"], max_steps=3000)
    up = engine.upload_rz
    print("funnel folded vs unfolded FP32, FP64 reference", flush=True)
    ref, _ = _fly(engine, up, pack, c["ions"], fl, "f64", None)
    plain, lp = _fly(engine, up, pack, c["ions"], fl, "f32", False)
    fold, lf = _fly(engine, up, pack, c["ions"], fl, "f32", True)
    assert lp == (3, 2, 2) and lf == (2, 1, 2)
    e_plain, s_plain = _err(plain, ref, 1.0)
    e_fold, s_fold = _err(fold, ref, 1.0)
    print(f"  error vs FP64: unfolded {e_plain:.3g}, folded {e_fold:.3g}",
          flush=True)
    assert s_fold >= s_plain - 0.002
    assert e_fold <= 2.0 * e_plain + 1e-6


def test_quadrupole_folded_flight(engine):
    """Quadrupole, vacuum, 40 RF cycles of 1000 ions."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.quad_3d(n_ions=1000, coarse=4, n_cycles=40)
    fl = Flight(c["dt"], c["t_max"], max_steps=10000)
    up = engine.upload_xyz
    print("quadrupole folded vs unfolded FP32, FP64 reference", flush=True)
    ref, _ = _fly(engine, up, c["pack"], c["ions"], fl, "f64", None)
    plain, _ = _fly(engine, up, c["pack"], c["ions"], fl, "f32", False)
    fold, lf = _fly(engine, up, c["pack"], c["ions"], fl, "f32", True)
    assert lf == (2, 1, 2)
    e_plain, s_plain = _err(plain, ref, 1.0)
    e_fold, s_fold = _err(fold, ref, 1.0)
    print(f"  error vs FP64: unfolded {e_plain:.3g}, folded {e_fold:.3g}",
          flush=True)
    assert s_fold >= s_plain - 0.002
    assert e_fold <= 2.0 * e_plain + 1e-6


def _slim_requests(c, n_req, n_each):
    from ion_gym_amd import configs
    rng = np.random.default_rng(11)
    ions, tunes = [], []
    for r in range(n_req):
        i, t = configs.slim_request(c, rng, n_ions=n_each, rng_key=r)
        ions.append(i)
        tunes.append(t)
    # one time step for all (the flight's), each request its own drive
    dt = min(t["dt"] for t in tunes)
    for t in tunes:
        t["dt"] = dt
    cat = {k: np.concatenate([i[k] for i in ions]) for k in ions[0]}
    return cat, tunes, np.repeat(np.arange(n_req), n_each), dt


def test_slim_folded_flight_with_request_tunes(engine):
    """SLIM board (coarse 4), vacuum, three requests in one ensemble, each
    with its own TW amplitude and speed (tunes re-laid onto the folded
    sets): 3000 steps."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.slim_3d(coarse=4, workers=4)
    ions, tunes, tof, dt = _slim_requests(c, 3, 700)
    fl = Flight(dt, 1.0, max_steps=3000)
    up = engine.upload_xyz
    print("SLIM folded vs unfolded FP32 with request tunes", flush=True)
    pk = c["pack"]
    ref, _ = _fly(engine, up, pk, ions, fl, "f64", None, tunes, tof)
    plain, _ = _fly(engine, up, pk, ions, fl, "f32", False, tunes, tof)
    fold, lf = _fly(engine, up, pk, ions, fl, "f32", True, tunes, tof)
    assert lf == (7, 6, 10)
    e_plain, s_plain = _err(plain, ref, 1.0)
    e_fold, s_fold = _err(fold, ref, 1.0)
    print(f"  error vs FP64: unfolded {e_plain:.3g}, folded {e_fold:.3g}",
          flush=True)
    assert s_fold >= s_plain - 0.002
    assert e_fold <= 2.0 * e_plain + 1e-6


def test_a_tune_breaking_a_linked_group_is_refused(engine):
    """A tune whose RF pair is not in the pack's amplitude ratio cannot fly
    the folded funnel (ig_fly says why); the unfolded pack flies it."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    from ion_gym_amd.serve import make_tune
    c = configs.funnel_rz(n_ions=64, seed=5, coarse=4)
    pack = c["model"].pack()
    bad = make_tune(pack, dt=c["dt"], t_max=c["t_max"],
                    set_scale=[1.0, 1.0, 0.5])
    good = make_tune(pack, dt=c["dt"], t_max=c["t_max"],
                     set_scale=[1.0, 0.5, 0.5])
    fl = Flight(c["dt"], c["t_max"], max_steps=50)
    up = engine.upload_rz
    tof = np.zeros(64, int)
    _fly(engine, up, pack, c["ions"], fl, "f32", True, [good], tof)
    _fly(engine, up, pack, c["ions"], fl, "f32", False, [bad], tof)
    with pytest.raises(RuntimeError, match="folded at upload"):
        _fly(engine, up, pack, c["ions"], fl, "f32", True, [bad], tof)


def test_composing_into_a_folded_pack_is_refused(engine):
    from ion_gym_amd import configs

    class NoBases:
        id = -1

    c = configs.funnel_rz(n_ions=8, coarse=4)
    fp = engine.upload_rz(c["model"].pack(), "f32")
    with pytest.raises(RuntimeError, match="folded"):
        engine.compose(NoBases(), fp, 1, [])
    fp.release()


def _rz_pack(kinds, oms, phs, duties, gate=False, seed=7):
    """A synthetic r-z pack with the given drive channels (random grids)."""
    nx, nr, h = 41, 11, 0.25
    nu = 2 * nr - 1
    rng = np.random.default_rng(seed)
    K = len(kinds)
    pack = dict(EzA=np.full((nx, nu), 3000.0) + rng.normal(size=(nx, nu)),
                EuA=rng.normal(size=(nx, nu)) * 200.0,
                EzK=rng.normal(size=(K, nx, nu)) * 500.0,
                EuK=rng.normal(size=(K, nx, nu)) * 500.0,
                ele=np.zeros((nx, nu)), label=np.zeros((nx, nu), np.uint8),
                u0=-(nr - 1) * h, mm=h, ch_kind=np.array(kinds),
                ch_om=np.array(oms, float), ch_ph=np.array(phs, float),
                ch_duty=np.array(duties, float), tau_gate=-1.0)
    if gate:
        pack.update(EzG=np.full((nx, nu), -2500.0), EuG=np.zeros((nx, nu)),
                    tau_gate=0.3)
    return pack


def test_what_folds_and_what_does_not(engine):
    """Linked channels fold whatever their position among the others;
    channels of different kinds, frequencies or quarter-period phases do
    not; a table waveform keeps the whole pack as given; the r-z gate set
    stays last."""
    import math
    w = 2 * math.pi
    cases = [
        # (kinds, oms, phases, duties, gate) -> (sets, drive channels)
        (([0, 0], [w, w], [0.0, math.pi], [0.5, 0.5], False), (2, 1)),
        (([0, 0], [w, w], [0.3, 0.3 + 3 * math.pi], [0.5] * 2, False),
         (2, 1)),                                   # odd multiple of pi
        (([0, 0], [w, w], [0.0, 2 * math.pi], [0.5] * 2, False), (2, 1)),
        (([2, 2], [w, w], [0.0, math.pi], [0.5] * 2, False), (2, 1)),
        (([2, 2], [w, w], [0.0, math.pi], [0.3, 0.3], False), (3, 2)),
        (([0, 1], [w, w], [0.0, math.pi], [0.5] * 2, False), (3, 2)),
        (([0, 0], [w, 2 * w], [0.0, math.pi], [0.5] * 2, False), (3, 2)),
        (([0, 0], [w, w], [0.0, math.pi / 2], [0.5] * 2, False), (3, 2)),
        (([1, 0, 0], [3 * w, w, w], [0.2, 0.1, 0.1 + math.pi],
          [0.5] * 3, False), (3, 2)),               # linked pair after another
        (([0, 0], [w, w], [0.0, math.pi], [0.5] * 2, True), (3, 1)),
    ]
    for (kinds, oms, phs, duties, gate), (n_sets, n_chan) in cases:
        fp = engine.upload_rz(_rz_pack(kinds, oms, phs, duties, gate), "f32")
        lay = fp.layout()
        assert (lay["n_sets"], lay["n_chan"]) == (n_sets, n_chan), \
            (kinds, phs, lay)
        fp.release()
    pk = _rz_pack([0, 3], [w, 0.0], [0.0, 0.0], [0.5, 0.5])
    pk.update(tab_t=np.array([0.0, 1.0]), tab_v=np.array([0.0, 1.0]),
              tab_off=np.array([0, 0, 2]))
    fp = engine.upload_rz(pk, "f32")
    assert fp.layout()["n_sets"] == 3
    fp.release()


def test_gated_pack_folded_flight(engine):
    """An r-z pack with a linked RF pair and a gate set: the folded flight
    against the unfolded one and FP64 (vacuum, 600 steps across the gate
    time)."""
    import math
    from ion_gym_amd.engine import Flight
    pack = _rz_pack([0, 0], [2 * math.pi * 3.0] * 2, [0.4, 0.4 + math.pi],
                    [0.5, 0.5], gate=True)
    rng = np.random.default_rng(3)
    n, m = 400, 150.0
    ions = dict(x=rng.uniform(1, 4, n), y=rng.uniform(-1.0, 1.0, n),
                z=rng.uniform(-1.0, 1.0, n), vx=rng.uniform(0.5, 2, n),
                vy=rng.normal(0, 0.2, n), vz=rng.normal(0, 0.2, n),
                tob=rng.uniform(0, 0.2, n), m_amu=np.full(n, m),
                acc=np.full(n, 1.602176634e-19 / (m * 1.66053906660e-27)),
                sigma_m2=np.zeros(n), seed=np.arange(n, dtype=np.int64))
    fl = Flight(1e-3, 8.0, max_steps=600)
    up = engine.upload_rz
    ref, _ = _fly(engine, up, pack, ions, fl, "f64", None)
    plain, lp = _fly(engine, up, pack, ions, fl, "f32", False)
    fold, lf = _fly(engine, up, pack, ions, fl, "f32", True)
    assert lp == (4, 2, 2) and lf == (3, 1, 2)
    e_plain, s_plain = _err(plain, ref, 1.0)
    e_fold, s_fold = _err(fold, ref, 1.0)
    print(f"  gated: error vs FP64 unfolded {e_plain:.3g}, folded "
          f"{e_fold:.3g}", flush=True)
    assert s_fold >= s_plain - 0.005
    assert e_fold <= 2.0 * e_plain + 1e-6


def test_planar_folded_flight(engine):
    """The planar quadrupole guide (RF pair in opposite phase): 2 stored
    sets, the folded flight within the unfolded FP32 error against FP64."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.planar_quad(n_ions=1000, coarse=2)
    fl = Flight(c["dt"], c["t_max"], max_steps=5000)
    up = engine.upload_planar
    ref, _ = _fly(engine, up, c["pack"], c["ions"], fl, "f64", None)
    plain, _ = _fly(engine, up, c["pack"], c["ions"], fl, "f32", False)
    fold, lf = _fly(engine, up, c["pack"], c["ions"], fl, "f32", True)
    assert lf == (2, 1, 2)
    e_plain, s_plain = _err(plain, ref, 1.0)
    e_fold, s_fold = _err(fold, ref, 1.0)
    print(f"  planar: error vs FP64 unfolded {e_plain:.3g}, folded "
          f"{e_fold:.3g}", flush=True)
    assert s_fold >= s_plain - 0.002
    assert e_fold <= 2.0 * e_plain + 1e-6


def test_idle_zero_drive_tune_flies_a_folded_pack(engine):
    """A tune row whose drive is identically zero (sin at zero frequency
    and phase, as a server's idle rows are) breaks no linked group."""
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    from ion_gym_amd.serve import make_tune
    c = configs.funnel_rz(n_ions=64, seed=5, coarse=4)
    pack = c["model"].pack()
    idle = make_tune(pack, dt=c["dt"], t_max=c["t_max"],
                     ch_om=np.zeros(2), ch_ph=np.zeros(2))
    fl = Flight(c["dt"], c["t_max"], max_steps=50)
    out, lay = _fly(engine, engine.upload_rz, pack, c["ions"], fl, "f32",
                    True, [idle], np.zeros(64, int))
    assert lay == (2, 1, 2)
    assert np.all(np.isfinite(out["x"]))
