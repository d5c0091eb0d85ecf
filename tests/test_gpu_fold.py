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
