"""SDS on the r-z route: ion-by-ion parity with the oracle's Philox flight
(FP64), transport against the model's own closed forms (drift K E,
diffusion 2 D t along the field and 4 D t across it, each within 1 %), and
several ensembles of different field strength in one launch via the tune
table.  The flights read ion_gym's own jump table (tests/golden)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_rz import assert_parity

pytestmark = pytest.mark.gpu


def _scaled_pack(c):
    pk = dict(c["model"].pack())
    pk["EzA"] = pk["EzA"] * c["scale"]
    pk["EuA"] = pk["EuA"] * c["scale"]
    return pk


def _flight(c, **kw):
    from ion_gym_amd.engine import Flight
    return Flight(c["dt"], c["t_max"], gas=c["gas"], **kw)


def test_sds_fp64_matches_oracle_ion_by_ion(engine):
    from ion_gym_amd import configs
    c = configs.drift_rz(n_ions=256, coarse=2)
    pk = _scaled_pack(c)
    fp = engine.upload_rz(pk, "f64")
    ens = engine.ensemble(256)
    ens.load(c["ions"])
    engine.fly(fp, _flight(c, max_steps=3000), ens)
    g = ens.read()
    o = O.fly_rz(pk, dict(dt=c["dt"], t_max=c["t_max"], max_steps=3000,
                          gas=c["gas"], rng=O.RNG_PHILOX), c["ions"])
    assert_parity(g, o, frac_ok=0.99)
    ens.release()
    fp.release()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_sds_drift_and_diffusion(engine, prec):
    """1e6 ions over 20,000 steps (~324 us): mean drift = K E,
    longitudinal variance = 2 D t and transverse = 4 D t (after the birth
    spread), each within 1 % (sampling errors 0.01 %, 0.14 % and 0.1 %)."""
    from ion_gym_amd import configs
    n = 1_000_000
    c = configs.drift_rz(n_ions=n, coarse=2)
    pk = _scaled_pack(c)
    fp = engine.upload_rz(pk, prec)
    ens = engine.ensemble(n)
    ens.load(c["ions"])
    engine.fly(fp, _flight(c, max_steps=20000), ens)
    g = ens.read()
    assert np.all(g["fate"] == -2)
    T = float(np.mean(g["t_us"]))
    vd = c["K_mm2_per_V_us"] * c["E_vpmm"]
    assert np.mean(g["x"] - 5.0) / T / vd == pytest.approx(1.0, abs=0.01)
    D = c["D"]
    assert np.var(g["x"]) / (2 * D * T) == pytest.approx(1.0, abs=0.01)
    v0 = np.var(c["ions"]["y"]) + np.var(c["ions"]["z"])
    vt = np.var(g["y"]) + np.var(g["z"]) - v0
    assert vt / (4 * D * T) == pytest.approx(1.0, abs=0.01)
    ens.release()
    fp.release()


def test_many_field_strengths_in_one_launch(engine):
    """Four ensembles at 5, 10, 15 and 20 Td share one launch through the
    tune table (set_scale on the static field): each drifts at its own
    K E within 1 % (sampling error 0.05 % at 5 Td)."""
    import ctypes as C
    from ion_gym_amd import _lib as L
    from ion_gym_amd import configs
    from ion_gym_amd.engine import _Keep, make_ions
    from ion_gym_amd.serve import _fill_tune, make_tune
    tds = [5.0, 10.0, 15.0, 20.0]
    per = 100_000
    base = configs.drift_rz(n_ions=per, coarse=2, e_td=10.0)
    fp = engine.upload_rz(base["model"].pack(), "f64")
    ens = engine.ensemble(per * len(tds))
    tunes = (L.Tune * len(tds))()
    cs = []
    for k, td in enumerate(tds):
        c = configs.drift_rz(n_ions=per, coarse=2, e_td=td, seed=4000 + k)
        cs.append(c)
        _fill_tune(tunes[k], make_tune({}, dt=c["dt"], t_max=c["t_max"],
                                       set_scale=[c["scale"]]))
    L.check(L.lib().ig_ensemble_set_tunes(engine.h, ens.id, len(tds), tunes))
    for k, c in enumerate(cs):
        keep = _Keep()
        io = make_ions(keep, c["ions"])
        L.check(L.lib().ig_ensemble_load_at(engine.h, ens.id, k * per, io, k))
    ens.n = per * len(tds)
    engine.fly(fp, _flight(base, max_steps=8000), ens)
    g = ens.read()
    for k, c in enumerate(cs):
        sl = slice(k * per, (k + 1) * per)
        T = float(np.mean(g["t_us"][sl]))
        vd = c["K_mm2_per_V_us"] * c["E_vpmm"]
        assert np.mean(g["x"][sl] - 5.0) / T / vd == pytest.approx(
            1.0, abs=0.01), tds[k]
    live = (C.c_int64 * len(tds))()
    L.check(L.lib().ig_ensemble_live_by_tune(engine.h, ens.id, len(tds),
                                             live))
    assert list(live) == [per] * len(tds)
    ens.release()
    fp.release()
