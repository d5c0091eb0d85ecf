"""ig_fly_noalloc, the flight window without device allocation, against
ig_fly through the C ABI: the same ensembles flown by both must agree bit
for bit (the live list, its cell order and the SDS table are only moved
into the device's fixed scratch region).  Two calls in a row, the second
with a smaller ensemble, so that what the first call left in the shared
region cannot leak into the second; hard sphere in FP64 and in the FP32
basis with the cell-ordered live list, SDS (its jump table copied into the
region) with the slot-ordered one; and the refusal beyond the limit."""
import ctypes as C

import numpy as np
import pytest

from ion_gym_amd import _lib as L

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz", "t_us", "fate", "face", "label",
        "ncol", "steps")


def _noalloc(engine, fp, flight, ens):
    from ion_gym_amd.engine import _Keep
    keep = _Keep()
    p = flight.struct(keep)
    return L.lib().ig_fly_noalloc(engine.h, fp.id, C.byref(p), ens.id)


def _both(engine, fp, flight, ions, windows):
    n = len(ions["x"])
    a, b = engine.ensemble(n), engine.ensemble(n)
    a.load(ions)
    b.load(ions)
    for _ in range(windows):
        engine.fly(fp, flight, a)
        assert _noalloc(engine, fp, flight, b) == 0, L.lib().ig_last_error()
    ga, gb = a.read(records=False), b.read(records=False)
    a.release()
    b.release()
    return ga, gb


def _same(ga, gb):
    for k in KEYS:
        assert np.array_equal(ga[k], gb[k]), k


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_flight_equals_ig_fly_on_two_calls(engine, prec):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.funnel_rz(n_ions=3000, seed=3, coarse=2)
    fp = engine.upload_rz(c["model"].pack(), prec)
    for n, compact in ((3000, 2), (700, 2), (500, 1)):
        ions = {k: v[:n] for k, v in c["ions"].items()}
        fl = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=400,
                    compact=compact)
        ga, gb = _both(engine, fp, fl, ions, 3)
        assert np.all(ga["steps"] == 1200)
        _same(ga, gb)
    fp.release()


def test_noalloc_sds_flight_equals_ig_fly(engine):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.drift_rz(n_ions=2000, coarse=2)
    pk = dict(c["model"].pack())
    pk["EzA"] = pk["EzA"] * c["scale"]
    pk["EuA"] = pk["EuA"] * c["scale"]
    fp = engine.upload_rz(pk, "f64")
    for n in (2000, 600):
        ions = {k: v[:n] for k, v in c["ions"].items()}
        fl = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=500,
                    compact=1)
        ga, gb = _both(engine, fp, fl, ions, 2)
        _same(ga, gb)
    fp.release()


def test_noalloc_refuses_more_ions_than_its_limit(engine):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    n = L.NOALLOC_MAX_IONS + 1
    c = configs.einzel_rz(n_ions=n, coarse=4)
    fp = engine.upload_rz(c["model"].pack(), "f64")
    ens = engine.ensemble(n)
    ens.load(c["ions"])
    fl = Flight(c["dt"], c["t_max"], max_steps=10, compact=1)
    rc = _noalloc(engine, fp, fl, ens)
    assert rc == L.E_ARG
    assert b"no-allocation limit" in L.lib().ig_last_error()
    assert ens.steps() == 0                      # nothing was flown
    ens.release()
    fp.release()
