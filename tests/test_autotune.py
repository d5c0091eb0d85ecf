"""The launch autotuner's search and cache, against a stand-in engine whose
launch times are a known function of the settings (no GPU)."""
import json

import numpy as np

from ion_gym_amd import _lib as L
from ion_gym_amd import autotune as AT
from ion_gym_amd.engine import Flight


class _Ens:
    def __init__(self, n):
        self.n = n
        self.loads = 0

    def load(self, ions):
        assert len(ions["x"]) == self.n
        self.loads += 1

    def release(self):
        pass


class _Engine:
    device = 0

    def __init__(self, cost):
        self.cost = cost
        self.flights = []
        self.last = None

    def ensemble(self, n):
        return _Ens(n)

    def fly(self, fields, flight, ens):
        self.flights.append((flight.block, flight.occupancy,
                             flight.coll_batch, flight.max_steps))
        self.last = self.cost(flight)

    def last_fly_ms(self):
        return self.last


class _Fields:
    route = L.ROUTE_RZ
    n_chan = 2

    def __init__(self, precision):
        self.precision = precision


def _ions(n):
    return {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz")}


def test_candidates_follow_the_kernel_variants():
    hs64 = AT.candidates("f64", L.GAS_HS)
    assert {c["occupancy"] for c in hs64} == {1, 2, 3}
    assert {c["coll_batch"] for c in hs64} == set(AT.COLL_BATCHES)
    hs32 = AT.candidates("f32", L.GAS_HS)
    assert {c["coll_batch"] for c in hs32} == set(AT.COLL_BATCHES)
    vac = AT.candidates("f32", L.GAS_VACUUM)
    assert {c["occupancy"] for c in vac} == {1, 2, 3}
    assert {c["coll_batch"] for c in vac} == {0}
    assert {c["block"] for c in vac} == set(AT.BLOCKS)
    # the planar FP32-basis vacuum step has a fourth, 128-VGPR variant; the
    # 3-D SDS kernel has one
    pl = AT.candidates("f32", L.GAS_VACUUM, L.ROUTE_PLANAR)
    assert {c["occupancy"] for c in pl} == {1, 2, 3, 4}
    assert {c["occupancy"] for c in AT.candidates("f64", L.GAS_VACUUM,
                                                  L.ROUTE_PLANAR)} == {1, 2, 3}
    assert {c["occupancy"] for c in AT.candidates("f32", L.GAS_SDS,
                                                  L.ROUTE_XYZ)} == {1}


def test_autotune_picks_the_fastest_and_caches_it(tmp_path):
    def cost(fl):
        return (abs(fl.block - 128) / 64.0 + abs(fl.occupancy - 2)
                + abs(fl.coll_batch - 16) / 16.0 + 1.0)
    eng = _Engine(cost)
    fl = Flight(1e-3, 10.0, gas=dict(model=L.GAS_HS, T_k=300.0, P_pa=100.0,
                                      c_star=1.0, c_bar=1.0, sig1d=1.0,
                                      m_gas=4.0))
    path = str(tmp_path / "launch.json")
    r = AT.autotune(eng, _Fields("f64"), fl, _ions(5000), sample=1000,
                    window=50, cache_path=path, device=("stub", 7))
    assert (r["block"], r["occupancy"], r["coll_batch"]) == (128, 2, 16)
    assert r["ms"] == 1.0 and not r["cached"]
    assert len(r["table"]) == len(AT.candidates("f64", L.GAS_HS, 0))
    # every measured window is the requested one, over the sample
    assert {f[3] for f in eng.flights} == {50}
    # the flight passed in is not modified
    assert fl.block == 0 and fl.max_steps == 0
    data = json.load(open(path))
    assert len(data) == 1 and "stub|7|" in next(iter(data))
    n_fly = len(eng.flights)
    again = AT.autotune(eng, _Fields("f64"), fl, _ions(5000), sample=1000,
                        window=50, cache_path=path, device=("stub", 7))
    assert again["cached"] and len(eng.flights) == n_fly
    t = AT.tuned(fl, again)
    assert (t.block, t.occupancy, t.coll_batch) == (128, 2, 16)
    # another precision is another key
    AT.autotune(eng, _Fields("f32"), fl, _ions(100), cache_path=path,
                device=("stub", 7))
    assert len(json.load(open(path))) == 2


def test_launch_copy_rejects_other_fields():
    fl = Flight(1.0, 2.0)
    try:
        fl.launch(dt=3.0)
    except ValueError:
        pass
    else:
        raise AssertionError("dt is not a launch setting")
