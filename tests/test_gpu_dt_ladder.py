"""The reference's step-size ladder of its hard-sphere drift tube
(tests/golden/dt_ladder/drift_ckpt_dt{0.25,0.5,1,2}.csv: the reference's
notebooks/data/mason_schamp_dt_ladder/, 500 ions of C60, m/z 720, in He at
1 Torr flown by ion_gym at dt 0.25, 0.5, 1 and 2 ns, with the times each
ion crossed x = 26 and 59 mm).

The ladder's deck is not in the reference, only its flights.  Its drift
speed at dt 0.25 ns is the certified winner deck's (72.2 against 72.0
m/s: the same ion, gas and 0.2 V/mm field; tests/test_gpu_winner_drift.py
holds that deck against its own flights), and the ladder shows the
reference integrator's step-size bias: the drift speed rises by 0.3, 1.5
and 3.9 % at 0.5, 1 and 2 ns (at most one collision per step, with
probability 1 - exp(-h) of the step's hazard h).  The engine flies the
winner deck (configs.drift_winner_rz, births from the deck's source) at
the same four steps in both precisions and measures the drift speed
through the uniform field between x = 32 and 53 mm (linear crossing times
between 1 us records): at every step it is within four standard errors
of the difference of the ladder's (26 -> 59 mm), and the 2 ns speed is
above the 0.25 ns one by the ladder's margin within four standard
errors.  The comparison is distributional: the reference drew its
collisions from numba's MT19937 stream."""
import csv
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LADDER = os.path.join(os.path.dirname(__file__), "golden", "dt_ladder")
STEPS = (0.25, 0.5, 1.0, 2.0)
N_GPU = 3000


def _ladder(dt):
    name = f"drift_ckpt_dt{dt:g}.csv"
    rows = list(csv.DictReader(open(os.path.join(LADDER, name),
                                    encoding="utf-8")))
    f = lambda k: np.array([float(r[k]) if r[k] != "" else np.nan  # noqa
                            for r in rows])
    return _speed(f("t26"), f("t59"), 33.0)


def _speed(ta, tb, dx):
    ok = np.isfinite(ta) & np.isfinite(tb)
    d = tb[ok] - ta[ok]
    return dx / d.mean(), dx * d.std() / d.mean() ** 2 / math.sqrt(ok.sum())


def _crossing(t, x, X):
    j = np.nonzero((x[:-1] < X) & (x[1:] >= X))[0]
    if not len(j):
        return np.nan
    a = j[0]
    return t[a] + (X - x[a]) / (x[a + 1] - x[a]) * (t[a + 1] - t[a])


def _engine_speed(engine, prec, dt):
    from ion_gym_amd import configs
    from ion_gym_amd.births import generate_births
    from ion_gym_amd.physics.build_rz import make_rz_fly_fn
    c = configs.drift_winner_rz(n_ions=N_GPU, dt_ns=dt)
    spec = c["spec"]
    spec.precision = prec
    # every ion that reaches x = 53 mm does so well before 1100 us
    spec.integration.t_max_us = 1100.0
    births = generate_births(c["source"], engine)
    fly_fn, cols = make_rz_fly_fn(c["model"], births, spec)
    got = fly_fn.batch(range(N_GPU))
    ix = cols.index("x")
    t32 = np.array([_crossing(tr[:, 0], tr[:, ix], 32.0) for tr, _ in got])
    t53 = np.array([_crossing(tr[:, 0], tr[:, ix], 53.0) for tr, _ in got])
    return _speed(t32, t53, 21.0)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_step_size_ladder_matches_the_references(engine, prec):
    got = {}
    for dt in STEPS:
        v, s = _engine_speed(engine, prec, dt)
        vl, sl = _ladder(dt)
        got[dt] = (v, s, vl, sl)
        print(f"[ladder] {prec} dt {dt} ns: engine {v * 1e3:.2f} +- "
              f"{s * 1e3:.2f} m/s, reference {vl * 1e3:.2f} +- "
              f"{sl * 1e3:.2f}", flush=True)
        assert abs(v - vl) <= 4 * math.hypot(s, sl), (dt, v, vl, s, sl)
    v0, s0, l0, e0 = got[0.25]
    v2, s2, l2, e2 = got[2.0]
    assert abs((v2 - v0) - (l2 - l0)) <= 4 * math.sqrt(
        s0 ** 2 + s2 ** 2 + e0 ** 2 + e2 ** 2)
    assert v2 > v0 + 4 * math.hypot(s0, s2)   # the bias is resolved
