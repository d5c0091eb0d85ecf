"""Device births (ig_births, ig_ensemble_births_at) against the oracle's
restatement of generate_births on the same seed: every distribution, with
directed energies, thermal and per-axis spreads, drifts and mass lists.

Tolerance: the positions of the disc and grid sources pass through cos,
sin (and hypot, atan2 for the grid), which the device and the host C
library may round differently in the last place, so those are compared
to 4e-15 relative; a standard normal from the ziggurat's tail (about one
draw in 4000) goes through log1p, with the same caveat.  Everything else
-- the draw order, the stream position of every ion, uniforms, energies,
ziggurat normals, truncation decisions -- must match bit for bit, so at
least 99.9% of all values are identical and the rest within 4e-15."""
import warnings

import numpy as np
import pytest

from oracle import births as OB
from ion_gym_amd.physics.flight_spec import SourceSpec

pytestmark = pytest.mark.gpu


def _match(g, o, trig_cols=()):
    assert g.shape == o.shape
    assert np.all(np.isfinite(g))
    same = g == o
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1.0)
    assert rel.max() <= 4e-15, (np.unravel_index(rel.argmax(), rel.shape),
                                g[rel.argmax() // 7], o[rel.argmax() // 7])
    exact_cols = [c for c in range(7) if c not in trig_cols]
    assert same[:, exact_cols].mean() >= 0.999
    return same.mean()


SOURCES = {
    "disc_thermal": dict(distribution="disc", n_ions=3000, r_mm=2.5,
                         x0_mm=0.0, ke_lo=0.0, ke_hi=0.0,
                         temperature_k=300.0, mz_list=[200.0, 2000.0],
                         tob_span_us=1.0),
    "disc_beam": dict(distribution="disc", n_ions=2000, r_mm=1.0, axis="z",
                      z0_mm=3.0, ke_lo=0.5, ke_hi=4.5,
                      direction=[0.3, -0.2, 1.0], mz_list=[556.0],
                      tob_span_us=2.0),
    "point_spreads": dict(distribution="point", n_ions=2500, ke_lo=1.0,
                          ke_hi=2.0, direction=[1.0, 1.0, 0.0],
                          temperature_k=50.0, dv_fwhm_ms=[10.0, 0.0, 3.0],
                          v_drift_mm_us=[0.0, 0.0, 2.06],
                          mz_list=[100.0, 300.0, 900.0]),
    "line": dict(distribution="line", n_ions=777, len_mm=12.0, axis="y",
                 ke_lo=0.0, ke_hi=0.0, temperature_k=10.0, mz_list=[50.0]),
    "grid": dict(distribution="grid", n_ions=1000, r_mm=3.0, ke_lo=0.1,
                 ke_hi=1.9, mz_list=[120.0, 322.0]),
    "box": dict(distribution="box", n_ions=4000, box_mm=[1.0, 2.0, 0.5],
                ke_lo=0.0, ke_hi=3.0, direction=[0.0, 0.0, -1.0],
                mz_list=[400.0]),
    "gaussian": dict(distribution="gaussian", n_ions=3000, x0_mm=5.0,
                     y0_mm=5.0, fwhm_mm=[1.0, 0.5, 0.2],
                     trunc_mm=[0.6, 0.3, 0.1], temperature_k=300.0,
                     ke_lo=0.0, ke_hi=0.0, mz_list=[100.0]),
}


@pytest.mark.parametrize("name", sorted(SOURCES))
def test_device_births_match_oracle(engine, name):
    from ion_gym_amd import births as B
    src = SourceSpec(**SOURCES[name], seed=17)
    g = B.generate_births(src, engine)
    o = OB.generate_births(src, 17)
    trig = (0, 1, 2) if src.distribution in ("disc", "grid") else ()
    frac = _match(g, o, trig)
    print(f"{name}: {frac:.5f} of values bit-identical")


def test_device_births_large_source_matches_oracle(engine):
    """200,000 thermal disc births: a stream of ~1.4M words and 18 jump
    levels."""
    from ion_gym_amd import births as B
    src = SourceSpec(distribution="disc", n_ions=100_000, r_mm=5.0,
                     x0_mm=0.0, ke_lo=0.0, ke_hi=0.0, temperature_k=300.0,
                     mz_list=[200.0, 1000.0], tob_span_us=1.0, seed=3)
    g = B.generate_births(src, engine)
    o = OB.generate_births(src, 3)
    _match(g, o, (0, 1, 2))


def test_zero_direction_on_device(engine):
    from ion_gym_amd import births as B
    kw = dict(distribution="disc", n_ions=4, x0_mm=5.0, y0_mm=3.0,
              z0_mm=1.0, r_mm=0.5, axis="x", ke_lo=0.02, ke_hi=0.05,
              mz_list=[120.0, 322.0], tob_span_us=0.0, seed=3)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        rest = B.generate_births(SourceSpec(direction=[0, 0, 0], **kw),
                                 engine)
    assert any("nowhere to point" in str(c.message) for c in caught)
    aimed = B.generate_births(SourceSpec(direction=[1, 0, 0], **kw), engine)
    assert np.array_equal(rest[:, 3:6], np.zeros((8, 3)))
    assert np.array_equal(rest[:, 0:3], aimed[:, 0:3])


def test_truncation_failure_is_raised(engine, monkeypatch):
    from ion_gym_amd import births as B
    monkeypatch.setattr(SourceSpec, "gaussian_layout",
                        lambda self: ([1.0, 0.0, 0.0], [1e-300, 0.0, 0.0]))
    src = SourceSpec(distribution="gaussian", n_ions=50, fwhm_mm=[1, 0, 0],
                     trunc_mm=[1, 0, 0], seed=1)
    with pytest.raises(RuntimeError, match="truncated Gaussian"):
        B.generate_births(src, engine)


def test_births_into_ensemble_slots(engine):
    """Births drawn straight into ensemble slots equal the host rows, and
    the slots fly like ions loaded from those rows."""
    from ion_gym_amd import births as B
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    src = SourceSpec(distribution="disc", n_ions=1500, r_mm=2.0, x0_mm=1.0,
                     ke_lo=0.0, ke_hi=0.0, temperature_k=300.0,
                     mz_list=[400.0], tob_span_us=0.5, seed=8)
    rows = B.generate_births(src, engine)
    c = configs.funnel_rz(n_ions=1500, seed=3, coarse=2)
    ions = {k: v for k, v in c["ions"].items()
            if k not in ("x", "y", "z", "vx", "vy", "vz", "tob")}
    ens = engine.ensemble(4000)
    ens.load(ions)
    B.births_into(engine, ens, 0, src)
    got = ens.read(records=False)
    for j, k in enumerate(("x", "y", "z", "vx", "vy", "vz")):
        assert np.array_equal(got[k], rows[:, j]), k
    assert np.array_equal(got["t_us"], rows[:, 6])
    ref = engine.ensemble(4000)
    full = dict(ions)
    for j, k in enumerate(("x", "y", "z", "vx", "vy", "vz", "tob")):
        full[k] = rows[:, j].copy()
    ref.load(full)
    fp = engine.upload_rz(c["model"].pack(), "f32")
    fl = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=300)
    engine.fly(fp, fl, ens)
    engine.fly(fp, fl, ref)
    a, b = ens.read(records=False), ref.read(records=False)
    for k in ("x", "y", "z", "vx", "vy", "vz", "t_us", "fate", "ncol"):
        assert np.array_equal(a[k], b[k]), k
    ens.release()
    ref.release()
    fp.release()
