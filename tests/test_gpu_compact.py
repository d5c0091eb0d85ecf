"""Live-ion compaction (ig_flight_params.compact) against the same windows
without it: a wave-ballot kernel packs the ensemble's flying ions into a
dense index list before each window (mode 1), optionally put in cell
order by a counting sort over grid blocks (mode 2), so lanes map to other
ions than in the plain launch.  An ion's flight does not depend on its lane, so every
field of every ion stays bit-identical after every window -- on r-z with
hard-sphere gas (FP64 and the FP32 basis), on the 3-D quadrupole with
gas, on the planar guide, in a partly filled ensemble, and inside a
captured graph with eager work between replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz", "t_us", "fate", "ncol", "steps")
FLYING = -2


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


def _windows(engine, fp, fl, ions, cap, n_win, graph_every=0, mode=1):
    plain = engine.ensemble(cap)
    comp = engine.ensemble(cap)
    plain.load(ions)
    comp.load(ions)
    flc = fl.launch(compact=mode)
    fl = fl.launch(compact=False)
    lives = []
    for w in range(n_win):
        engine.fly(fp, fl, plain)
        if graph_every and w % graph_every == 0:
            engine.fly_graph(fp, flc, comp, 1)
        else:
            engine.fly(fp, flc, comp)
        engine.sync()
        a, b = plain.read(), comp.read()
        _same(a, b)
        lives.append(int((a["fate"] == FLYING).sum()))
    plain.release()
    comp.release()
    return lives


def _einzel_gas(n):
    """The einzel lens with helium at 0.01 Torr: ~25 collisions across the
    lens, so some ions reach the tubes and most leave the grid within
    ~6000 steps."""
    from ion_gym_amd import configs
    from ion_gym_amd.physics.collision3d import gas_kernel_scalars
    c = configs.einzel_rz(n_ions=n, seed=5, coarse=2)
    g0 = gas_kernel_scalars("He", T_k=300.0, P_torr=0.01)
    c["gas"] = dict(model=1, **{k: g0[k] for k in ("T_k", "P_pa", "c_star",
                                                   "c_bar", "sig1d",
                                                   "m_gas")})
    c["ions"] = dict(c["ions"], sigma_m2=np.full(n, 2e-18))
    return c


@pytest.mark.parametrize("prec,mode", [("f64", 1), ("f32", 1), ("f64", 2),
                                       ("f32", 2)])
def test_compact_rz_hard_sphere_bitwise(engine, prec, mode):
    from ion_gym_amd.engine import Flight
    c = _einzel_gas(6000)
    fp = engine.upload_rz(c["model"].pack(), prec)
    fl = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=500)
    lives = _windows(engine, fp, fl, c["ions"], 6000, 16, mode=mode)
    # ions are lost during the windows, so the lists really were compacted
    assert lives[-1] < lives[0] / 2
    fp.release()


def test_compact_partly_filled_ensemble_and_graph(engine):
    from ion_gym_amd.engine import Flight
    c = _einzel_gas(3000)
    fp = engine.upload_rz(c["model"].pack(), "f32")
    fl = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=500)
    # capacity above the load: the empty slots are never listed
    lives = _windows(engine, fp, fl, c["ions"], 5000, 16, graph_every=2,
                     mode=2)
    assert lives[-1] < lives[0] / 2
    fp.release()


def test_compact_3d_hard_sphere_bitwise(engine):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    from ion_gym_amd.physics.collision3d import gas_kernel_scalars
    c = configs.quad_3d(n_ions=2000, coarse=4, n_cycles=20)
    g0 = gas_kernel_scalars("N2", T_k=300.0, P_torr=0.05)
    gas = dict(model=1, **{k: g0[k] for k in ("T_k", "P_pa", "c_star",
                                              "c_bar", "sig1d", "m_gas")})
    ions = dict(c["ions"])
    ions["sigma_m2"] = np.full(len(ions["x"]), 2e-18)
    for prec in ("f64", "f32"):
        fp = engine.upload_xyz(c["pack"], prec)
        fl = Flight(c["dt"], c["t_max"], gas=gas, max_steps=150)
        _windows(engine, fp, fl, ions, 2000, 6)
        _windows(engine, fp, fl, ions, 2000, 6, mode=2)
        fp.release()


def test_compact_planar_bitwise(engine):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.planar_quad(n_ions=3000, coarse=2, n_cycles=30, rf_v=700.0)
    fp = engine.upload_planar(c["pack"], "f64")
    fl = Flight(c["dt"], c["t_max"], max_steps=200)
    lives = _windows(engine, fp, fl, c["ions"], 3000, 8, mode=2)
    assert lives[-1] < lives[0]
    fp.release()
