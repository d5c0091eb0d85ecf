"""Graph-captured window loops (ig_fly_graph) against the same windows
launched one by one: bit-identical after every replay of the cached
graph, with eager launches, reads and synchronisations between replays,
and re-captured when an argument changes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz", "t_us", "fate", "ncol", "steps")


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_graph_replays_equal_eager_windows(engine, prec):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.funnel_rz(n_ions=4000, seed=3, coarse=2)
    fp = engine.upload_rz(c["model"].pack(), prec)
    fl = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=50)
    eager = engine.ensemble(4000)
    graph = engine.ensemble(4000)
    eager.load(c["ions"])
    graph.load(c["ions"])
    for rep in range(3):
        for _ in range(8):
            engine.fly(fp, fl, eager)
        engine.fly_graph(fp, fl, graph, 8)
        engine.sync()
        a, b = eager.read(), graph.read()
        _same(a, b)
        assert np.all(b["steps"] == 400 * (rep + 1))
    assert b["ncol"].sum() > 0
    # another window length is another graph
    fl2 = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=30)
    for _ in range(5):
        engine.fly(fp, fl2, eager)
    engine.fly_graph(fp, fl2, graph, 5)
    _same(eager.read(), graph.read())
    eager.release()
    graph.release()
    fp.release()


def test_graph_sds_drift_equals_eager(engine):
    from ion_gym_amd import configs
    from ion_gym_amd.engine import Flight
    c = configs.drift_rz(n_ions=3000, coarse=2)
    pk = dict(c["model"].pack())
    pk["EzA"] = pk["EzA"] * c["scale"]
    pk["EuA"] = pk["EuA"] * c["scale"]
    fp = engine.upload_rz(pk, "f64")
    fl = Flight(c["dt"], c["t_max"], gas=c["gas"], max_steps=100)
    eager = engine.ensemble(3000)
    graph = engine.ensemble(3000)
    eager.load(c["ions"])
    graph.load(c["ions"])
    for _ in range(2):
        for _ in range(6):
            engine.fly(fp, fl, eager)
        engine.fly_graph(fp, fl, graph, 6)
        _same(eager.read(), graph.read())
    eager.release()
    graph.release()
    fp.release()
