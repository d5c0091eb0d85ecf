"""The engine's 2-D Laplace solve (ig_solve2d) against the oracle's LU of
the same system: exact quadratic fields, electrode geometries in both
symmetries with both fold kinds, the funnel's r-z grid, and the device
side's own refusal of a free node on a closed edge."""
import ctypes as C

import numpy as np
import pytest

from oracle import solve2d as O

pytestmark = pytest.mark.gpu

REL = 1e-9  # max |device - LU| over max |val|


@pytest.fixture(scope="module")
def eng():
    from ion_gym_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _close(p, q, scale):
    err = float(np.abs(p - q).max())
    assert err <= REL * scale, err


def test_cylindrical_quadratic(eng):
    from ion_gym_amd.physics import solver2d as S
    z, r = np.meshgrid(np.arange(81) * 0.25, np.arange(41) * 0.25,
                       indexing="ij")
    ex = r * r - 2.0 * z * z
    f = np.zeros(ex.shape, bool)
    f[0], f[-1], f[:, -1] = True, True, True
    val = np.where(f, ex, 0.0)
    print("2-D cylindrical quadratic 81x41", flush=True)
    p = S.solve_laplace(f, val, 0.25, "cylindrical", neumann_edges=(),
                        engine=eng)
    _close(p, ex, np.abs(val).max())


@pytest.mark.parametrize("sym", ["planar", "cylindrical"])
@pytest.mark.parametrize("ghost", ["mirror", {"z1": "ghost_linear",
                                              "u1": "ghost_linear"}])
def test_electrodes_match_lu(eng, sym, ghost):
    from ion_gym_amd.configs import _grid_labels
    from ion_gym_amd.physics import solver2d as S
    h = 0.1
    lab = _grid_labels(301, 81, h, [(1, 3.0, 6.0, 2.0, 3.0),
                                    (2, 10.0, 14.0, 2.5, 4.0),
                                    (3, 20.0, 20.5, 0.0, 6.0)])
    edges = ("z0", "z1", "u1") if sym == "cylindrical" else \
        ("z0", "z1", "u0", "u1")
    solve = S.laplace_factor(lab > 0, h, sym, edges, ghost, engine=eng)
    for volts in ({1: 100.0, 2: -40.0, 3: 0.0}, {1: 0.0, 2: 0.0, 3: 1e4}):
        val = sum(np.where(lab == k, v, 0.0) for k, v in volts.items())
        print(f"2-D {sym} {ghost} {volts}", flush=True)
        p = solve(val)
        q, _ = O.solve(lab > 0, val, h, sym, edges, ghost)
        _close(p, q, np.abs(val).max())
        assert solve.sweeps > 0


def test_funnel_grid_matches_lu(eng):
    from ion_gym_amd.configs import _grid_labels
    from ion_gym_amd.physics import solver2d as S
    h = 0.1
    elec = [(i + 1, 10.0 + i, 10.5 + i, 12.5 - 11.5 * i / 99.0,
             15.5 - 11.5 * i / 99.0) for i in range(100)]
    lab = _grid_labels(1201, 201, h, elec)
    val = np.where(lab > 0, 100.0 - (lab - 1) * 1.0, 0.0)
    print("2-D funnel 1201x201 cylindrical", flush=True)
    import time
    solve = S.laplace_factor(lab > 0, h, "cylindrical", ("z0", "z1", "u1"),
                             engine=eng)
    solve(val)
    t0 = time.perf_counter()
    p = solve(val)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    q, _ = O.solve(lab > 0, val, h, "cylindrical", ("z0", "z1", "u1"))
    t_lu = time.perf_counter() - t0
    print(f"device {t_dev:.4f} s ({solve.sweeps} pairs), LU {t_lu:.3f} s, "
          f"max err {np.abs(p - q).max():.3e}", flush=True)
    _close(p, q, 100.0)


def test_device_refuses_a_free_node_on_a_closed_edge(eng):
    from ion_gym_amd import _lib as L
    f = np.zeros((6, 4), np.uint8)
    f[:, -1] = 1
    val = np.zeros(f.shape)
    d = L.Solve2dDesc()
    d.nz, d.nu = f.shape
    d.fixed = L.ptr(f, L._u8p)
    d.val = L.ptr(val)
    d.h, d.symmetry, d.neumann, d.ghost_linear = 1.0, 0, 1 | 4, 0
    d.tol, d.omega, d.max_sweeps = 1e-9, 1.5, 10
    phi = np.empty(f.shape)
    res = L.Solve3dResult()
    with pytest.raises(L.EngineError, match="closed edge z1"):
        L.check(L.lib().ig_solve2d(eng.h, C.byref(d), L.ptr(phi),
                                   C.byref(res)), "ig_solve2d")
