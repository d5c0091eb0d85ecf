"""The engine's 3-D Laplace solver (ig_solve3d) against the oracle, bit for
bit: SOR with both fractional-leg stencils, the V-cycle (thick and 1-node
plates, a mirror-symmetric scene, leg fractions, a coarsest grid too big
for one workgroup), the cascade (plates, leg fractions, odd and even
axes), plus the basis drivers and the refusals."""
import numpy as np
import pytest

from oracle import solve3d as O
from tests.test_solver3d_oracle import _coax, _plate, coax_theta

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from ion_gym_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _same(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a, b), float(np.abs(a - b).max())


def _coax_case():
    inner, ring, exact, _ = _coax(81, 40.0, 6.3, 9)
    val = np.where(inner, 100.0, 0.0) + np.where(ring, exact, 0.0)
    return inner | ring, val, coax_theta(inner, 40.0, 6.3)


@pytest.mark.parametrize("stencil", ["ghost_linear", "sw"])
def test_sor_matches_oracle_bitwise(eng, stencil):
    from ion_gym_amd.physics import solver3d as S
    fixed, val, th = _coax_case()
    print(f"SOR coax 81x81x9 {stencil}", flush=True)
    p, sw, d = S.solve3d(fixed, val, theta=th, tol=1e-8, stencil=stencil,
                         engine=eng)
    q, sw_o, d_o, st = O.solve3d(fixed, val, theta=th, tol=1e-8,
                                 omega=S.optimal_omega(fixed.shape),
                                 stencil=stencil)
    assert st == 0 and sw == sw_o and d == d_o
    _same(p, q)
    # plain node-Dirichlet plate, from a nonzero initial guess
    fixed, val = _plate(n=37, nz=21)
    phi0 = np.linspace(0.0, 50.0, fixed.size).reshape(fixed.shape)
    p, sw, d = S.solve3d(fixed, val, tol=1e-7, phi0=phi0, stencil=stencil,
                         engine=eng)
    q, sw_o, d_o, st = O.solve3d(fixed, val, tol=1e-7, phi0=phi0,
                                 omega=S.optimal_omega(fixed.shape),
                                 stencil=stencil)
    assert st == 0 and sw == sw_o and d == d_o
    _same(p, q)


@pytest.mark.parametrize("thin", [False, True])
def test_vcycle_matches_oracle_bitwise(eng, thin):
    from ion_gym_amd.physics import solver3d as S
    fixed, val = _plate(thin=thin)
    print(f"V-cycle plate 49x49x25 thin={thin}", flush=True)
    p, sw, d = S.solve3d_vmg(fixed, val, tol=1e-6, engine=eng)
    q, sw_o, d_o, st = O.solve3d_vmg(fixed, val, tol=1e-6,
                                     sym_axes=S.symmetric_axes(fixed, val))
    assert st == 0 and sw == sw_o and d == d_o
    _same(p, q)


def test_vcycle_symmetric_and_fractional_match_oracle(eng):
    from ion_gym_amd.physics import solver3d as S
    fixed = np.zeros((41, 33, 21), bool)
    val = np.zeros(fixed.shape)
    fixed[10:31, 15:18, 5:16] = True
    val[fixed] = 1e4
    fixed[:, 0, :] = True
    assert S.symmetric_axes(fixed, val) == [0, 2]
    print("V-cycle symmetric 41x33x21", flush=True)
    p, sw, d = S.solve3d_vmg(fixed, val, tol=1e-7, engine=eng)
    q, sw_o, d_o, st = O.solve3d_vmg(fixed, val, tol=1e-7, sym_axes=(0, 2))
    assert st == 0 and sw == sw_o and d == d_o
    _same(p, q)
    assert np.array_equal(p, np.flip(p, axis=0))
    assert np.array_equal(p, np.flip(p, axis=2))
    fixed, val, th = _coax_case()
    for stencil in ("ghost_linear", "sw"):
        print(f"V-cycle coax {stencil}", flush=True)
        p, sw, d = S.solve3d_vmg(fixed, val, theta=th, tol=1e-8,
                                 stencil=stencil, engine=eng)
        q, sw_o, d_o, st = O.solve3d_vmg(fixed, val, theta=th, tol=1e-8,
                                         stencil=stencil)
        assert st == 0 and sw == sw_o and d == d_o
        _same(p, q)


def test_vcycle_large_coarsest_grid_matches_oracle(eng):
    """371 x 371 x 3: the coarsest error grid (186 x 186 x 2, 69,192
    nodes) is relaxed by chunked device sweeps, not one workgroup."""
    from ion_gym_amd.physics import solver3d as S
    fixed = np.zeros((371, 371, 3), bool)
    val = np.zeros(fixed.shape)
    fixed[100:270, 180:184, :] = True
    val[fixed] = 1e4
    fixed[:, 0, :] = True
    fixed[0, :, :] = True
    print("V-cycle 371x371x3", flush=True)
    p, sw, d = S.solve3d_vmg(fixed, val, tol=1e-6, engine=eng)
    q, sw_o, d_o, st = O.solve3d_vmg(fixed, val, tol=1e-6,
                                     sym_axes=S.symmetric_axes(fixed, val))
    assert st == 0 and sw == sw_o and d == d_o
    _same(p, q)


@pytest.mark.parametrize("thin", [False, True])
def test_cascade_matches_oracle_bitwise(eng, thin):
    from ion_gym_amd.physics import solver3d as S
    fixed, val = _plate(n=48 if thin else 49, thin=thin)
    print(f"cascade plate {fixed.shape} thin={thin}", flush=True)
    p, sw, d = S.solve3d_mg(fixed, val, tol=1e-7, engine=eng)
    q, sw_o, d_o, st, lv = O.solve3d_mg(fixed, val, tol=1e-7,
                                        omega=S.optimal_omega(fixed.shape))
    assert st == 0 and lv == 0 and sw == sw_o and d == d_o
    _same(p, q)


def _slab_129x65x129():
    fixed = np.zeros((129, 65, 129), bool)
    val = np.zeros(fixed.shape)
    fixed[32:97, 31:34, 32:97] = True
    val[fixed] = 1e4
    fixed[:, 0, :] = True
    fixed[:, -1, :] = True
    return fixed, val


@pytest.mark.parametrize("route", ["sor", "cascade"])
def test_two_stage_pair_maximum_matches_oracle_bitwise(eng, route):
    """129x65x129: 17 x 129 = 2193 workgroups per colour sweep, so a pair's
    4386 partial maxima exceed the one-workgroup limit of 4096 and the
    convergence check takes the two-stage reduction."""
    from ion_gym_amd.physics import solver3d as S
    fixed, val = _slab_129x65x129()
    half = -(-65 * 65 // 256)
    assert 2 * half * 129 > 4096
    om = S.optimal_omega(fixed.shape)
    print(f"{route} slab 129x65x129 two-stage maximum", flush=True)
    if route == "sor":
        p, sw, d = S.solve3d(fixed, val, tol=1e-2, engine=eng)
        q, sw_o, d_o, st = O.solve3d(fixed, val, tol=1e-2, omega=om)
    else:
        p, sw, d = S.solve3d_mg(fixed, val, tol=1e-2, engine=eng)
        q, sw_o, d_o, st, _ = O.solve3d_mg(fixed, val, tol=1e-2, omega=om)
    assert st == 0 and sw == sw_o and d == d_o
    _same(p, q)


@pytest.mark.parametrize("stencil", ["ghost_linear", "sw"])
def test_cascade_fractions_match_oracle_bitwise(eng, stencil):
    from ion_gym_amd.physics import solver3d as S
    inner, ring, exact, _ = _coax(81, 40.0, 6.3, 21)
    val = np.where(inner, 100.0, 0.0) + np.where(ring, exact, 0.0)
    fixed, th = inner | ring, coax_theta(inner, 40.0, 6.3)
    print(f"cascade coax 81x81x21 {stencil}", flush=True)
    p, sw, d = S.solve3d_mg(fixed, val, theta=th, tol=1e-8, stencil=stencil,
                            engine=eng)
    q, sw_o, d_o, st, lv = O.solve3d_mg(fixed, val, theta=th, tol=1e-8,
                                        omega=S.optimal_omega(fixed.shape),
                                        stencil=stencil)
    assert st == 0 and sw == sw_o and d == d_o
    _same(p, q)
    with pytest.raises(S.SolveNotConverged,
                       match="cap of 3 sweeps on a 21x21x6 grid"):
        S.solve3d_mg(fixed, val, theta=th, tol=1e-8, max_sweeps=3,
                     engine=eng)


def test_exact_harmonic_and_bases(eng):
    from ion_gym_amd.physics import solver3d as S
    n = 65
    i, j, k = np.meshgrid(*[np.arange(n, dtype=float)] * 3, indexing="ij")
    exact = i * i + j * j - 2 * k * k
    fixed = np.zeros((n, n, n), bool)
    fixed[0], fixed[-1] = True, True
    fixed[:, 0], fixed[:, -1] = True, True
    fixed[:, :, 0], fixed[:, :, -1] = True, True
    print("exact harmonic 65^3 V-cycle", flush=True)
    p, _, _ = S.solve3d_vmg(fixed, np.where(fixed, exact, 0.0), tol=1e-9,
                            engine=eng)
    assert np.abs(p - exact)[~fixed].max() < 1e-6
    # two electrodes: the bases superpose to the both-at-v_basis solve
    a = np.zeros((49, 41, 33), bool)
    b = np.zeros_like(a)
    a[10:20, 15:25, 10:20] = True
    b[30:40, 15:25, 12:22] = True
    wall = np.zeros_like(a)
    wall[:, 0, :] = True
    masks = {1: a, 2: b, 3: wall}
    print("bases 49x41x33", flush=True)
    bm = S.solve_bases_mg(masks, tol=1e-7, only=(1, 2), engine=eng)
    bs = S.solve_bases(masks, tol=1e-7, only=(1, 2), engine=eng)
    assert set(bm) == {1, 2}
    for e in (1, 2):
        assert np.abs(bm[e] - bs[e]).max() < 5e-3
    both = np.where(a | b, 1e4, 0.0)
    pb, _, _ = S.solve3d_vmg(a | b | wall, both, tol=1e-7, engine=eng)
    assert np.abs(bm[1] + bm[2] - pb).max() < 1e-3


def test_refusals(eng):
    from ion_gym_amd.physics import solver3d as S
    fixed, val = _plate(n=21, nz=9)
    with pytest.raises(ValueError, match="unknown stencil"):
        S.solve3d(fixed, val, stencil="star", engine=eng)
    with pytest.raises(ValueError, match="theta must be"):
        S.solve3d(fixed, val, theta=np.ones((6, 3, 3, 3), np.float32),
                  engine=eng)
    with pytest.raises(S.SolveNotConverged, match="SOR hit its cap of 3"):
        S.solve3d(fixed, val, tol=1e-9, max_sweeps=3, engine=eng)
    with pytest.raises(S.SolveNotConverged, match="exhausted max_cycles=1"):
        S.solve3d_vmg(fixed, val, tol=1e-12, max_cycles=1, engine=eng)
    from ion_gym_amd._lib import EngineError
    with pytest.raises(EngineError, match="at least 2 nodes"):
        S.solve3d_vmg(np.zeros((9, 9, 1), bool), np.zeros((9, 9, 1)),
                      engine=eng)
