"""The oracle's 2-D Laplace solve (restating solver2d.assemble_laplace and
its LU) against fields the discrete system reproduces exactly, against the
independent r-z assembly of the benchmark bases, and the refusals of the
reference and of the engine's host side."""
import numpy as np
import pytest

from oracle import solve2d as O


def _grid(nz, nu, h):
    return np.meshgrid(np.arange(nz) * h, np.arange(nu) * h, indexing="ij")


def test_cylindrical_quadratic_is_exact_with_the_axis_row():
    """r^2 - 2 z^2 is harmonic in r-z, and the discrete cylindrical stencil
    (1 +- h/2r off the axis, -6/4 on it) reproduces it exactly."""
    z, r = _grid(41, 21, 0.5)
    ex = r * r - 2.0 * z * z
    f = np.zeros(ex.shape, bool)
    f[0], f[-1], f[:, -1] = True, True, True
    p, _ = O.solve(f, np.where(f, ex, 0.0), 0.5, "cylindrical",
                   neumann_edges=())
    assert np.abs(p - ex).max() < 1e-9


def test_planar_quadratic_is_exact():
    z, u = _grid(33, 25, 0.25)
    ex = u * u - z * z + 3.0 * u
    f = np.zeros(ex.shape, bool)
    f[0], f[-1], f[:, 0], f[:, -1] = True, True, True, True
    p, _ = O.solve(f, np.where(f, ex, 0.0), 0.25, neumann_edges=())
    assert np.abs(p - ex).max() < 1e-10


@pytest.mark.parametrize("sym", ["planar", "cylindrical"])
@pytest.mark.parametrize("ghost", ["mirror", "ghost_linear",
                                   {"u1": "ghost_linear"}])
def test_field_linear_in_z_across_open_edges(sym, ghost):
    """Plates at both z ends, both u edges open: the uniform field is
    the solution whichever way the missing neighbours fold."""
    z, _ = _grid(30, 12, 0.2)
    ex = 5.0 - 40.0 * z
    f = np.zeros(ex.shape, bool)
    f[0], f[-1] = True, True
    p, _ = O.solve(f, np.where(f, ex, 0.0), 0.2, sym,
                   neumann_edges=("u0", "u1"), edge_ghost=ghost)
    assert np.abs(p - ex).max() < 1e-10


def test_matches_the_benchmark_rz_assembly():
    from ion_gym_amd.configs import _grid_labels
    from ion_gym_amd.physics.solver2d import solve_rz
    h = 0.2
    lab = _grid_labels(121, 41, h, [(1, 4.0, 8.0, 3.0, 4.0),
                                    (2, 12.0, 16.0, 3.0, 4.0)])
    (a,) = solve_rz(lab, [{1: 10.0, 2: -250.0}], h)
    val = np.where(lab == 1, 10.0, 0.0) + np.where(lab == 2, -250.0, 0.0)
    p, _ = O.solve(lab > 0, val, h, "cylindrical",
                   neumann_edges=("z0", "z1", "u1"))
    assert np.abs(p - a).max() < 1e-10


def test_refusals():
    f = np.zeros((9, 5), bool)
    f[:, -1] = True
    with pytest.raises(ValueError, match="closed edge z1"):
        O.assemble(f, 1.0, neumann_edges=("z0", "u0"))
    from ion_gym_amd.physics import solver2d as S
    with pytest.raises(ValueError, match="closed edge z1"):
        S.laplace_factor(f, 1.0, neumann_edges=("z0", "u0"))
    # the axis needs no ghost: u0 closed is fine when cylindrical
    f2 = f.copy()
    f2[0], f2[-1] = True, True
    O.assemble(f2, 1.0, "cylindrical", neumann_edges=())
    S.laplace_factor(f2, 1.0, "cylindrical", neumann_edges=())
    with pytest.raises(ValueError, match="closed edge u0"):
        S.laplace_factor(f2, 1.0, "planar", neumann_edges=())
    with pytest.raises(ValueError, match="edge_ghost must be one of"):
        S.laplace_factor(f, 1.0, edge_ghost="reflect")
    with pytest.raises(ValueError, match="unknown edge"):
        S.laplace_factor(f, 1.0, edge_ghost={"z2": "mirror"})
    with pytest.raises(ValueError, match="symmetry must be"):
        S.laplace_factor(f, 1.0, symmetry="spherical")


def test_sor_omega_limits():
    from ion_gym_amd.physics.solver2d import sor_omega
    w = sor_omega((101, 101))
    assert 1.9 < w < 2.0
    assert sor_omega((101, 101), ("z0",)) > w
    assert sor_omega((2, 2)) < w
