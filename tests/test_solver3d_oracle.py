"""The oracle's 3-D Laplace solver (restating solver3d.py and
multigrid3d.py) against the reference's own solver gates at their sizes
where the CPU allows (tests/test_solver3d_analytic.py, test_multigrid.py,
test_solver3d_nz1.py), and its pieces against independent numpy
formulations (the reference's vectorised full weighting and prolongation,
a padded-array residual, the reshape block-any)."""
import numpy as np
import pytest

from oracle import solve3d as S


def shell(nx, ny, nz):
    f = np.zeros((nx, ny, nz), bool)
    f[0], f[-1] = True, True
    f[:, 0], f[:, -1] = True, True
    f[:, :, 0], f[:, :, -1] = True, True
    return f


def test_exact_quadratic_harmonic():
    """S1-a: x^2 + y^2 - 2 z^2 is in the null space of the 7-point
    Laplacian: reproduced to solver tolerance."""
    n = 21
    i, j, k = np.meshgrid(*[np.arange(n, dtype=float)] * 3, indexing="ij")
    exact = i * i + j * j - 2 * k * k
    fixed = shell(n, n, n)
    phi, sw, d, st = S.solve3d(fixed, np.where(fixed, exact, 0.0), tol=1e-9)
    assert st == 0 and d < 1e-9
    assert np.abs(phi - exact)[~fixed].max() < 1e-6


def test_second_order():
    """S1-b: sin(ax) sin(by) sinh(cz), c^2 = a^2 + b^2: error ~ h^2."""
    errs = []
    for n in (9, 17):
        h = 1.0 / (n - 1)
        a = b = np.pi
        c = np.sqrt(a * a + b * b)
        x = np.arange(n) * h
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        exact = np.sin(a * X) * np.sin(b * Y) * np.sinh(c * Z)
        fixed = shell(n, n, n)
        phi, _, _, st = S.solve3d(fixed, np.where(fixed, exact, 0.0),
                                  tol=1e-11)
        assert st == 0
        errs.append(np.abs(phi - exact)[~fixed].max())
    assert np.log2(errs[0] / errs[1]) > 1.8, errs


def _coax(N, c, r, nz):
    ii, jj = np.meshgrid(np.arange(N, dtype=float), np.arange(N, dtype=float),
                         indexing="ij")
    rr = np.hypot(ii - c, jj - c)
    inner = (rr <= r)[:, :, None] & np.ones((1, 1, nz), bool)
    ring = (rr >= 32)[:, :, None] & np.ones((1, 1, nz), bool)
    exact = np.log(np.maximum(rr, r) / 32.0) / np.log(r / 32.0) * 100.0
    exact = exact[:, :, None] * np.ones((1, 1, nz))
    return inner, ring, exact, rr


def test_fold_against_full_domain():
    """S1-c: a problem symmetric about x = 0 and y = 0 solved on the full
    domain and on the quadrant (even reflection at the low faces) agree."""
    n, nz, r = 41, 9, 6.3
    inner, ring, exact, _ = _coax(2 * n - 1, 40.0, r, nz)
    val = np.where(inner, 100.0, 0.0) + np.where(ring, exact, 0.0)
    phi_f, _, _, st = S.solve3d(inner | ring, val, tol=1e-8)
    assert st == 0
    inner_q, ring_q, exact_q, _ = _coax(n, 0.0, r, nz)
    val_q = np.where(inner_q, 100.0, 0.0) + np.where(ring_q, exact_q, 0.0)
    phi_q, _, _, st = S.solve3d(inner_q | ring_q, val_q, tol=1e-8)
    assert st == 0
    assert np.abs(phi_q - phi_f[40:, 40:, :]).max() < 1e-5


def coax_theta(inner, c, r):
    """Leg fractions to the exact circle of radius r about (c, c): for a
    vacuum node whose axis neighbour is inside, the distance along the axis
    to the surface (clamped to [1e-3, 1], float32), as the reference's
    bisection on the exact solid gives it (solver3d.py:75-128)."""
    nx, ny, nz = inner.shape
    th = np.ones((6, nx, ny, nz), np.float32)
    for d, (ax, s) in enumerate(((0, -1), (0, 1), (1, -1), (1, 1))):
        for i in range(nx):
            for j in range(ny):
                if inner[i, j, 0]:
                    continue
                ni, nj = (i + s, j) if ax == 0 else (i, j + s)
                if not (0 <= ni < nx and 0 <= nj < ny) or not inner[ni, nj, 0]:
                    continue
                along = (i - c) if ax == 0 else (j - c)
                across = (j - c) if ax == 0 else (i - c)
                t = abs(along) - np.sqrt(r * r - across * across)
                th[d, i, j, :] = np.float32(np.clip(t, 1e-3, 1.0))
    return th


def test_fractional_legs_beat_node_dirichlet():
    """S1-d: coaxial capacitor with a non-lattice inner radius: the
    Shortley-Weller legs beat the node-Dirichlet solve by >= 10x in max
    interior error, the linear-ghost legs by >= 3x."""
    n, nz, r = 81, 9, 6.3
    inner, ring, exact, rr = _coax(n, 40.0, r, nz)
    fixed = inner | ring
    val = np.where(inner, 100.0, 0.0) + np.where(ring, exact, 0.0)
    vac = ~fixed & (rr[:, :, None] > r) & (rr[:, :, None] < 32)
    phi0, _, _, _ = S.solve3d(fixed, val, tol=1e-8)
    e0 = np.abs(phi0 - exact)[vac].max()
    th = coax_theta(inner, 40.0, r)
    phi1, _, _, st = S.solve3d(fixed, val, theta=th, tol=1e-8, stencil="sw")
    assert st == 0
    e1 = np.abs(phi1 - exact)[vac].max()
    assert e1 * 10 <= e0, (e0, e1)
    phi2, _, _, st = S.solve3d(fixed, val, theta=th, tol=1e-8)
    e2 = np.abs(phi2 - exact)[vac].max()
    assert e2 * 3 <= e0, (e0, e2)


def _plate(n=49, nz=25, thin=False):
    fixed = np.zeros((n, n, nz), bool)
    val = np.zeros(fixed.shape)
    w = 1 if thin else 3
    fixed[n // 4:3 * n // 4, n // 2:n // 2 + w, nz // 4:3 * nz // 4] = True
    val[fixed] = 1e4
    fixed[:, 0, :] = True
    fixed[:, -1, :] = True
    return fixed, val


@pytest.mark.parametrize("thin", [False, True])
def test_vcycle_reaches_the_sor_fixed_point(thin):
    """Multigrid gates 1-2: the V-cycle converges (block-any coarse masks
    keep a 1-node plate) to SOR's field within the convergence floor."""
    fixed, val = _plate(thin=thin)
    p_sor, _, _, st = S.solve3d(fixed, val, tol=1e-6)
    assert st == 0
    p_v, sw, d, st = S.solve3d_vmg(fixed, val, tol=1e-6)
    assert st == 0 and d < 1e-6
    assert np.abs(p_v - p_sor).max() < 5e-4


def test_vcycle_accelerates():
    """Multigrid gate 4 (129 x 65 x 33): far fewer fine sweeps than SOR."""
    n = 129
    fixed = np.zeros((n, 65, 33), bool)
    val = np.zeros(fixed.shape)
    fixed[10:119, 30:33, 8:25] = True
    val[fixed] = 1e4
    fixed[:, 0, :] = True
    p_sor, sw_sor, _, st = S.solve3d(fixed, val, tol=1e-6)
    assert st == 0
    p_v, sw_v, d, st = S.solve3d_vmg(fixed, val, tol=1e-6)
    assert st == 0
    assert np.abs(p_v - p_sor).max() < 5e-4
    assert sw_v < 0.5 * sw_sor, (sw_v, sw_sor)


def test_vcycle_symmetric_projection_keeps_the_fixed_point():
    """A problem symmetric in x: projecting each cycle onto the symmetric
    subspace converges to SOR's field and leaves it exactly symmetric."""
    fixed = np.zeros((41, 33, 21), bool)
    val = np.zeros(fixed.shape)
    fixed[10:31, 15:18, 5:16] = True
    val[fixed] = 1e4
    fixed[:, 0, :] = True
    p_sor, _, _, _ = S.solve3d(fixed, val, tol=1e-7)
    p_v, _, d, st = S.solve3d_vmg(fixed, val, tol=1e-7, sym_axes=(0,))
    assert st == 0
    assert np.array_equal(p_v, np.flip(p_v, axis=0))
    assert np.abs(p_v - p_sor).max() < 5e-5


def _axis_potential(ex, ey, nz, N=51, C=25):
    metal = np.zeros((N, N, nz), bool)
    metal[0, :, :] = metal[-1, :, :] = True
    metal[:, 0, :] = metal[:, -1, :] = True
    metal[ex - 2:ex + 3, ey - 2:ey + 3, :] = True
    val = np.zeros((N, N, nz))
    val[ex - 2:ex + 3, ey - 2:ey + 3, :] = 1e4
    phi, _, _, st = S.solve3d(metal, val, tol=1e-8, max_sweeps=20000,
                              omega=1.0)
    assert st == 0
    return phi[C, C, nz // 2], phi


def test_size_one_axis_is_the_planar_problem():
    """nz = 1 gates: x/y symmetry, equality with nz = 3 and 5 for a
    z-uniform geometry, and the pure 5-point mean at an interior node."""
    vx, phi = _axis_potential(40, 25, 1)
    vy, _ = _axis_potential(25, 40, 1)
    assert abs(vx - vy) < 1e-4
    v3, _ = _axis_potential(40, 25, 3)
    v5, _ = _axis_potential(40, 25, 5)
    assert abs(vx - v3) < 1e-3 and abs(vx - v5) < 1e-3
    p = phi[:, :, 0]
    i, j = 20, 30
    mean = (p[i - 1, j] + p[i + 1, j] + p[i, j - 1] + p[i, j + 1]) / 4.0
    assert abs(p[i, j] - mean) < 1e-6


# ---- the pieces against independent formulations ----------------------

def _restrict_fw_numpy(fine):
    """The reference's vectorised full weighting (multigrid3d.py:529-550):
    edge-replicated pad, 27 strided terms in (di, dj, dk) order, / 8."""
    p = np.pad(fine, 1, mode="edge")
    out = None
    for di, wi in ((-1, .5), (0, 1.), (1, .5)):
        for dj, wj in ((-1, .5), (0, 1.), (1, .5)):
            for dk, wk in ((-1, .5), (0, 1.), (1, .5)):
                s = p[1 + di:p.shape[0] - 1 + di, 1 + dj:p.shape[1] - 1 + dj,
                      1 + dk:p.shape[2] - 1 + dk][::2, ::2, ::2]
                term = (wi * wj * wk) * s
                out = term if out is None else out + term
    return out / 8.0


def _prolongate_numpy(phic, shape_f):
    """The reference's numpy prolongation (multigrid3d.py:138-172)."""
    nx, ny, nz = shape_f
    xf = np.minimum(np.arange(nx) / 2.0, phic.shape[0] - 1)
    yf = np.minimum(np.arange(ny) / 2.0, phic.shape[1] - 1)
    zf = np.minimum(np.arange(nz) / 2.0, phic.shape[2] - 1)
    x0 = np.floor(xf).astype(int)
    x1 = np.minimum(x0 + 1, phic.shape[0] - 1)
    y0 = np.floor(yf).astype(int)
    y1 = np.minimum(y0 + 1, phic.shape[1] - 1)
    z0 = np.floor(zf).astype(int)
    z1 = np.minimum(z0 + 1, phic.shape[2] - 1)
    tx = (xf - x0)[:, None, None]
    ty = (yf - y0)[None, :, None]
    tz = (zf - z0)[None, None, :]
    c = lambda a, b, d: phic[np.ix_(a, b, d)]  # noqa: E731
    c00 = c(x0, y0, z0) * (1 - tx) + c(x1, y0, z0) * tx
    c10 = c(x0, y1, z0) * (1 - tx) + c(x1, y1, z0) * tx
    c01 = c(x0, y0, z1) * (1 - tx) + c(x1, y0, z1) * tx
    c11 = c(x0, y1, z1) * (1 - tx) + c(x1, y1, z1) * tx
    c0 = c00 * (1 - ty) + c10 * ty
    c1 = c01 * (1 - ty) + c11 * ty
    return c0 * (1 - tz) + c1 * tz


@pytest.mark.parametrize("shape", [(9, 8, 7), (10, 5, 12), (2, 3, 17)])
def test_transfer_operators_match_numpy_forms(shape):
    rng = np.random.default_rng(sum(shape))
    fine = rng.standard_normal(shape)
    assert np.array_equal(S.restrict_fw(fine), _restrict_fw_numpy(fine))
    cs = tuple((n + 1) // 2 for n in shape)
    coarse = rng.standard_normal(cs)
    fixed = rng.random(shape) < 0.2
    base = rng.standard_normal(shape)
    got = S.prolong_add(coarse, base.copy(), fixed)
    want = base + np.where(fixed, 0.0, _prolongate_numpy(coarse, shape))
    assert np.array_equal(got, want)
    # block-any coarse mask by reshape (multigrid3d.py:507-526)
    px, py, pz = [n + (n & 1) for n in shape]
    fp = np.zeros((px, py, pz), bool)
    fp[:shape[0], :shape[1], :shape[2]] = fixed
    blk = fp.reshape(px // 2, 2, py // 2, 2, pz // 2, 2).any(axis=(1, 3, 5))
    assert np.array_equal(S.coarsen_any(fixed), blk)


@pytest.mark.parametrize("stencil", ["ghost_linear", "sw"])
def test_residual_matches_padded_numpy(stencil):
    """The residual num - den phi from an even-reflection padded array."""
    rng = np.random.default_rng(7)
    shape = (7, 9, 6)
    phi = rng.standard_normal(shape)
    fixed = rng.random(shape) < 0.15
    th = rng.uniform(0.05, 1.0, (6,) + shape).astype(np.float32)
    th[rng.random((6,) + shape) < 0.7] = 1.0
    p = np.pad(phi, 1, mode="reflect")
    v = [p[:-2, 1:-1, 1:-1], p[2:, 1:-1, 1:-1], p[1:-1, :-2, 1:-1],
         p[1:-1, 2:, 1:-1], p[1:-1, 1:-1, :-2], p[1:-1, 1:-1, 2:]]
    t = [th[d] for d in range(6)]
    if stencil == "ghost_linear":
        w = [1.0 / t[d].astype(np.float64) for d in range(6)]
        num = (v[0] * w[0] + v[1] * w[1] + v[2] * w[2] + v[3] * w[3]
               + v[4] * w[4] + v[5] * w[5])
        den = w[0] + w[1] + w[2] + w[3] + w[4] + w[5]
    else:
        f64 = np.float64
        s = [t[0] + t[1], t[2] + t[3], t[4] + t[5]]
        num = (v[0] / (t[0] * s[0]).astype(f64) + v[1] / (t[1] * s[0]).astype(f64)
               + v[2] / (t[2] * s[1]).astype(f64) + v[3] / (t[3] * s[1]).astype(f64)
               + v[4] / (t[4] * s[2]).astype(f64) + v[5] / (t[5] * s[2]).astype(f64))
        den = (1.0 / (t[0] * t[1]).astype(f64) + 1.0 / (t[2] * t[3]).astype(f64)
               + 1.0 / (t[4] * t[5]).astype(f64))
    want = np.where(fixed, 0.0, num - den * phi)
    assert np.array_equal(S.residual(phi, fixed, th, stencil), want)


def test_optimal_omega_formula():
    """Young's factor (solver3d.py:201-226): a cube reduces to
    2 / (1 + sin(pi / (n + 1))); a size-1 axis is left out."""
    n = 33
    want = 2 / (1 + np.sin(np.pi / (n + 1)))
    assert abs(S.optimal_omega((n, n, n)) - want) < 1e-14
    assert abs(S.optimal_omega((n, n, 1)) - want) < 1e-14
    assert S.optimal_omega((1, 1, 1)) == 1.0


# ---- the cascade (multigrid3d.solve3d_mg) ------------------------------

def _coarsen_mask_numpy(fixed, val):
    """The reference's vectorised coarsening (multigrid3d.py:84-135):
    injection, then block-any by reshape with the block's largest-|val|
    fixed node by argmax over the (di, dj, dk) axes moved last."""
    nx, ny, nz = fixed.shape
    cs = tuple((n + 1) // 2 for n in fixed.shape)
    fc = np.zeros(cs, bool)
    vc = np.zeros(cs)
    fc[...] = fixed[::2, ::2, ::2]
    vc[...] = val[::2, ::2, ::2]
    px, py, pz = [n + (n & 1) for n in fixed.shape]
    fpad = np.zeros((px, py, pz), bool)
    fpad[:nx, :ny, :nz] = fixed
    vpad = np.zeros((px, py, pz))
    vpad[:nx, :ny, :nz] = np.where(fixed, val, 0.0)
    fb = fpad.reshape(px // 2, 2, py // 2, 2, pz // 2, 2)
    vb = vpad.reshape(px // 2, 2, py // 2, 2, pz // 2, 2)
    anyf = fb.any(axis=(1, 3, 5))
    mag = np.where(fb, np.abs(vb), -1.0).transpose(0, 2, 4, 1, 3, 5)
    flat = vb.transpose(0, 2, 4, 1, 3, 5).reshape(cs + (8,))
    idx = np.argmax(mag.reshape(cs + (8,)), axis=-1)
    vany = np.take_along_axis(flat, idx[..., None], axis=-1)[..., 0]
    rescue = anyf & ~fc
    fc[rescue] = True
    vc[rescue] = vany[rescue]
    return fc, vc


@pytest.mark.parametrize("shape", [(9, 8, 7), (10, 5, 12), (3, 3, 2)])
def test_cascade_transfers_match_numpy_forms(shape):
    rng = np.random.default_rng(7 + sum(shape))
    fixed = rng.random(shape) < 0.15
    # integer values with repeated magnitudes, so argmax ties occur
    val = rng.integers(-3, 4, shape).astype(float)
    fc, vc = S.coarsen_mask(fixed, val)
    wf, wv = _coarsen_mask_numpy(fixed, val)
    assert np.array_equal(fc, wf)
    assert np.array_equal(vc[fc], wv[wf])
    cs = fc.shape
    coarse = rng.standard_normal(cs)
    assert np.array_equal(S.prolongate(coarse, shape),
                          _prolongate_numpy(coarse, shape))


def test_coarsen_mask_cases():
    """Injection wins when the injection node is metal; a free injection
    node takes the block's largest |val|, the first in (di, dj, dk) order
    on a tie; a 1-node plate on an odd plane survives coarsening."""
    f = np.zeros((4, 4, 4), bool)
    v = np.zeros(f.shape)
    f[0, 0, 0], v[0, 0, 0] = True, 1.0          # injection node
    f[1, 1, 1], v[1, 1, 1] = True, -9.0         # larger, ignored
    f[2, 0, 1], v[2, 0, 1] = True, 4.0          # block (1,0,0): tie,
    f[3, 0, 0], v[3, 0, 0] = True, -4.0         # the first (dk=1) wins
    f[0, 2, 2], v[0, 2, 2] = True, 2.0          # block (0,1,1): injection
    f[1, 3, 3], v[1, 3, 3] = True, -6.0         # node metal, keeps 2
    fc, vc = S.coarsen_mask(f, v)
    assert fc[0, 0, 0] and vc[0, 0, 0] == 1.0
    assert fc[1, 0, 0] and vc[1, 0, 0] == 4.0
    assert fc[0, 1, 1] and vc[0, 1, 1] == 2.0
    assert fc.sum() == 3
    plate = np.zeros((9, 7, 5), bool)
    plate[5, 1:6, 1:4] = True
    fc, vc = S.coarsen_mask(plate, np.where(plate, 7.0, 0.0))
    assert fc.shape == (5, 4, 3)
    assert fc[2, 1:3, 1:2].all() and fc.sum() == fc[2].sum()
    assert (vc[fc] == 7.0).all()


@pytest.mark.parametrize("shape", [(9, 7, 5), (8, 6, 4)])
def test_prolongate_reproduces_affine_fields(shape):
    """A coarse field affine in the coarse coordinates comes back exactly
    at (i/2, j/2, k/2), each clamped to the last coarse node (dyadic
    values, so the arithmetic is exact)."""
    cs = tuple((n + 1) // 2 for n in shape)
    ci, cj, ck = np.meshgrid(*[np.arange(c, dtype=float) for c in cs],
                             indexing="ij")
    a = 3.0 + 2.0 * ci - 5.0 * cj + 0.5 * ck
    got = S.prolongate(a, shape)
    xs = [np.minimum(np.arange(n) / 2.0, c - 1) for n, c in zip(shape, cs)]
    fi, fj, fk = np.meshgrid(*xs, indexing="ij")
    assert np.array_equal(got, 3.0 + 2.0 * fi - 5.0 * fj + 0.5 * fk)


def test_cascade_reaches_the_sor_fixed_point():
    """solve3d_mg converges to plain SOR's answer at the same tolerance
    with fewer fine sweeps; on a grid too small to coarsen it is plain
    SOR, bit for bit."""
    fixed, val = _plate()
    p, sw, d, st, lv = S.solve3d_mg(fixed, val, tol=1e-8)
    q, sw_s, d_s, st_s = S.solve3d(fixed, val, tol=1e-8)
    assert st == 0 and st_s == 0 and lv == 0 and d < 1e-8
    assert sw < sw_s
    assert np.abs(p - q).max() < 1e-5
    f9, v9 = _plate(n=21, nz=9)
    p, sw, d, st, lv = S.solve3d_mg(f9, v9, tol=1e-7)
    q, sw_s, d_s, st_s = S.solve3d(f9, v9, tol=1e-7)
    assert (sw, d, st) == (sw_s, d_s, st_s) and np.array_equal(p, q)


def test_cascade_exact_quadratic_and_cap():
    n = 33
    i, j, k = np.meshgrid(*[np.arange(n, dtype=float)] * 3, indexing="ij")
    exact = i * i + j * j - 2 * k * k
    fixed = shell(n, n, n)
    phi, sw, d, st, lv = S.solve3d_mg(fixed, np.where(fixed, exact, 0.0),
                                      tol=1e-10)
    assert st == 0 and np.abs(phi - exact)[~fixed].max() < 1e-6
    fixed, val = _plate()
    _, sw, d, st, lv = S.solve3d_mg(fixed, val, tol=1e-9, max_sweeps=3)
    # the coarsest of 49x49x25 -> 25x25x13 -> 13x13x7 fails first
    assert st == 1 and lv == 2 and sw == 3 and d >= 1e-9
