"""CPU restatement of the reference's birth generator.

TEST INFRASTRUCTURE ONLY: the checker the device births
(ion_gym_amd.births) are compared against.  The product package never
imports this module.

generate_births follows ion_gym/physics/sim_build.py:82-277 step for
step: one numpy Generator (np.random.default_rng(seed), PCG64) drawn ion
after ion, per distribution (:144-206), then the directed energy
(:217-231), the thermal spread (:232-234), the per-axis spreads
(:235-251), the drift (:252-262) and the birth time (:274).  A pure-Python
loop, for the small sources the tests fly.

zig_normals restates numpy's standard normal (the Marsaglia-Tsang
ziggurat as numpy lays out its 256 layers) on a raw PCG64 word stream, so
that the layer tables the engine is given can be pinned against numpy's
own draws.
"""
import math

import numpy as np

KB = 1.380649e-23
E_CHG = 1.602176634e-19
KG_AMU = 1.66053906660e-27
FWHM_PER_SIGMA = 2.0 * math.sqrt(2.0 * math.log(2.0))
GAUSSIAN_BIRTH_TRIES = 1000
ZIG_R = 3.6541528853610088
ZIG_INV_R = 0.27366123732975828


def generate_births(src, seed):
    """(N, 7) births of a SourceSpec-like object (sim_build.py:82-277),
    distributions point/disc/line/grid/box/gaussian."""
    rng = np.random.default_rng(seed)
    n_per = src.n_ions
    n_mz = max(1, len(src.mz_list))
    n = n_per * n_mz
    out = np.zeros((n, 7))
    ax = {"x": 0, "y": 1, "z": 2}[src.axis]
    p0 = np.array([src.x0_mm, src.y0_mm, src.z0_mm], dtype=float)
    if src.distribution == "gaussian":
        sig_g, trc_g = src.gaussian_layout()
    for i in range(n):
        pos = p0.copy()
        dist = src.distribution
        if dist in ("disc", "grid"):
            if dist == "disc":
                rr = src.r_mm * math.sqrt(rng.uniform())
                th = rng.uniform(0, 2 * math.pi)
            else:
                g = int(math.ceil(math.sqrt(n_per)))
                ib = i % n_per
                gi, gj = ib % g, ib // g
                u = (gi / max(g - 1, 1) - 0.5) * 2 * src.r_mm
                v = (gj / max(g - 1, 1) - 0.5) * 2 * src.r_mm
                rr = math.hypot(u, v)
                th = math.atan2(v, u)
            a, b = [k for k in range(3) if k != ax]
            pos[a] += rr * math.cos(th)
            pos[b] += rr * math.sin(th)
        elif dist == "line":
            ib = i % n_per
            pos[ax] += src.len_mm * (ib / max(n_per - 1, 1))
        elif dist == "box":
            bx = (list(src.box_mm) + [0.0, 0.0, 0.0])[:3]
            for k in range(3):
                pos[k] += bx[k] * (rng.random() - 0.5)
        elif dist == "gaussian":
            for k in range(3):
                if sig_g[k] <= 0.0:
                    continue
                for _ in range(GAUSSIAN_BIRTH_TRIES):
                    dpos = rng.normal() * sig_g[k]
                    if abs(dpos) <= trc_g[k]:
                        break
                else:
                    raise RuntimeError(
                        f"generate_births: no draw of the truncated Gaussian "
                        f"on axis {'xyz'[k]} fell inside the truncation")
                pos[k] += dpos
        elif dist != "point":
            raise ValueError(f"unknown distribution {dist!r}")
        out[i, 0:3] = pos
        mz = src.mz_list[(i // n_per) % n_mz]
        m_kg = mz * KG_AMU
        v = np.zeros(3)
        if src.ke_hi > 0 or src.ke_lo > 0:
            ke = rng.uniform(src.ke_lo, src.ke_hi)
            sp = math.sqrt(2 * ke * E_CHG / m_kg) / 1000.0
            d = np.array(src.direction, float)
            nrm = float(np.linalg.norm(d))
            if nrm > 0.0:
                v = v + sp * (d / nrm)
        if src.temperature_k > 0:
            sig = math.sqrt(KB * src.temperature_k / m_kg) / 1000.0
            v = v + rng.normal(size=3) * sig
        dvf = list(src.dv_fwhm_ms or [])
        if any(x for x in dvf):
            sig3 = np.array(dvf, float) * 1e-3 / FWHM_PER_SIGMA
            v = v + rng.normal(size=3) * sig3
        vd = list(getattr(src, "v_drift_mm_us", None) or [])
        if any(x for x in vd):
            v = v + np.array(vd, float)
        out[i, 3:6] = v
        out[i, 6] = rng.uniform(0, src.tob_span_us)
    return out


def zig_normals(words, n, k, w, f):
    """The first n standard normals numpy draws from the 64-bit words
    (random_standard_normal): layer i = low 8 bits, sign bit 8, 52-bit
    magnitude m above; x = m w_i returned when m < k_i; layer 0 beyond
    that is the tail past r (two log1p draws per try), any other layer
    accepts x with probability (f_{i-1} - f_i) u + f_i < exp(-x^2 / 2).
    Returns (normals, words used)."""
    out = []
    p = 0
    W = [int(v) for v in words]

    def nd():
        nonlocal p
        v = (W[p] >> 11) * (1.0 / 9007199254740992.0)
        p += 1
        return v

    while len(out) < n:
        r = W[p]
        p += 1
        i = r & 0xFF
        r >>= 8
        sign = r & 1
        m = (r >> 1) & 0x000FFFFFFFFFFFFF
        x = m * float(w[i])
        if sign:
            x = -x
        if m < int(k[i]):
            out.append(x)
            continue
        if i == 0:
            while True:
                xx = -ZIG_INV_R * math.log1p(-nd())
                yy = -math.log1p(-nd())
                if yy + yy > xx * xx:
                    out.append(-(ZIG_R + xx) if (m >> 8) & 1 else ZIG_R + xx)
                    break
            continue
        if (float(f[i - 1]) - float(f[i])) * nd() + float(f[i]) < \
                math.exp(-0.5 * x * x):
            out.append(x)
    return np.array(out), p


DOMAIN_EDGE_TOL = 1e-9


def refuse_2d(ele, h, a_mm, b_mm):
    """(fate, label) per birth of a 2-D route (r-z: a = x, b = hypot(y,
    z) on the half grid; planar: a = x, b = y): outside nodes 0..n-1 on
    either axis, 1e-9 node tolerance (birth_domain.py:27-38) -> 1;
    else nearest-node metal (raster2d.py:624-646) -> 0 with the label;
    else -1 (flies)."""
    ele = np.asarray(ele)
    n0, n1 = ele.shape
    fate = np.full(len(a_mm), -1, np.int32)
    label = np.zeros(len(a_mm), np.int32)
    for q in range(len(a_mm)):
        g0, g1 = a_mm[q] / h, b_mm[q] / h
        if not (-DOMAIN_EDGE_TOL <= g0 <= n0 - 1 + DOMAIN_EDGE_TOL
                and -DOMAIN_EDGE_TOL <= g1 <= n1 - 1 + DOMAIN_EDGE_TOL):
            fate[q] = 1
            continue
        gi, gj = int(math.floor(g0 + 0.5)), int(math.floor(g1 + 0.5))
        if 0 <= gi < n0 and 0 <= gj < n1 and ele[gi, gj] > 0:
            fate[q] = 0
            label[q] = int(ele[gi, gj])
    return fate, label


def refuse_3d(metal, h, x, y, z):
    """fate per 3-D birth (tracer3d.py:192-230): outside the domain in
    node units x / h with the 1e-9 tolerance -> 1; inside metal by the
    all-eight-corners rule of _in_metal_3d (tracer3d.py:174-189: outside
    the box is not metal, the cell index clamped to n - 2) -> 0; else -1."""
    m = np.asarray(metal).astype(bool)
    nx, ny, nz = m.shape
    inv_h = 1.0 / float(h)
    fate = np.full(len(x), -1, np.int32)
    for q in range(len(x)):
        g = (x[q] * inv_h, y[q] * inv_h, z[q] * inv_h)
        if not all(-DOMAIN_EDGE_TOL <= gv <= n - 1 + DOMAIN_EDGE_TOL
                   for gv, n in zip(g, (nx, ny, nz))):
            fate[q] = 1
            continue
        if any(gv < 0.0 or gv > n - 1 for gv, n in zip(g, (nx, ny, nz))):
            continue
        i, j, k = (min(int(gv), n - 2) for gv, n in zip(g, (nx, ny, nz)))
        if m[i:i + 2, j:j + 2, k:k + 2].all():
            fate[q] = 0
    return fate
