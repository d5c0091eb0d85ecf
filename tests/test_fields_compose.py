"""Host field composition: the vectorised electrode-aware 3-D gradient
against the reference's per-node loop rule (tracer3d.py:54-90) written
out literally on a small grid, and the r-z term lists against a direct
accumulation."""
import numpy as np

from ion_gym_amd.physics import fields_3d, fields_rz


def _loop_rule(phi, ele, h):
    d = np.gradient(phi, h)
    dx, dy, dz = (np.array(v) for v in d)
    nx, ny, nz = phi.shape
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if not ele[i, j, k]:
                    continue
                for arr, di, dj, dk, n, c in ((dx, 1, 0, 0, nx, i),
                                              (dy, 0, 1, 0, ny, j),
                                              (dz, 0, 0, 1, nz, k)):
                    vp = c + 1 < n and not ele[i + di, j + dj, k + dk]
                    vm = c - 1 >= 0 and not ele[i - di, j - dj, k - dk]
                    if vp and not vm:
                        arr[i, j, k] = (phi[i + di, j + dj, k + dk]
                                        - phi[i, j, k]) / h
                    elif vm and not vp:
                        arr[i, j, k] = (phi[i, j, k]
                                        - phi[i - di, j - dj, k - dk]) / h
    return -dx, -dy, -dz


def test_3d_electrode_aware_gradient_matches_the_loop_rule():
    rng = np.random.default_rng(0)
    phi = rng.normal(size=(9, 8, 11))
    ele = rng.uniform(size=phi.shape) < 0.3
    ele[0, :, :] = True          # a metal face on the array edge
    got = fields_3d.build_field_aware_3d(phi, ele, 0.05)
    want = _loop_rule(phi, ele, 0.05)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    # plain gradient where there is no metal
    plain = fields_3d.build_field_aware_3d(phi, np.zeros_like(ele), 0.05)
    assert np.array_equal(plain[0], -np.gradient(phi, 0.05)[0])


def test_rz_terms_reproduce_the_accumulation():
    rng = np.random.default_rng(1)
    bases = {i: rng.normal(size=(6, 5)) * 1e4 for i in (3, 1, 7, 2)}
    dc = {1: 10.0, 2: -5.0, 7: 1.5}
    members = {1: ("rf",), 3: ("rf", "gate"), 7: ("dc2",)}
    groups = [fields_rz.DriveGroup("rf", "sin", 1e6, 50.0, 0.0, 0.5, 2.0),
              fields_rz.DriveGroup("gate", "table", 0.0, 3.0, 0.0, 0.5, 0.0,
                                   [0.0, 4.0], [0.0, 1.0]),
              fields_rz.DriveGroup("dc2", "sin", 0.0, 0.0, 0.0, 0.5, 7.0)]
    keys = sorted(bases)
    tA, tK, tG, drives, tau = fields_rz.compose_terms_rz(keys, dc, members,
                                                         groups)
    assert [d.name for d in drives] == ["rf"] and tau == 4.0
    # the reference's loop, written out
    A = np.zeros((6, 5))
    G = np.zeros((6, 5))
    B = np.zeros((6, 5))
    for idx in keys:
        fa = bases[idx] / 1e4
        A = A + dc.get(idx, 0.0) * fa
        for g in members.get(idx, ()):
            drv = {x.name: x for x in groups}[g]
            if drv.waveform == "table":
                A = A + (drv.amplitude_v * 0.0 + drv.offset_v) * fa
                G = G + drv.amplitude_v * (1.0 - 0.0) * fa
                continue
            if drv.offset_v:
                A = A + drv.offset_v * fa
            if drv.amplitude_v == 0.0:
                continue
            B += drv.amplitude_v * fa
    fas = [bases[i] / 1e4 for i in keys]
    assert np.array_equal(fields_rz.apply_terms(fas, tA, A.shape), A)
    assert np.array_equal(fields_rz.apply_terms(fas, tG, G.shape), G)
    assert np.array_equal(fields_rz.apply_terms(fas, tK[0], B.shape), B)
