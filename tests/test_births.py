"""Births without a device: the standard-normal tables the engine is given
reproduce numpy's draws bit for bit; the oracle's generate_births is
pinned by the reference's own birth tests (truncation, seeds, the
zero-direction stream, refusals), by the draw order of the uniform-only
sources against an independent vectorised reading of the same stream,
and by the distributions' closed forms; the source descriptor refuses
what the reference refuses."""
import math
import warnings

import numpy as np
import pytest

from oracle import births as OB
from ion_gym_amd import births as B
from ion_gym_amd.physics.flight_spec import SourceSpec


@pytest.mark.parametrize("seed", [0, 1, 7, 123456])
def test_ziggurat_tables_reproduce_numpy_normals(seed):
    k, w, f = B.ziggurat_tables()
    assert k[1] == 0 and f[0] == 1.0
    n = 200_000
    ref = np.random.default_rng(seed).standard_normal(n)
    words = np.random.PCG64(seed).random_raw(int(n * 1.05) + 64)
    got, used = OB.zig_normals(words, n, k, w, f)
    assert np.array_equal(got, ref)
    # the tail beyond r was exercised, and a normal takes ~1.02 words
    assert np.sum(np.abs(ref) > OB.ZIG_R) > 20
    assert n < used < 1.05 * n
    # the next word numpy draws is the one after the last consumed
    g = np.random.default_rng(seed)
    g.standard_normal(n)
    assert int(g.bit_generator.random_raw()) == int(words[used])


def _gauss(seed=4):
    # tests/test_gaussian_births.py:19-28 of the reference
    return SourceSpec(distribution="gaussian", n_ions=500, x0_mm=5.0,
                      y0_mm=5.0, fwhm_mm=[1.0, 0.5, 0.0],
                      trunc_mm=[0.6, 0.3, 0.0], mz_list=[100.0], seed=seed)


def test_oracle_gaussian_births_inside_truncation_and_seeded():
    b = OB.generate_births(_gauss(), 4)
    assert np.all(np.abs(b[:, 0] - 5.0) <= 0.6)
    assert np.all(np.abs(b[:, 1] - 5.0) <= 0.3)
    assert np.all(b[:, 2] == 0.0)
    assert np.std(b[:, 0]) > np.std(b[:, 1]) > 0.0
    assert np.array_equal(b, OB.generate_births(_gauss(), 4))
    assert not np.array_equal(b, OB.generate_births(_gauss(5), 5))


def test_oracle_refuses_a_layout_no_draw_satisfies(monkeypatch):
    monkeypatch.setattr(SourceSpec, "gaussian_layout",
                        lambda self: ([1.0, 0.0, 0.0], [1e-300, 0.0, 0.0]))
    with pytest.raises(RuntimeError, match="truncated Gaussian on axis x"):
        OB.generate_births(_gauss(), 4)


def _zero_dir(direction=(0.0, 0.0, 0.0)):
    # tests/test_zero_direction.py:37-45 of the reference (the source part)
    return SourceSpec(distribution="disc", n_ions=4, x0_mm=5.0, y0_mm=3.0,
                      z0_mm=1.0, r_mm=0.5, axis="x",
                      direction=list(direction), ke_lo=0.02, ke_hi=0.05,
                      mz_list=[120.0, 322.0], tob_span_us=0.0, seed=3)


def test_oracle_zero_direction_keeps_the_stream():
    at_rest = OB.generate_births(_zero_dir(), 3)
    aimed = OB.generate_births(_zero_dir((1.0, 0.0, 0.0)), 3)
    assert np.array_equal(at_rest[:, 3:6], np.zeros((8, 3)))
    assert np.array_equal(at_rest[:, 0:3], aimed[:, 0:3])
    assert np.abs(aimed[:, 3]).min() > 0.0
    assert _zero_dir().zero_direction_note().startswith("WARNING")
    assert _zero_dir((1.0, 0.0, 0.0)).zero_direction_note() is None


def test_oracle_draw_order_of_uniform_sources():
    """disc + energy + birth time take exactly four uniforms per ion, in
    the order radius, angle, energy, time; box three then energy, time."""
    n = 300
    src = SourceSpec(distribution="disc", n_ions=n, x0_mm=2.0, y0_mm=0.5,
                     z0_mm=-1.0, r_mm=3.0, axis="y", ke_lo=1.0, ke_hi=5.0,
                     direction=[0.0, 0.0, 2.0], mz_list=[250.0],
                     tob_span_us=4.0)
    b = OB.generate_births(src, 11)
    u = np.random.default_rng(11).random(4 * n).reshape(n, 4)
    rr = 3.0 * np.sqrt(u[:, 0])
    th = 2 * math.pi * u[:, 1]
    # axis y: the disc lies in the (x, z) plane
    assert np.allclose(b[:, 0], 2.0 + rr * np.cos(th), rtol=0, atol=1e-14)
    assert np.all(b[:, 1] == 0.5)
    assert np.allclose(b[:, 2], -1.0 + rr * np.sin(th), rtol=0, atol=1e-14)
    ke = 1.0 + 4.0 * u[:, 2]
    sp = np.sqrt(2 * ke * OB.E_CHG / (250.0 * OB.KG_AMU)) / 1000.0
    assert np.array_equal(b[:, 5], sp) and np.all(b[:, 3:5] == 0.0)
    assert np.array_equal(b[:, 6], 4.0 * u[:, 3])
    box = SourceSpec(distribution="box", n_ions=n, box_mm=[1.0, 2.0, 4.0],
                     x0_mm=0.0, ke_lo=0.0, ke_hi=0.0, tob_span_us=1.0)
    bb = OB.generate_births(box, 5)
    u = np.random.default_rng(5).random(4 * n).reshape(n, 4)
    for k, e in enumerate((1.0, 2.0, 4.0)):
        assert np.array_equal(bb[:, k], 0.0 + e * (u[:, k] - 0.5))
    assert np.array_equal(bb[:, 6], u[:, 3])


def test_oracle_distributions_closed_forms():
    n = 20000
    disc = SourceSpec(distribution="disc", n_ions=n, x0_mm=0.0, r_mm=2.0,
                      ke_lo=0.0, ke_hi=0.0, temperature_k=300.0,
                      mz_list=[100.0], tob_span_us=3.0)
    b = OB.generate_births(disc, 2)
    r2 = (b[:, 1] ** 2 + b[:, 2] ** 2) / 4.0
    # uniform over the disc: r^2 / R^2 is U(0, 1)
    assert abs(r2.mean() - 0.5) < 0.01 and r2.max() <= 1.0
    sig = math.sqrt(OB.KB * 300.0 / (100.0 * OB.KG_AMU)) / 1000.0
    for k in (3, 4, 5):
        assert abs(b[:, k].std() / sig - 1.0) < 0.03
    assert 0.0 <= b[:, 6].min() and b[:, 6].max() < 3.0
    # masses in contiguous blocks of n_ions
    two = SourceSpec(distribution="point", n_ions=5000, ke_lo=0.0,
                     ke_hi=0.0, temperature_k=300.0, mz_list=[50.0, 800.0])
    bt = OB.generate_births(two, 1)
    s_lo, s_hi = bt[:5000, 3].std(), bt[5000:, 3].std()
    assert abs(s_lo / s_hi - 4.0) < 0.25
    line = SourceSpec(distribution="line", n_ions=11, len_mm=5.0, axis="z",
                      z0_mm=1.0, ke_lo=0.0, ke_hi=0.0)
    bl = OB.generate_births(line, 0)
    assert np.allclose(bl[:, 2], 1.0 + np.arange(11) * 0.5)
    grid = SourceSpec(distribution="grid", n_ions=9, r_mm=1.0, x0_mm=0.0,
                      ke_lo=0.0, ke_hi=0.0)
    bg = OB.generate_births(grid, 0)
    want = sorted((a, c) for a in (-1.0, 0.0, 1.0) for c in (-1.0, 0.0, 1.0))
    got = sorted((round(y, 12) + 0.0, round(z, 12) + 0.0)
                 for y, z in bg[:, 1:3])
    assert got == want


def test_descriptor_refusals():
    with pytest.raises(ValueError, match="not one this generator places"):
        B.births_desc(SourceSpec(distribution="ring"), 0)
    with pytest.raises(ValueError, match="dv_fwhm_ms must be a 3-vector"):
        B.births_desc(SourceSpec(dv_fwhm_ms=[1.0, 2.0]), 0)
    with pytest.raises(ValueError, match="cannot be negative"):
        B.births_desc(SourceSpec(dv_fwhm_ms=[1.0, -2.0, 0.0]), 0)
    with pytest.raises(ValueError, match="requires fwhm_mm"):
        B.births_desc(SourceSpec(distribution="gaussian"), 0)
    with pytest.raises(ValueError, match="below 1 sigma"):
        B.births_desc(SourceSpec(distribution="gaussian",
                                 fwhm_mm=[1.0, 0.0, 0.0],
                                 trunc_mm=[0.1, 0.0, 0.0]), 0)
    with pytest.raises(ValueError, match="v_drift_mm_us must be"):
        B.births_desc(SourceSpec(v_drift_mm_us=[1.0]), 0)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        B.births_desc(_zero_dir(), 0)
    assert any("nowhere to point" in str(c.message) for c in caught)


def test_descriptor_carries_numpy_seeding():
    d, _ = B.births_desc(SourceSpec(seed=42), None)
    st = np.random.PCG64(42).state["state"]
    assert (d.pcg_state[0] << 64 | d.pcg_state[1]) == st["state"]
    assert (d.pcg_inc[0] << 64 | d.pcg_inc[1]) == st["inc"]
    assert d.distribution == 1 and d.n_per == 100 and d.has_ke == 1


def test_births_file_rows(tmp_path):
    p = tmp_path / "b.csv"
    p.write_text("x,y,z,vx,vy,vz\n1,2,3,4,5,6\n7,8,9,10,11,12\n")
    src = SourceSpec(distribution="file", births_file="b.csv", n_ions=2)
    out = B.read_births_file(src, deck_path=str(tmp_path / "deck.json"))
    assert out.shape == (2, 7) and out[1, 5] == 12.0 and out[1, 6] == 0.0
    with pytest.raises(ValueError, match="has 2 rows"):
        B.read_births_file(SourceSpec(distribution="file",
                                      births_file=str(p), n_ions=3))
    q = tmp_path / "c.csv"
    q.write_text("x,y,z,vx,vy\n1,2,3,4,5\n")
    with pytest.raises(ValueError, match="missing column"):
        B.read_births_file(SourceSpec(distribution="file",
                                      births_file=str(q), n_ions=1))


def test_oracle_refusals_on_hand_made_grids():
    ele = np.zeros((5, 4), np.int32)
    ele[2, 1] = 7
    h = 0.5
    # node 2 spans [1.75, 2.25) in mm on axis 0 by the nearest-node rule
    a = np.array([1.0, 1.0, 1.126, 0.0, 2.0 + 5e-10, 2.0 + 1e-8, -1e-10,
                  -1e-8, 0.3])
    b = np.array([0.5, 0.74, 0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.5])
    fate, label = OB.refuse_2d(ele, h, a, b)
    # nodes: (2, 1) metal, (2, 1.48) -> (2, 1) metal, (2.252, 1) -> (2, 1),
    # origin, the last node + 1e-9 tolerance in and out, below zero in
    # and out, (0.6, 1) vacuum
    assert list(fate) == [0, 0, 0, -1, -1, 1, -1, 1, -1]
    assert list(label[:3]) == [7, 7, 7]
    m = np.zeros((3, 3, 3), bool)
    m[1:, 1:, 1:] = True
    x = np.array([1.5, 1.5, 0.5, 2.0, 2.0 + 1e-8, 1.0])
    y = np.array([1.5, 1.5, 1.5, 2.0, 2.0, 1.0])
    z = np.array([1.5, 0.5, 1.5, 2.0, 2.0, 1.0])
    # a fully metal cell; two cells with metal on one face only; the far
    # corner clamps into the metal cell; past the box; the shared corner
    assert list(OB.refuse_3d(m, 1.0, x, y, z)) == [0, -1, -1, 0, 1, 0]
