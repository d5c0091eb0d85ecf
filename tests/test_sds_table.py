"""The SDS jump table: ion_gym's own sds_jump_icdf.dat (a fixture under
tests/golden, ion_gym/physics/sds_jump_icdf.dat of the reference, GPL v3)
read by the reference's format rules, the chi-3 fallback's error against
it stated per row, and the kernels' jump normalisation on both tables.

Pins: the reader against the file's own layout (5 rows of 1002 points,
minimum .. 99.9 % .. maximum, point 1001 repeating the maximum, rows
non-decreasing); the draw against the closed form <jump^2> = 6 D dt that
fixes mfp_mm (sds.py:401); the chi-3 law against the reference's rows.
"""
import math
import os

import numpy as np
import pytest

from ion_gym_amd.physics import sds

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sds_jump_icdf.dat")


@pytest.fixture(scope="module")
def ref():
    return sds.load_diffusion_statistics(GOLDEN)


def test_the_reference_table_reads_in_its_layout(ref):
    assert ref.shape == (5, 1002)
    assert np.all(ref[:, 1001] == ref[:, 1000])          # max repeated
    assert np.all(np.diff(ref[:, :1001], axis=1) >= 0)
    # the walks are 1e5 collisions of unit mean free path: the median
    # distance is of order sqrt(1e5) and grows with the mass ratio
    med = ref[:, 500]
    assert np.all(np.diff(med) > 0)
    assert 0.5 * math.sqrt(1e5) < med[0] < 2 * math.sqrt(1e5)


def test_the_reader_refuses_a_token_by_position(tmp_path):
    p = tmp_path / "bad.dat"
    text = open(GOLDEN, encoding="utf-8").read().replace("1680.59", "x1", 1)
    p.write_text(text, encoding="utf-8")
    with pytest.raises(ValueError, match="positionally"):
        sds.load_diffusion_statistics(str(p))
    p.write_text("; only a comment\n1 2 3\n", encoding="utf-8")
    with pytest.raises(ValueError, match="expected 5010"):
        sds.load_diffusion_statistics(str(p))


def test_the_tests_fly_the_reference_table(ref):
    # tests/conftest.py points ION_GYM_SDS_TABLE at the fixture
    assert np.array_equal(sds.default_statistics(), ref)


@pytest.mark.parametrize("table", ["reference", "chi3"])
@pytest.mark.parametrize("mz", [28.0, 300.0, 3000.0, 50000.0])
def test_mean_square_jump_is_6_D_dt(ref, table, mz):
    st = ref if table == "reference" else sds.chi3_statistics()
    p = sds.transport_params(mz, 1, 28.0134, 0.366, 300.0, 4.0,
                             gas_name="N2",
                             species=sds.declared_K0(mz, "N2", 1.5),
                             stats=st)
    dt = 1e-3
    rng = np.random.default_rng(5)
    d = sds.icdf_steps_at(st, p["log_mr_ratio"], rng.random(400000))
    ms = np.mean((p["V_mm_us"] / p["mfp_mm"] * dt / 1e5) * d * d
                 * p["mfp_mm"] ** 2)
    assert ms == pytest.approx(6 * p["D_mm2_per_us"] * dt, rel=0.01)


# the chi-3 law over the reference's table, both normalised to the
# kernels' mean square, at quantiles 1 % .. 99.9 %: the largest relative
# deviation per row (mass ratio 10^k)
CHI3_MAX_DEV = {0: 0.005, 1: 0.008, 2: 0.008, 3: 0.021, 4: 0.135}


@pytest.mark.parametrize("row", range(5))
def test_chi3_fallback_error_per_row(ref, row):
    chi = sds.chi3_statistics()
    u = np.array([0.01, 0.05, 0.1, 0.25, 0.5, 0.75, 0.9, 0.95, 0.99, 0.999])
    r = sds.icdf_steps_at(ref, float(row), u) / math.sqrt(
        sds.mean_sq_jump_steps(float(row), ref))
    c = sds.icdf_steps_at(chi, float(row), u) / math.sqrt(
        sds.mean_sq_jump_steps(float(row), chi))
    dev = np.max(np.abs(c / r - 1.0))
    assert dev <= CHI3_MAX_DEV[row]
    if row == 4:
        assert dev > 0.10          # the heavy-ion tail is where it departs
