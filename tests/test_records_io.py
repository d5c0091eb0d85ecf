"""Trajectory files and results banks: round trips in both precisions,
the member layout ion_gym reads (t<i>, time<i>, summary table, _meta as
JSON bytes), skipped ions, and the bank's refusals."""
import json

import numpy as np
import pytest

from ion_gym_amd import records_io as R

COLS = ["t", "x", "y", "z", "vx", "vy", "vz", "speed"]


def _results():
    rng = np.random.default_rng(1)
    out = []
    for i in (0, 3, 7):
        tr = rng.normal(size=(5 + i, len(COLS)))
        tr[:, 0] = np.linspace(0.1, 12.3456789012345, 5 + i)
        out.append((i, tr, dict(kind=i % 3, tof=float(tr[-1, 0]),
                                mz=100.0 + i, x_end=tr[-1, 1],
                                y_end=tr[-1, 2], z_end=tr[-1, 3],
                                n_col=None)))
    out.append((9, None, {}))       # never flown: skipped
    return out


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_trajectory_file_round_trip(tmp_path, dtype):
    res = _results()
    p = R.save_trajectories(res, COLS, str(tmp_path / "run.traj.npz"),
                            label="funnel", dtype=dtype, geometry_key="abc")
    trajs, meta = R.load_trajectories(p)
    assert sorted(trajs) == [0, 3, 7] and meta["n_skipped"] == 1
    assert meta["columns"] == COLS and meta["precision"] == dtype
    for i, tr, _ in res[:3]:
        if dtype == "float64":
            assert np.array_equal(trajs[i], tr)
        else:
            # the time column is exact, the rest float32-rounded
            assert np.array_equal(trajs[i][:, 0], tr[:, 0])
            assert np.array_equal(trajs[i][:, 1:],
                                  tr[:, 1:].astype(np.float32))
    with np.load(p, allow_pickle=False) as z:
        assert list(z["summary_cols"]) == R.SUMMARY_COLS
        s = z["summary"]
        assert s.shape == (3, 7) and list(s[:, 0]) == [0.0, 3.0, 7.0]
        assert list(s[:, 6]) == [100.0, 103.0, 107.0]
        assert list(z["mz_source"]) == ["summary"] * 3
        assert str(z["label"][0]) == "funnel"
        meta2 = json.loads(bytes(z["_meta"]).decode())
        assert meta2["kind"] == "ion_gym_trajectories"
        assert ("time3" in z.files) == (dtype == "float32")


def test_trajectory_refusals(tmp_path):
    with pytest.raises(ValueError, match="does not match"):
        R.trajectory_members([(0, np.zeros((3, 2)), {})], COLS)
    with pytest.raises(ValueError, match="precision"):
        R.trajectory_members([], COLS, dtype="float16")
    p = tmp_path / "x.npz"
    np.savez(p, a=np.zeros(3))
    with pytest.raises(R.TrajectoryFormatError):
        R.read_trajectory_meta(str(p))


def test_bank_round_trip_and_refusals(tmp_path):
    arrays = dict(tof=np.arange(5.0), fate=np.array([6, 6, 0, 1, 6]))
    p = R.bank_run(tmp_path / "b1", meta=dict(operating_point={"U": 1.0}),
                   arrays=arrays, seed=3)
    assert p.endswith(".npz")
    meta, got = R.load_bank(p)
    assert meta["operating_point"] == {"U": 1.0} and meta["seed"] == 3
    assert np.array_equal(got["tof"], arrays["tof"])
    with pytest.raises(R.BankError, match="exists"):
        R.bank_run(p, meta=dict(operating_point={}), arrays=arrays)
    R.bank_run(p, meta=dict(operating_point={}), arrays=arrays,
               overwrite=True)
    with pytest.raises(R.BankError, match="operating_point"):
        R.bank_run(tmp_path / "b2", meta={}, arrays=arrays)
    with pytest.raises(R.BankError, match="no arrays"):
        R.bank_run(tmp_path / "b3", meta=dict(operating_point={}), arrays={})
    q = tmp_path / "plain.npz"
    np.savez(q, a=np.zeros(2))
    with pytest.raises(R.BankError, match="not a results bank"):
        R.load_bank(q)
