"""The ensemble driver's contract (ion_gym tests/test_ensemble_driver.py
analogue): index order, identical results at any worker count, a failing
ion named, stop finishing only in-flight work, zero ions, and the batch
path taken when the fly function carries .batch."""
import threading
import time

import numpy as np
import pytest

from ion_gym_amd.physics import ensemble_driver as ED


def _fly(i):
    rng = np.random.default_rng(1000 + i)
    time.sleep(rng.uniform(0, 0.002))
    traj = np.c_[np.arange(5.0) + i, rng.normal(size=(5, 2))]
    return traj, {"kind": i % 3, "tof": float(i), "x_end": traj[-1, 1]}


def _batched():
    calls = []

    def fly(i):
        return _fly(i)

    def batch(idx):
        calls.append(list(idx))
        return [_fly(i) for i in idx]

    fly.batch = batch
    return fly, calls


@pytest.mark.parametrize("workers", [1, 4])
def test_results_in_index_order_and_worker_invariant(workers):
    prog = ED.run(40, _fly, n_workers=workers, check_every=7)
    assert [r.index for r in prog.results] == list(range(40))
    ref = ED.run(40, _fly, n_workers=1)
    for a, b in zip(prog.results, ref.results):
        assert np.array_equal(a.traj, b.traj) and a.summary == b.summary
    assert prog.n_terminated == sum(1 for i in range(40) if i % 3 == 0)


def test_progress_once_per_check_every():
    seen = []
    ED.run(30, _fly, check_every=10, on_progress=lambda p: seen.append(
        p.done))
    assert seen == [10, 20, 30]


@pytest.mark.parametrize("workers", [1, 3])
def test_failing_ion_is_named(workers):
    def bad(i):
        if i == 13:
            raise ValueError("boom")
        return _fly(i)
    with pytest.raises((RuntimeError, ValueError), match="13|boom"):
        ED.fly_all(30, bad, n_workers=workers)


def test_stop_finishes_in_flight_only():
    tok = ED.StopToken()

    def slow(i):
        if i == 5:
            tok.stop("test stop")
        return _fly(i)
    prog = ED.run(100, slow, stop_token=tok, check_every=1)
    assert prog.stopped_early and prog.stop_reason == "test stop"
    assert prog.done <= 7


def test_zero_ions_and_fly_all_refusal():
    assert ED.fly_all(0, _fly, n_workers=2) == []
    with pytest.raises(ValueError):
        ED.fly_all(-1, _fly, n_workers=1)


def test_batch_path_is_taken_and_equals_per_ion():
    fly, calls = _batched()
    out = ED.fly_all(50, fly, n_workers=1, engine_batch=16)
    assert [len(c) for c in calls] == [16, 16, 16, 2]
    ref = [_fly(i) for i in range(50)]
    for (a, sa), (b, sb) in zip(out, ref):
        assert np.array_equal(a, b) and sa == sb


def test_batch_path_stops_between_batches_and_reports_failures():
    fly, calls = _batched()
    tok = ED.StopToken()
    seen = []

    def on_prog(p):
        seen.append(p.done)
        tok.stop("enough")
    prog = ED.run(100, fly, stop_token=tok, on_progress=on_prog,
                  engine_batch=10)
    assert prog.stopped_early and prog.done == 10

    def bad_batch(idx):
        raise ValueError("device refused")
    fly.batch = bad_batch
    with pytest.raises(RuntimeError, match="ions 0..9"):
        ED.fly_all(20, fly, n_workers=1, engine_batch=10)


def test_run_threaded_handle():
    h = ED.run_threaded(20, _fly, n_workers=2)
    h.join(30)
    assert h.done and h.error is None
    assert h.final.done == 20
    h2 = ED.run_threaded(10, lambda i: 1 / 0)
    h2.join(30)
    assert h2.done and isinstance(h2.error, Exception)


def test_viability_transmission_aborts():
    v = ED.viability_transmission(0.9, after=10)
    prog = ED.run(60, _fly, viability=v, check_every=10)
    assert prog.stopped_early and "transmission" in prog.stop_reason
