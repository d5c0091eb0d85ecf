"""Ensemble driver: run / fly_all / run_threaded over a fly function.

The public contract of ion_gym/physics/ensemble_driver.py:44-571, kept
name for name: IonResult, EnsembleProgress, StopToken, RunHandle,
default_workers, run, fly_all, run_threaded and the viability predicates,
with the same guarantees -- results in index order in the final progress,
identical at any worker count, a failed ion named in the re-raised error,
a stop that finishes only the work in flight, progress once per batch.

What changes is how ions are flown.  When the fly function carries
``fly_fn.batch(indices)`` (the engine-backed fly functions of
ion_gym_amd.physics.build_rz do), the driver hands the engine whole
batches: one device launch flies thousands of ions, and the stop token and
viability predicates are consulted between batches.  A plain fly function
is flown ion by ion, on one thread or a pool of daemon threads, as the
reference does.
"""
import queue
import sys
import threading
import time
import traceback
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

# ions handed to one fly_fn.batch call when the caller does not choose
DEFAULT_ENGINE_BATCH = 65536


@dataclass
class IonResult:
    index: int
    traj: Optional[np.ndarray]
    summary: dict
    full: Optional[np.ndarray] = None


@dataclass
class EnsembleProgress:
    done: int
    total: int
    elapsed_s: float
    n_terminated: int
    stopped_early: bool = False
    stop_reason: str = ""
    results: list = field(default_factory=list)
    n_workers: int = 1
    in_flight: int = 0
    provenance: Optional[dict] = None

    @property
    def fraction(self) -> float:
        return self.done / self.total if self.total else 1.0

    @property
    def rate_hz(self) -> float:
        return self.done / self.elapsed_s if self.elapsed_s > 0 else 0.0

    @property
    def eta_s(self) -> float:
        r = self.rate_hz
        return (self.total - self.done) / r if r > 0 else float("nan")


def _decimate(traj, stride):
    """The record thinned to every stride-th row, the terminal row always
    kept (it holds the fate's state); a copy, never a view of the
    flight's buffer (ensemble_driver.py:90-98)."""
    if traj is None:
        return None
    n = traj.shape[0]
    if stride <= 1 or n <= 2:
        return traj.copy()
    keep = np.zeros(n, bool)
    keep[::stride] = True
    keep[n - 1] = True
    return traj[keep]


class StopToken:
    """A cooperative stop shared between threads: ``stop(reason)`` from any
    thread, ``stopped`` and ``reason`` polled by the driver between
    batches (ensemble_driver.py:101-116).  The latest stop's reason is
    the one reported."""

    __slots__ = ("_lock", "_why")

    def __init__(self):
        self._lock = threading.Lock()
        self._why = None

    def stop(self, reason: str = "user requested"):
        with self._lock:
            self._why = str(reason)

    @property
    def stopped(self) -> bool:
        return self._why is not None

    @property
    def reason(self) -> str:
        return self._why or ""


def default_workers() -> int:
    import os
    return max(1, (os.cpu_count() or 2) - 2)


class _Bank:
    def __init__(self, n_ions, decimate, keep_full, keep_traj):
        self.n_ions = n_ions
        self.decimate = decimate
        self.keep_full = keep_full
        self.keep_traj = keep_traj
        self.results = []
        self.n_term = 0
        self.t0 = time.perf_counter()

    def add(self, i, traj, summary):
        if summary.get("kind", -1) == 0:
            self.n_term += 1
        self.results.append(IonResult(
            index=int(i),
            traj=(_decimate(traj, self.decimate) if self.keep_traj
                  else None),
            summary=summary,
            full=(traj.copy() if (self.keep_full and self.keep_traj
                                  and traj is not None) else None)))

    def progress(self, n_workers=1, in_flight=0):
        p = EnsembleProgress(done=len(self.results), total=self.n_ions,
                             elapsed_s=time.perf_counter() - self.t0,
                             n_terminated=self.n_term, results=self.results)
        p.n_workers = n_workers
        p.in_flight = in_flight
        return p


def _iter_engine(n_ions, fly_fn, *, check_every, bank, viability,
                 stop_token, engine_batch):
    """Batches of engine_batch ions through fly_fn.batch, in index order;
    one progress per batch (check_every paces only per-ion flights)."""
    step = max(1, int(engine_batch))
    for i0 in range(0, n_ions, step):
        if stop_token is not None and stop_token.stopped:
            p = bank.progress()
            p.stopped_early, p.stop_reason = True, stop_token.reason
            yield p
            return
        idx = list(range(i0, min(n_ions, i0 + step)))
        try:
            out = fly_fn.batch(idx)
        except Exception as e:
            raise RuntimeError(
                f"ensemble worker failed on ions {idx[0]}..{idx[-1]} "
                f"({len(bank.results)}/{n_ions} completed before the "
                f"failure): {type(e).__name__}: {e}") from e
        for i, (traj, summary) in zip(idx, out):
            bank.add(i, traj, summary)
        p = bank.progress()
        reason = None
        if stop_token is not None and stop_token.stopped:
            reason = stop_token.reason
        elif viability is not None:
            reason = viability(bank.results, p)
        if reason:
            p.stopped_early, p.stop_reason = True, reason
            yield p
            return
        yield p


def _iter_sequential(n_ions, fly_fn, *, check_every, bank, viability,
                     stop_token):
    try:
        for i in range(n_ions):
            if stop_token is not None and stop_token.stopped:
                p = bank.progress()
                p.stopped_early, p.stop_reason = True, stop_token.reason
                yield p
                return
            traj, summary = fly_fn(i)
            bank.add(i, traj, summary)
            if (i + 1) % check_every == 0 or i + 1 == n_ions:
                p = bank.progress()
                reason = None
                if stop_token is not None and stop_token.stopped:
                    reason = stop_token.reason
                elif viability is not None:
                    reason = viability(bank.results, p)
                if reason:
                    p.stopped_early, p.stop_reason = True, reason
                    yield p
                    return
                yield p
    except KeyboardInterrupt:
        p = bank.progress()
        p.stopped_early, p.stop_reason = True, "KeyboardInterrupt"
        yield p


def _iter_pool(n_ions, fly_fn, *, check_every, bank, viability, stop_token,
               n_workers):
    """Daemon-thread pool, at most 2*n_workers ions outstanding; a failure
    or a stop halts submission and drains what is in flight."""
    work, done = queue.SimpleQueue(), queue.SimpleQueue()
    halt = threading.Event()

    def worker():
        for i in iter(work.get, None):
            if halt.is_set() or (stop_token is not None
                                 and stop_token.stopped):
                done.put((i, None, None, "skipped"))
                continue
            try:
                traj, summary = fly_fn(i)
                done.put((i, traj, summary, None))
            except BaseException as e:  # noqa: BLE001 -- re-raised below
                done.put((i, None, None, e))

    threads = [threading.Thread(target=worker, daemon=True,
                                name=f"ion-worker-{k}")
               for k in range(n_workers)]
    for t in threads:
        t.start()
    failure = None
    reason = None
    try:
        nxt, out, last = 0, 0, 0
        while True:
            stopping = failure is not None or reason is not None or (
                stop_token is not None and stop_token.stopped)
            while not stopping and nxt < n_ions and out < 2 * n_workers:
                work.put(nxt)
                nxt += 1
                out += 1
            if out == 0:
                break
            i, traj, summary, err = done.get()
            out -= 1
            if err == "skipped":
                continue
            if err is not None:
                if failure is None:
                    failure = (i, err)
                halt.set()
                continue
            bank.add(i, traj, summary)
            if len(bank.results) - last < check_every and out:
                continue
            last = len(bank.results)
            p = bank.progress(n_workers, out)
            if reason is None and stop_token is not None \
                    and stop_token.stopped:
                reason = stop_token.reason
            if reason is None and failure is None and viability is not None:
                reason = viability(bank.results, p)
                if reason:
                    if stop_token is not None:
                        stop_token.stop(reason)
                    halt.set()
            if reason is None and failure is None:
                yield p
        if failure is not None:
            i, e = failure
            bank.results.sort(key=lambda r: r.index)
            raise RuntimeError(
                f"ensemble worker failed on ion {i} "
                f"({len(bank.results)}/{n_ions} completed before the "
                f"failure): {type(e).__name__}: {e}") from e
        if reason is None and stop_token is not None and stop_token.stopped:
            reason = stop_token.reason
        bank.results.sort(key=lambda r: r.index)
        p = bank.progress(n_workers, 0)
        if reason is not None:
            p.stopped_early, p.stop_reason = True, reason
        yield p
    finally:
        for _ in threads:
            work.put(None)


def _iter_batches(n_ions, fly_fn, *, check_every=25, decimate=1,
                  keep_full=False, keep_traj=True, viability=None,
                  stop_token=None, n_workers=1,
                  engine_batch=DEFAULT_ENGINE_BATCH):
    """The shared loop, yielding an EnsembleProgress per batch; the final
    yield holds every result in index order."""
    bank = _Bank(n_ions, decimate, keep_full, keep_traj)
    if getattr(fly_fn, "batch", None) is not None:
        yield from _iter_engine(n_ions, fly_fn, check_every=check_every,
                                bank=bank, viability=viability,
                                stop_token=stop_token,
                                engine_batch=engine_batch)
    elif int(n_workers) > 1:
        yield from _iter_pool(n_ions, fly_fn, check_every=check_every,
                              bank=bank, viability=viability,
                              stop_token=stop_token,
                              n_workers=int(n_workers))
    else:
        yield from _iter_sequential(n_ions, fly_fn,
                                    check_every=check_every, bank=bank,
                                    viability=viability,
                                    stop_token=stop_token)


def run(n_ions: int, fly_fn: Callable[[int], tuple], *, check_every=25,
        decimate=1, keep_full=False, keep_traj=True, viability=None,
        stop_token: Optional[StopToken] = None, on_progress=None,
        n_workers=1, engine_batch=DEFAULT_ENGINE_BATCH):
    """Synchronous run; returns the final EnsembleProgress."""
    last = None
    for prog in _iter_batches(n_ions, fly_fn, check_every=check_every,
                              decimate=decimate, keep_full=keep_full,
                              keep_traj=keep_traj, viability=viability,
                              stop_token=stop_token, n_workers=n_workers,
                              engine_batch=engine_batch):
        last = prog
        if on_progress is not None:
            on_progress(prog)
        if (stop_token is not None and stop_token.stopped
                and not prog.stopped_early):
            prog.stopped_early = True
            prog.stop_reason = stop_token.reason
            return prog
    if last is None:
        last = EnsembleProgress(0, n_ions, 0.0, 0)
    return last


class RunHandle:
    """Handle of a background run: .latest, .final, .error, .stop(),
    .done, .join()."""

    def __init__(self):
        self.latest = None
        self.final = None
        self.error = None
        self.error_tb = None
        self._stop = StopToken()
        self._thread = None

    def stop(self, reason: str = "user requested"):
        self._stop.stop(reason)

    @property
    def done(self) -> bool:
        return self.final is not None or self.error is not None

    def join(self, timeout=None):
        if self._thread is not None:
            self._thread.join(timeout)
        return self.final


def fly_all(n_ions: int, fly_fn, *, n_workers: int, keep_traj=True,
            on_progress=None, check_every=25,
            engine_batch=DEFAULT_ENGINE_BATCH) -> list:
    """[(traj, summary)] for ions 0..n_ions-1 in index order; raises unless
    every ion returned."""
    n_ions = int(n_ions)
    if n_ions < 0:
        raise ValueError(f"fly_all: n_ions must be >= 0, got {n_ions}")
    prog = run(n_ions, fly_fn, keep_traj=keep_traj, n_workers=n_workers,
               on_progress=on_progress, check_every=check_every,
               engine_batch=engine_batch)
    res = sorted(prog.results, key=lambda r: r.index)
    if len(res) != n_ions or prog.stopped_early:
        raise RuntimeError(
            f"fly_all: {len(res)} of {n_ions} ions returned"
            + (f" (stopped: {prog.stop_reason})" if prog.stopped_early
               else ""))
    return [(r.traj, r.summary) for r in res]


def run_threaded(n_ions: int, fly_fn, *, check_every=25, decimate=1,
                 keep_full=False, keep_traj=True, viability=None,
                 n_workers=1, engine_batch=DEFAULT_ENGINE_BATCH) -> RunHandle:
    """run() on a daemon thread; returns a RunHandle whose .latest follows
    the progress, .stop() asks the run to stop, and .final (with .error
    and .error_tb when the run raised: the traceback is also written to
    stderr, and the progress so far is kept) marks it done
    (ensemble_driver.py:495-533)."""
    handle = RunHandle()

    def track(prog):
        handle.latest = prog

    def body():
        try:
            handle.final = run(n_ions, fly_fn, check_every=check_every,
                               decimate=decimate, keep_full=keep_full,
                               keep_traj=keep_traj, viability=viability,
                               stop_token=handle._stop, on_progress=track,
                               n_workers=n_workers,
                               engine_batch=engine_batch)
        except Exception as exc:  # noqa: BLE001 -- kept on the handle
            tb = "".join(traceback.format_exception(type(exc), exc,
                                                    exc.__traceback__))
            handle.error, handle.error_tb = exc, tb
            handle.final = handle.latest
            sys.stderr.write(tb)
            sys.stderr.flush()

    handle._thread = threading.Thread(target=body, daemon=True,
                                      name="ensemble-run")
    handle._thread.start()
    return handle


# ------------------------------------------------------------ viability
# Predicates a run consults after each batch: each returns the reason to
# stop (a string) or None (ensemble_driver.py:535-571).

class viability_transmission:
    """Stop once, after at least ``after`` ions, the fraction that ended in
    an impact (kind 0) is below ``min_fraction``."""

    def __init__(self, min_fraction: float, after: int = 50):
        self.min_fraction = float(min_fraction)
        self.after = int(after)

    def __call__(self, results, prog):
        if prog.done < self.after:
            return None
        frac = prog.n_terminated / prog.done
        if frac >= self.min_fraction:
            return None
        return (f"transmission {frac:.0%} < {self.min_fraction:.0%} "
                f"after {prog.done} ions")


class viability_beam_shrinks:
    """Stop when the beam is not confined: over the ions so far with a
    finite final radius (``radius_of(summary)``), the mean of the last
    third exceeds 1 / min_ratio times the mean of the first third."""

    def __init__(self, radius_of, after: int = 50, min_ratio: float = 0.9):
        self.radius_of = radius_of
        self.after = int(after)
        self.limit = 1.0 / float(min_ratio)

    def __call__(self, results, prog):
        if prog.done < self.after:
            return None
        radii = np.fromiter((self.radius_of(r.summary) for r in results),
                            np.float64, len(results))
        radii = radii[np.isfinite(radii)]
        if radii.size < self.after:
            return None
        third = radii.size // 3
        first, last = radii[:third].mean(), radii[radii.size - third:].mean()
        if first <= 0 or last / first <= self.limit:
            return None
        return (f"beam not confined: final radius grew "
                f"{last / first:.2f}x across the run")
