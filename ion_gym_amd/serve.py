"""Continuous-batching flight service.

Stands in for the reference's interactive callers -- the workbench flying
an ensemble per click (ion_gym/ui/sim_app.py:7002, run_threaded) and the
CMA-ES optimizer flying a candidate tune per evaluation
(ion_gym/physics/optimize.py:273-310) -- which in the reference each build
a fly function and fly their ensemble alone.

Here the field pack stays resident and a fixed slab of ion slots stays
resident beside it.  Requests arrive at any time; each brings its ions and
its tune (drive amplitudes/frequencies/phases, field-set scales, step,
time limit, RNG key).  Between device windows the server

  * admits waiting requests in arrival order while a contiguous run of free
    slots and the HBM budget allow (each ion costs ``BYTES_PER_ION`` of
    device state), giving each a row of the slab's tune table;
  * flies one window of ``window`` steps over every live ion of every
    admitted request in ONE launch (ig_fly over the slab);
  * asks the device how many ions of each tune still fly
    (ig_ensemble_live_by_tune) and, for every request with none left, reads
    its slots back, frees them and its tune row, and records its latency.

Requests that arrive while others fly are admitted at the next window, so
their work overlaps on the device.  An ion's result does not depend on the
mix: its random stream is keyed by (seed, request key) and counted per ion,
the kernels' per-ion operation order does not depend on the lane, batch or
window, and nothing is reduced across ions.
"""
import collections
import threading
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib as L

# device bytes one resident ion costs (state, per-ion scalars, counters)
BYTES_PER_ION = 8 * 19 + 8 * 3 + 8 * 4 + 4 * 6


@dataclass
class FlightRequest:
    """One request: its ions' per-ion arrays and its tune.  With a
    ``source`` (a SourceSpec), ``ions`` may omit the births (x..vz, tob):
    they are drawn on the device into the request's slots at admission
    (ion_gym_amd.births), so the request's births never exist on the
    host."""
    ions: dict
    tune: dict
    source: Optional[object] = None
    rid: int = -1
    t_submit: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    result: Optional[dict] = None
    steps: int = 0
    refused: int = 0

    @property
    def n(self):
        return len(self.ions["x"] if "x" in self.ions else
                   self.ions["m_amu"])

    @property
    def latency(self):
        return self.t_done - self.t_submit


def make_tune(fields_pack, *, dt, t_max, rng_key=0, set_scale=None,
              ch_kind=None, ch_om=None, ch_ph=None, ch_duty=None,
              ch_amp=None, ch_off=None):
    """A tune dict, defaulting every waveform field to the pack's own."""
    K = len(fields_pack.get("ch_kind", ()))
    pick = (lambda v, key, d: np.asarray(v if v is not None else
                                         fields_pack.get(key, d), float))
    return dict(
        dt=float(dt), t_max=float(t_max), rng_key=int(rng_key),
        set_scale=(np.ones(K + 2) if set_scale is None
                   else np.asarray(set_scale, float)),
        ch_kind=np.asarray(ch_kind if ch_kind is not None
                           else fields_pack.get("ch_kind", np.zeros(0)),
                           np.int32),
        ch_om=pick(ch_om, "ch_om", np.zeros(K)),
        ch_ph=pick(ch_ph, "ch_ph", np.zeros(K)),
        ch_duty=pick(ch_duty, "ch_duty", np.full(K, 0.5)),
        ch_amp=pick(ch_amp, "ch_amp", np.ones(K)),
        ch_off=pick(ch_off, "ch_off", np.zeros(K)))


def _fill_tune(t, d):
    t.dt = d["dt"]
    t.t_max = d["t_max"]
    t.rng_key = d["rng_key"] & 0xFFFFFFFFFFFFFFFF
    sc = d["set_scale"]
    for s in range(L.MAX_CHANNELS + 2):
        t.set_scale[s] = float(sc[s]) if s < len(sc) else 1.0
    K = len(d["ch_kind"])
    for c in range(L.MAX_CHANNELS):
        t.ch_kind[c] = int(d["ch_kind"][c]) if c < K else 0
        t.ch_om[c] = float(d["ch_om"][c]) if c < K else 0.0
        t.ch_ph[c] = float(d["ch_ph"][c]) if c < K else 0.0
        t.ch_duty[c] = float(d["ch_duty"][c]) if c < K else 0.5
        t.ch_amp[c] = float(d["ch_amp"][c]) if c < K else 1.0
        t.ch_off[c] = float(d["ch_off"][c]) if c < K else 0.0


class FlightServer:
    """Serve flight requests on one engine and one resident field pack."""

    def __init__(self, engine, fields, flight, *, capacity,
                 max_requests=256, window=2000, hbm_budget=None,
                 refuse=True):
        self.engine = engine
        # the fly functions' birth refusals at admission (an ion born in
        # metal or outside the domain is never flown; face -2)
        self.refuse = bool(refuse)
        self.fields = fields
        self.flight = flight
        self.flight.max_steps = int(window)
        self.capacity = int(capacity)
        self.window = int(window)
        self.hbm_budget = (int(hbm_budget) if hbm_budget is not None
                           else self.capacity * BYTES_PER_ION)
        self.slab = engine.ensemble(self.capacity)
        self.max_requests = int(max_requests)
        self.tunes = (L.Tune * self.max_requests)()
        idle = dict(dt=flight.dt, t_max=flight.t_max, rng_key=0,
                    set_scale=np.ones(L.MAX_CHANNELS + 2),
                    ch_kind=np.zeros(0, np.int32), ch_om=np.zeros(0),
                    ch_ph=np.zeros(0), ch_duty=np.zeros(0),
                    ch_amp=np.zeros(0), ch_off=np.zeros(0))
        for row in range(self.max_requests):   # rows not in use yet
            _fill_tune(self.tunes[row], idle)
        self.free_tunes = list(range(self.max_requests))[::-1]
        self.free = [(0, self.capacity)]          # sorted (offset, length)
        self.pending = collections.deque()
        self.active = {}                          # tune row -> request
        self.slots = {}                           # tune row -> (off, n)
        self.done = []
        self.hbm_used = 0
        self.next_rid = 0
        self.lock = threading.Lock()
        self.windows = 0
        self.ion_steps = 0
        self.device_s = 0.0

    # ------------------------------------------------------------- intake
    def submit(self, req: FlightRequest):
        with self.lock:
            req.rid = self.next_rid
            self.next_rid += 1
            if not req.t_submit:
                req.t_submit = time.perf_counter()
            if req.n > self.capacity:
                raise ValueError(f"request {req.rid}: {req.n} ions exceed "
                                 f"the slab capacity {self.capacity}")
            self.pending.append(req)
        return req.rid

    def _take_slots(self, n):
        for k, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (off + n, ln - n)
                return off
        return None

    def _give_slots(self, off, n):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged

    def _admit(self):
        dirty = False
        while True:
            with self.lock:
                if not self.pending:
                    break
                req = self.pending[0]
                need = req.n * BYTES_PER_ION
                if (not self.free_tunes
                        or self.hbm_used + need > self.hbm_budget):
                    break
                off = self._take_slots(req.n)
                if off is None:
                    break
                self.pending.popleft()
            row = self.free_tunes.pop()
            _fill_tune(self.tunes[row], req.tune)
            dirty = True
            self.slots[row] = (off, req.n)
            self.active[row] = req
            self.hbm_used += need
            req.t_admit = time.perf_counter()
            self._set_tunes()   # the new row must exist before the load
            dirty = False
            from .engine import _Keep, make_ions
            keep = _Keep()
            io = make_ions(keep, req.ions)
            L.check(L.lib().ig_ensemble_load_at(self.engine.h, self.slab.id,
                                                off, io, row),
                    "ig_ensemble_load_at")
            if req.source is not None and "x" not in req.ions:
                from .births import births_into
                births_into(self.engine, self.slab, off, req.source)
            if self.refuse:
                from .births import refuse_births
                req.refused = refuse_births(self.engine, self.fields,
                                            self.slab, off, req.n)
        if dirty:
            self._set_tunes()

    def _set_tunes(self):
        L.check(L.lib().ig_ensemble_set_tunes(self.engine.h, self.slab.id,
                                              self.max_requests, self.tunes),
                "ig_ensemble_set_tunes")

    # --------------------------------------------------------------- step
    def step(self):
        """Admit, fly one window, retire finished requests.  Returns the
        requests completed by this step."""
        self._admit()
        if not self.active:
            return []
        import ctypes as C
        t0 = time.perf_counter()
        # step total after admission (loading resets the new slots' counts)
        steps0 = self.slab.steps()
        self.engine.fly(self.fields, self.flight, self.slab)
        live = (C.c_int64 * self.max_requests)()
        L.check(L.lib().ig_ensemble_live_by_tune(self.engine.h, self.slab.id,
                                                 self.max_requests, live),
                "ig_ensemble_live_by_tune")
        self.device_s += time.perf_counter() - t0
        self.ion_steps += self.slab.steps() - steps0
        self.windows += 1
        finished = []
        for row, req in list(self.active.items()):
            if live[row] == 0:
                off, n = self.slots.pop(row)
                res = self.engine_read(off, n)
                req.result = res
                req.steps = int(res["steps"].sum())
                req.t_done = time.perf_counter()
                del self.active[row]
                self.free_tunes.append(row)
                self._give_slots(off, n)
                self.hbm_used -= n * BYTES_PER_ION
                finished.append(req)
        self.done.extend(finished)
        return finished

    def engine_read(self, off, n):
        from .engine import read_slice
        return read_slice(self.engine, self.slab, off, n)

    @property
    def busy(self):
        return bool(self.active or self.pending)

    def drain(self, max_windows=1_000_000):
        """Step until every submitted request has finished."""
        for _ in range(max_windows):
            if not self.busy:
                return
            self.step()
        raise RuntimeError("FlightServer.drain: window budget exhausted")

    def run_arrivals(self, arrivals, max_windows=1_000_000, progress=None,
                     every=50):
        """Replay (arrival_time_s, request) pairs against the wall clock:
        each request is submitted once its arrival time has passed, while
        the server keeps flying.  ``progress(n_done)`` is called each time
        another ``every`` requests have finished.  Returns the finished
        requests."""
        arrivals = sorted(arrivals, key=lambda a: a[0])
        t0 = time.perf_counter()
        k = 0
        reported = 0
        for _ in range(max_windows):
            if progress is not None and len(self.done) >= reported + every:
                reported = len(self.done) - len(self.done) % every
                progress(len(self.done))
            now = time.perf_counter() - t0
            while k < len(arrivals) and arrivals[k][0] <= now:
                req = arrivals[k][1]
                req.t_submit = t0 + arrivals[k][0]
                self.submit(req)
                k += 1
            if not self.busy:
                if k >= len(arrivals):
                    break
                time.sleep(max(0.0, min(0.01, arrivals[k][0] - now)))
                continue
            self.step()
        return self.done

    def release(self):
        self.slab.release()


def latency_stats(done):
    lat = np.array([r.latency for r in done])
    if len(lat) == 0:
        return dict(p50_s=None, p99_s=None, n=0)
    return dict(p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), n=int(len(lat)))
