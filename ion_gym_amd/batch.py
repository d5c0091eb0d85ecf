"""Many ensembles of uneven size and tune in one launch.

The reference flies an ensemble per ``build_run`` and per call of the
ensemble driver (ion_gym/physics/ensemble_driver.py:381-529); a sweep over
field strengths, drive tunes or species is a loop of such runs.  Here the
ensembles of a sweep share one device slab: each ensemble's ions occupy a
contiguous slot range and carry the index of the ensemble's row in the
slab's tune table (its drive waveforms, per-set field scales, time step,
time limit and RNG key).  One kernel launch advances every live ion of
every ensemble by a window of steps, and a loop of windows is captured
once as a HIP graph and replayed (ig_fly_graph), so a long batched flight
costs one host call per ``windows_per_call`` windows.  Results are per
ion and do not depend on which other ensembles share the launch.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .serve import _fill_tune, make_tune


class EnsembleBatch:
    """Ensembles (births + tune) resident in one slab on one engine.

    ``ensembles`` is a list of (ions, tune) pairs: ``ions`` a births dict
    as ``Ensemble.load`` takes it, ``tune`` a dict of ``make_tune``
    arguments (dt, t_max and optionally rng_key, set_scale and the
    waveform fields; unset waveform fields default to ``pack``'s)."""

    def __init__(self, engine, fields, flight, ensembles, pack=None):
        from .engine import _Keep, make_ions
        self.engine = engine
        self.fields = fields
        self.flight = flight
        self.sizes = [len(ions["x"]) for ions, _ in ensembles]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])[:-1]
        self.n = int(sum(self.sizes))
        self.slab = engine.ensemble(self.n)
        self.tunes = (L.Tune * len(ensembles))()
        for k, (_, t) in enumerate(ensembles):
            _fill_tune(self.tunes[k], make_tune(pack or {}, **t))
        L.check(L.lib().ig_ensemble_set_tunes(engine.h, self.slab.id,
                                              len(ensembles), self.tunes),
                "ig_ensemble_set_tunes")
        for k, (ions, _) in enumerate(ensembles):
            keep = _Keep()
            io = make_ions(keep, ions)
            L.check(L.lib().ig_ensemble_load_at(engine.h, self.slab.id,
                                                int(self.offsets[k]), io, k),
                    "ig_ensemble_load_at")
        self.slab.n = self.n

    def __len__(self):
        return len(self.sizes)

    def fly(self, windows=1, graph=True):
        """Advance every live ion by ``windows`` windows of
        ``flight.max_steps`` steps (asynchronous)."""
        if graph:
            self.engine.fly_graph(self.fields, self.flight, self.slab,
                                  windows)
        else:
            for _ in range(windows):
                self.engine.fly(self.fields, self.flight, self.slab)

    def live(self):
        """Live ions per ensemble."""
        out = (C.c_int64 * len(self))()
        L.check(L.lib().ig_ensemble_live_by_tune(self.engine.h, self.slab.id,
                                                 len(self), out),
                "ig_ensemble_live_by_tune")
        return np.array(out, dtype=np.int64)

    def run(self, max_windows=100_000, windows_per_call=16, graph=True):
        """Fly until every ion of every ensemble has its fate (checking
        between calls of ``windows_per_call`` windows)."""
        if self.flight.max_steps <= 0:
            raise ValueError("EnsembleBatch.run needs flight.max_steps > 0 "
                             "(the window length)")
        done = 0
        while done < max_windows:
            self.fly(windows_per_call, graph=graph)
            done += windows_per_call
            if self.live().sum() == 0:
                return done
        raise RuntimeError("EnsembleBatch.run: window budget exhausted")

    def read(self, k):
        """Results of ensemble ``k`` (the ig_ensemble_read_at slice)."""
        from .engine import read_slice
        return read_slice(self.engine, self.slab, int(self.offsets[k]),
                          int(self.sizes[k]))

    def release(self):
        self.slab.release()
