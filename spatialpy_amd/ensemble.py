"""Ensembles of different static models in one persistent GPU launch.

The reference runs an ensemble as one engine process per trajectory, one
after another (spatialpy/solvers/solver.py:547-605).  Here every model
(its own domain size and chemistry, one trajectory each) gets one workgroup
of a single launch that runs all of its steps; trajectory m equals the
single-model run with seed seeds[m] bit for bit.
"""
import ctypes as C

import numpy as np

from . import _abi
from .core.errors import SimulationError
from .lowering import DescHandle


def _p(a, t):
    return _abi.ptr(a, t)


class Ensemble:
    def __init__(self, systems):
        self.lib = _abi.load_library()
        self.systems = list(systems)
        self.dhs = [DescHandle(sa) for sa in self.systems]
        arr = (C.POINTER(_abi.Desc) * len(self.dhs))(*[C.pointer(d.desc) for d in self.dhs])
        self._arr = arr
        h = C.c_void_p()
        _abi.check(self.lib, self.lib.ssasdpd_ensemble_create(arr, len(self.dhs), C.byref(h)),
                   "ssasdpd_ensemble_create")
        self.h = h
        k = C.c_int32()
        self.lib.ssasdpd_ensemble_num_snapshots(self.h, C.byref(k))
        self.K = k.value
        self.Smax = max(d.desc.n_species for d in self.dhs)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ssasdpd_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.systems)

    def set_tree_width(self, pl):
        _abi.check(self.lib, self.lib.ssasdpd_ensemble_set_tree_width(self.h, int(pl)), "set_tree_width")

    def autotune(self, seeds, probe_steps=2):
        """Time probe_steps steps at each event-tree width and keep the
        fastest.  Returns (chosen width, {width: GPU ms})."""
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        ms = np.zeros(4)
        ch = C.c_int32()
        _abi.check(self.lib, self.lib.ssasdpd_ensemble_autotune(self.h, _p(seeds, C.c_uint64), int(probe_steps),
                                                                C.byref(ch), _p(ms, C.c_double)), "autotune")
        return ch.value, {w: float(t) for w, t in zip((1, 2, 4, 8), ms) if t >= 0}

    def run(self, seeds, counts=False):
        """Species totals per snapshot [M][K][S_max]; with counts=True also
        each model's counts [K][n_m][S_m]."""
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        if seeds.size != len(self.systems):
            raise SimulationError("one seed per model")
        totals = np.zeros((len(self.systems), self.K, self.Smax), np.uint32)
        cnt = None
        cptr = None
        if counts:
            cnt = [np.zeros((self.K, d.desc.n, d.desc.n_species), np.uint32) for d in self.dhs]
            cptr = (C.POINTER(C.c_uint32) * len(cnt))(*[_p(c, C.c_uint32) for c in cnt])
        _abi.check(self.lib, self.lib.ssasdpd_ensemble_run(self.h, _p(seeds, C.c_uint64), _p(totals, C.c_uint32),
                                                           cptr), "ssasdpd_ensemble_run")
        return (totals, cnt) if counts else totals

    def stats(self):
        s = _abi.Stats()
        self.lib.ssasdpd_ensemble_stats(self.h, C.byref(s))
        return s
