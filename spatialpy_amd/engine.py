"""Python handle of the GPU engine (libssasdpd.so through the C ABI)."""
import ctypes as C

import numpy as np

from . import _abi
from .lowering import DescHandle, output_schedule


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def make_rng(seed=None, words=None, expo=None):
    """Random stream descriptor: seeded (GPU generator) or replay."""
    r = _abi.Rng()
    if words is None:
        r.kind = 0
        r.seed = int(seed if seed is not None else 0) & 0xFFFFFFFFFFFFFFFF
    else:
        words = np.ascontiguousarray(words, dtype=np.uint64)
        expo = np.ascontiguousarray(expo, dtype=np.float64)
        r.kind = 1
        r.words = _p(words, C.c_uint64)
        r.expo = _p(expo, C.c_double)
        r.n_words = words.size
    return r, (words, expo)


def mt19937_64_stream(seed, n):
    """std::mt19937_64(seed) words and their exponential variates (host libm)."""
    lib = _abi.load_library()
    w = np.zeros(n, np.uint64)
    e = np.zeros(n, np.float64)
    lib.ssasdpd_mt19937_64_stream(int(seed) & 0xFFFFFFFFFFFFFFFF, n, _p(w, C.c_uint64), _p(e, C.c_double))
    return w, e


class Engine:
    """One model on the GPU; mirrors the engine life cycle of the C ABI."""

    def __init__(self, sa, trace_capacity=0):
        self.lib = _abi.load_library()
        self.dh = DescHandle(sa)
        self.sa = sa
        h = C.c_void_p()
        _abi.check(self.lib, self.lib.ssasdpd_create(self.dh.byref(), C.byref(h)), "ssasdpd_create")
        self.h = h
        if trace_capacity:
            self.set_trace(trace_capacity)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ssasdpd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def S(self):
        return self.dh.desc.n_species

    @property
    def Sc(self):
        return self.dh.desc.n_chem_species

    @property
    def R(self):
        return self.dh.desc.n_reactions

    def reset(self, seed=None, words=None, expo=None):
        r, self._keep = make_rng(seed, words, expo)
        _abi.check(self.lib, self.lib.ssasdpd_reset(self.h, C.byref(r)), "ssasdpd_reset")

    def set_timeout(self, seconds):
        """Wall-clock limit of each later run / run_batch / advance call
        (None or 0: none); past it the call raises SimulationError with code
        SSASDPD_ERR_TIMEOUT at a step boundary."""
        _abi.check(self.lib, self.lib.ssasdpd_set_timeout(self.h, float(seconds or 0.0)), "set_timeout")

    def advance(self, n):
        _abi.check(self.lib, self.lib.ssasdpd_advance(self.h, n), "ssasdpd_advance")

    def set_trace(self, cap):
        _abi.check(self.lib, self.lib.ssasdpd_set_trace_capacity(self.h, cap), "trace")

    def trace(self, cap):
        ev = np.zeros(cap, _abi.EVENT_DTYPE)
        n = C.c_uint64()
        _abi.check(self.lib, self.lib.ssasdpd_get_trace(self.h, _p(ev, _abi.Event), cap, C.byref(n)), "trace")
        return ev[: min(cap, n.value)], n.value

    def stats(self):
        s = _abi.Stats()
        self.lib.ssasdpd_stats_get(self.h, C.byref(s))
        return s

    PHASES = ("bin_tree", "search", "sdpd", "rates_queue", "ssa", "forces")

    def phase_times(self):
        """GPU ms and launch counts per phase since the last reset."""
        ms = np.zeros(8)
        nl = np.zeros(8, np.uint64)
        k = self.lib.ssasdpd_phase_times(self.h, _p(ms, C.c_double), _p(nl, C.c_uint64), 8)
        return {name: (float(ms[i]), int(nl[i])) for i, name in enumerate(self.PHASES[:k])}

    def state(self):
        n = self.sa.n
        S, R, Sc = self.S, self.R, self.Sc
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "vt", "F", "Fbp")}
        out.update({k: np.zeros(n) for k in ("rho", "Frho", "bvf_phi", "nu", "old_rho", "srrate", "sdrate",
                                              "next_time")})
        out["C"] = np.zeros((n, Sc))
        out["Q"] = np.zeros((n, Sc))
        out["xx"] = np.zeros((n, S), np.uint32)
        out["rrate"] = np.zeros((n, R))
        out["Ddiag"] = np.zeros((n, S))
        st = _abi.State()
        for k, a in out.items():
            setattr(st, k, _p(a, C.c_uint32 if a.dtype == np.uint32 else C.c_double))
        _abi.check(self.lib, self.lib.ssasdpd_get_state(self.h, C.byref(st)), "get_state")
        return out

    def neighbors(self):
        n = self.sa.n
        rp = np.zeros(n + 1, np.int64)
        _abi.check(self.lib, self.lib.ssasdpd_get_neighbors(self.h, _p(rp, C.c_int64), None, None, None, None), "nbrs")
        m = int(rp[-1])
        idx = np.zeros(m, np.int32)
        dist, dwdr, dij = np.zeros(m), np.zeros(m), np.zeros(m)
        _abi.check(self.lib, self.lib.ssasdpd_get_neighbors(
            self.h, _p(rp, C.c_int64), _p(idx, C.c_int32), _p(dist, C.c_double), _p(dwdr, C.c_double),
            _p(dij, C.c_double)), "nbrs")
        return rp, idx, dist, dwdr, dij

    def eval_forces(self, fp32=False):
        """The pair-force kernel evaluated once on the current state:
        (F [n,3], Fbp [n,3], Frho [n])."""
        n = self.sa.n
        F = np.zeros((n, 3))
        Fbp = np.zeros((n, 3))
        Frho = np.zeros(n)
        _abi.check(self.lib, self.lib.ssasdpd_eval_forces(self.h, 1 if fp32 else 0, _p(F, C.c_double),
                                                          _p(Fbp, C.c_double), _p(Frho, C.c_double)),
                   "ssasdpd_eval_forces")
        return F, Fbp, Frho

    def run(self, seed=None, words=None, expo=None, fields=("xx", "C", "x", "v", "rho"), partial_ok=False):
        """Full trajectory with the reference's snapshot schedule.  With
        partial_ok, a run stopped by the time limit (set_timeout) returns the
        snapshots taken before it and res["timeout"] = True instead of
        raising."""
        sa = self.sa
        steps = output_schedule(sa.output_steps, sa.nt)
        k = steps.size
        n = sa.n
        res = {}
        o = _abi.Outputs()
        o.n_snapshots = k
        if "xx" in fields:
            res["xx"] = np.zeros((k, n, self.S), np.uint32)
            o.xx = _p(res["xx"], C.c_uint32)
        if "C" in fields:
            res["C"] = np.zeros((k, n, self.Sc))
            o.C = _p(res["C"], C.c_double)
        for f in ("x", "v"):
            if f in fields:
                res[f] = np.zeros((k, n, 3))
                setattr(o, f, _p(res[f], C.c_double))
        for f, attr in (("rho", "rho"), ("bvf_phi", "bvf_phi"), ("nu", "nu")):
            if f in fields:
                res[f] = np.zeros((k, n))
                setattr(o, attr, _p(res[f], C.c_double))
        res["steps"] = np.zeros(k, np.uint32)
        o.steps = _p(res["steps"], C.c_uint32)
        r, keep = make_rng(seed, words, expo)
        rc = self.lib.ssasdpd_run(self.h, C.byref(r), C.byref(o))
        res["timeout"] = partial_ok and rc == _abi.TIMEOUT
        if not res["timeout"]:
            _abi.check(self.lib, rc, "ssasdpd_run")
        return res

    def run_batch(self, seeds, fields=("xx", "C", "x", "v", "rho"), partial_ok=False):
        """Trajectories seeds[0..B) of the model in one batched run (seeded
        GPU streams).  Returns a dict with "xx" of shape [B][K][n][S] and the
        shared particle fields of shape [K][...] (the particle state does not
        depend on the molecule counts)."""
        sa = self.sa
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64) & np.uint64(0xFFFFFFFFFFFFFFFF))
        B = seeds.size
        steps = output_schedule(sa.output_steps, sa.nt)
        k = steps.size
        n = sa.n
        res = {}
        o = _abi.Outputs()
        o.n_snapshots = k
        if "C" in fields:
            res["C"] = np.zeros((k, n, self.Sc))
            o.C = _p(res["C"], C.c_double)
        for f in ("x", "v"):
            if f in fields:
                res[f] = np.zeros((k, n, 3))
                setattr(o, f, _p(res[f], C.c_double))
        for f in ("rho", "bvf_phi", "nu"):
            if f in fields:
                res[f] = np.zeros((k, n))
                setattr(o, f, _p(res[f], C.c_double))
        res["steps"] = np.zeros(k, np.uint32)
        o.steps = _p(res["steps"], C.c_uint32)
        counts = np.zeros((B, k, n, self.S), np.uint32)
        rc = self.lib.ssasdpd_run_batch(self.h, B, _p(seeds, C.c_uint64), C.byref(o), _p(counts, C.c_uint32))
        res["timeout"] = partial_ok and rc == _abi.TIMEOUT
        if not res["timeout"]:
            _abi.check(self.lib, rc, "ssasdpd_run_batch")
        res["xx"] = counts
        return res


def fr_search(pts, q, dim, h):
    """Stand-alone GPU fixed-radius search (lists in ANN order)."""
    lib = _abi.load_library()
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    m = q.shape[0]
    rp = np.zeros(m + 1, np.int64)
    _abi.check(lib, lib.ssasdpd_fr_search(_p(pts, C.c_double), pts.shape[0], _p(q, C.c_double), m, dim, h,
                                          _p(rp, C.c_int64), None, None), "fr_search")
    idx = np.zeros(rp[-1], np.int32)
    r2 = np.zeros(rp[-1], np.float64)
    _abi.check(lib, lib.ssasdpd_fr_search(_p(pts, C.c_double), pts.shape[0], _p(q, C.c_double), m, dim, h,
                                          _p(rp, C.c_int64), _p(idx, C.c_int32), _p(r2, C.c_double)), "fr_search")
    return rp, idx, r2


def fr_search_noalloc(pts, q, dim, h):
    """fr_search through ssasdpd_fr_search_noalloc: the same lists, with
    every device buffer in the library's fixed scratch array (no device
    allocation; at most NOALLOC_MAX_POINTS points and NOALLOC_MAX_QUERIES
    queries of at most NOALLOC_MAX_WIDTH neighbours)."""
    lib = _abi.load_library()
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    m = q.shape[0]
    rp = np.zeros(m + 1, np.int64)
    _abi.check(lib, lib.ssasdpd_fr_search_noalloc(_p(pts, C.c_double), pts.shape[0], _p(q, C.c_double), m, dim, h,
                                                  _p(rp, C.c_int64), None, None), "fr_search_noalloc")
    idx = np.zeros(rp[-1], np.int32)
    r2 = np.zeros(rp[-1], np.float64)
    _abi.check(lib, lib.ssasdpd_fr_search_noalloc(_p(pts, C.c_double), pts.shape[0], _p(q, C.c_double), m, dim, h,
                                                  _p(rp, C.c_int64), _p(idx, C.c_int32), _p(r2, C.c_double)),
               "fr_search_noalloc")
    return rp, idx, r2


def vm_eval(code, consts, start, x, vol, t, df=None, x_is_double=False):
    """Evaluate one compiled propensity program (spatialpy_amd.expression)
    on the GPU with the engine's interpreter for m argument sets: x [m][n_x]
    (integral values when x_is_double is False), vol and t [m], df [m][n_df]."""
    lib = _abi.load_library()
    code = np.ascontiguousarray(code, dtype=np.uint32)
    consts = np.ascontiguousarray(consts, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    m = x.shape[0]
    x = x.reshape(m, -1)
    vol = np.ascontiguousarray(np.broadcast_to(np.asarray(vol, dtype=np.float64), (m,)))
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (m,)))
    n_df = 0
    if df is not None:
        df = np.ascontiguousarray(df, dtype=np.float64).reshape(m, -1)
        n_df = df.shape[1]
    out = np.zeros(m)
    _abi.check(lib, lib.ssasdpd_vm_eval(_p(code, C.c_uint32), code.size, _p(consts, C.c_double) if consts.size else None,
                                        consts.size, int(start), int(bool(x_is_double)), x.shape[1], n_df, m,
                                        _p(x, C.c_double), _p(vol, C.c_double), _p(t, C.c_double),
                                        _p(df, C.c_double) if n_df else None, _p(out, C.c_double)), "vm_eval")
    return out
