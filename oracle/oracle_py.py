"""ctypes wrapper of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the spatialpy_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        from spatialpy_amd import _abi
        L = C.CDLL(path)
        P = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(_abi.Desc), C.c_int]
        L.oracle_create.restype = P
        L.oracle_destroy.argtypes = [P]
        L.oracle_reset.argtypes = [P, C.POINTER(_abi.Rng)]
        L.oracle_advance.argtypes = [P, C.c_uint32]
        L.oracle_run.argtypes = [P, C.POINTER(_abi.Rng), C.POINTER(_abi.Outputs)]
        L.oracle_stats.argtypes = [P, C.POINTER(_abi.Stats)]
        L.oracle_get_state.argtypes = [P, C.POINTER(_abi.State)]
        L.oracle_get_neighbors.argtypes = [P] + [C.POINTER(C.c_int64), C.POINTER(C.c_int32)] + [C.POINTER(C.c_double)] * 3
        L.oracle_eval_forces.argtypes = [P] + [C.POINTER(C.c_double)] * 3
        L.oracle_set_trace_capacity.argtypes = [P, C.c_uint64]
        L.oracle_get_trace.argtypes = [P, C.POINTER(_abi.Event), C.c_uint64, C.POINTER(C.c_uint64)]
        sig = [C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_double,
               C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.oracle_fr_search.argtypes = sig
        L.oracle_fr_brute.argtypes = sig
        L.oracle_tree_stats.argtypes = [C.POINTER(C.c_double), C.c_int32, C.c_int32] + [C.POINTER(C.c_int32)] * 4
        L.oracle_mt19937_64_stream.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
        L.oracle_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _fr(fn, pts, q, dim, h):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    m = q.shape[0]
    rp = np.zeros(m + 1, np.int64)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    fn(dp(pts), pts.shape[0], dp(q), m, dim, h, rp.ctypes.data_as(C.POINTER(C.c_int64)), None, None)
    idx = np.zeros(rp[-1], np.int32)
    r2 = np.zeros(rp[-1], np.float64)
    fn(dp(pts), pts.shape[0], dp(q), m, dim, h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
       idx.ctypes.data_as(C.POINTER(C.c_int32)), dp(r2))
    return rp, idx, r2


def fr_search(pts, q, dim, h):
    """kd-tree restatement: lists in ANN order."""
    return _fr(lib().oracle_fr_search, pts, q, dim, h)


def fr_brute(pts, q, dim, h):
    """Brute force: acceptance rule only, ordered by (r2, index)."""
    return _fr(lib().oracle_fr_brute, pts, q, dim, h)


def tree_stats(pts, dim):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = [C.c_int32() for _ in range(4)]
    lib().oracle_tree_stats(pts.ctypes.data_as(C.POINTER(C.c_double)), pts.shape[0], dim, *[C.byref(o) for o in out])
    return dict(n_split=out[0].value, n_leaf=out[1].value, n_empty=out[2].value, depth=out[3].value)


def mt_stream(seed, n):
    w = np.zeros(n, np.uint64)
    e = np.zeros(n, np.float64)
    lib().oracle_mt19937_64_stream(seed, n, w.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   e.ctypes.data_as(C.POINTER(C.c_double)))
    return w, e


class Oracle:
    """One oracle simulation (same life cycle as the engine handle)."""

    def __init__(self, sa, gauss_seidel=True):
        from spatialpy_amd.lowering import DescHandle
        self.dh = DescHandle(sa)
        self.sa = sa
        self.h = lib().oracle_create(self.dh.byref(), int(gauss_seidel))
        if not self.h:
            raise RuntimeError(lib().oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    @staticmethod
    def _rng(seed=None, words=None, expo=None):
        from spatialpy_amd import _abi
        r = _abi.Rng()
        if words is None:
            r.kind = 0
            r.seed = int(seed or 0) & 0xFFFFFFFFFFFFFFFF
        else:
            r.kind = 1
            r.words = words.ctypes.data_as(C.POINTER(C.c_uint64))
            r.expo = expo.ctypes.data_as(C.POINTER(C.c_double))
            r.n_words = words.size
        return r

    def reset(self, seed=None, words=None, expo=None):
        self._keep = (words, expo)
        lib().oracle_reset(self.h, C.byref(self._rng(seed, words, expo)))

    def advance(self, n):
        rc = lib().oracle_advance(self.h, n)
        if rc:
            raise RuntimeError(f"oracle advance rc={rc}: {lib().oracle_last_error().decode()}")

    def set_trace(self, cap):
        lib().oracle_set_trace_capacity(self.h, cap)

    def trace(self, cap):
        from spatialpy_amd import _abi
        ev = np.zeros(cap, _abi.EVENT_DTYPE)
        n = C.c_uint64()
        lib().oracle_get_trace(self.h, ev.ctypes.data_as(C.POINTER(_abi.Event)), cap, C.byref(n))
        return ev[: min(cap, n.value)], n.value

    def stats(self):
        from spatialpy_amd import _abi
        s = _abi.Stats()
        lib().oracle_stats(self.h, C.byref(s))
        return s

    def state(self):
        from spatialpy_amd import _abi
        sa = self.sa
        n = sa.n
        S = self.dh.desc.n_species
        R = self.dh.desc.n_reactions
        Sc = self.dh.desc.n_chem_species
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "vt", "F", "Fbp")}
        out.update({k: np.zeros(n) for k in ("rho", "Frho", "bvf_phi", "nu", "old_rho", "srrate", "sdrate", "next_time")})
        out["C"] = np.zeros((n, Sc)); out["Q"] = np.zeros((n, Sc))
        out["xx"] = np.zeros((n, S), np.uint32); out["rrate"] = np.zeros((n, R)); out["Ddiag"] = np.zeros((n, S))
        st = _abi.State()
        for k, a in out.items():
            setattr(st, k, a.ctypes.data_as(C.POINTER(C.c_uint32 if a.dtype == np.uint32 else C.c_double)))
        lib().oracle_get_state(self.h, C.byref(st))
        return out

    def neighbors(self):
        n = self.sa.n
        rp = np.zeros(n + 1, np.int64)
        L = lib()
        L.oracle_get_neighbors(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)), None, None, None, None)
        m = int(rp[-1])
        idx = np.zeros(m, np.int32)
        dist, dwdr, dij = np.zeros(m), np.zeros(m), np.zeros(m)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        L.oracle_get_neighbors(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                               idx.ctypes.data_as(C.POINTER(C.c_int32)), dp(dist), dp(dwdr), dp(dij))
        return rp, idx, dist, dwdr, dij

    def eval_forces(self):
        """pairwiseForce once on the current state: (F [n,3], Fbp [n,3], Frho [n])."""
        n = self.sa.n
        F, Fbp, Frho = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        rc = lib().oracle_eval_forces(self.h, dp(F), dp(Fbp), dp(Frho))
        assert rc == 0
        return F, Fbp, Frho

    def run(self, seed=None, words=None, expo=None):
        """Full trajectory with the reference's snapshot schedule."""
        from spatialpy_amd import _abi
        from spatialpy_amd.lowering import output_schedule
        sa = self.sa
        steps = output_schedule(sa.output_steps, sa.nt)
        k = steps.size
        n = sa.n
        S = self.dh.desc.n_species
        Sc = self.dh.desc.n_chem_species
        res = dict(xx=np.zeros((k, n, S), np.uint32), C=np.zeros((k, n, Sc)), x=np.zeros((k, n, 3)),
                   v=np.zeros((k, n, 3)), rho=np.zeros((k, n)), bvf_phi=np.zeros((k, n)), nu=np.zeros((k, n)),
                   steps=np.zeros(k, np.uint32))
        o = _abi.Outputs()
        o.n_snapshots = k
        o.xx = res["xx"].ctypes.data_as(C.POINTER(C.c_uint32))
        o.C = res["C"].ctypes.data_as(C.POINTER(C.c_double))
        o.x = res["x"].ctypes.data_as(C.POINTER(C.c_double))
        o.v = res["v"].ctypes.data_as(C.POINTER(C.c_double))
        o.rho = res["rho"].ctypes.data_as(C.POINTER(C.c_double))
        o.bvf_phi = res["bvf_phi"].ctypes.data_as(C.POINTER(C.c_double))
        o.nu = res["nu"].ctypes.data_as(C.POINTER(C.c_double))
        o.steps = res["steps"].ctypes.data_as(C.POINTER(C.c_uint32))
        self._keep = (words, expo)
        rc = lib().oracle_run(self.h, C.byref(self._rng(seed, words, expo)), C.byref(o))
        if rc:
            raise RuntimeError(f"oracle run rc={rc}: {lib().oracle_last_error().decode()}")
        return res
