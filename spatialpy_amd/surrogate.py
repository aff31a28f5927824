"""The learned surrogate for the SDPD accelerations (no counterpart in the
reference): a message-passing network trained and evaluated on the GPU by
this package's own kernels, attached to an Engine with periodic checks
against the plain SDPD and fallback."""
import ctypes as C

import numpy as np

from . import _abi
from .engine import Engine


def _p(a, t):
    return _abi.ptr(a, t)


class Surrogate:
    def __init__(self, seed=0):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_create(int(seed), C.byref(h)), "ssasdpd_mpn_create")
        self.h = h
        self.P = int(self.lib.ssasdpd_mpn_num_params())
        self.n_samples = 0
        self.scales = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.ssasdpd_mpn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters ----------------------------------------------------------
    def params(self):
        w = np.zeros(self.P, np.float32)
        _abi.check(self.lib, self.lib.ssasdpd_mpn_get_params(self.h, _p(w, C.c_float)), "get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.P
        _abi.check(self.lib, self.lib.ssasdpd_mpn_set_params(self.h, _p(w, C.c_float)), "set_params")

    def set_scales(self, rho0, v_scale, rho_scale, nu_scale, acc_scale):
        self.scales = np.array([rho0, v_scale, rho_scale, nu_scale, acc_scale], np.float32)
        _abi.check(self.lib, self.lib.ssasdpd_mpn_set_scales(self.h, _p(self.scales, C.c_float)), "set_scales")

    # data ------------------------------------------------------------------
    def add_sample(self, x, X0, deg, nidx, inv_h, y):
        x = np.ascontiguousarray(x, np.float32)
        X0 = np.ascontiguousarray(X0, np.float32)
        deg = np.ascontiguousarray(deg, np.int32)
        nidx = np.ascontiguousarray(nidx, np.int32)
        y = np.ascontiguousarray(y, np.float32)
        n, W = X0.shape[0], nidx.shape[0]
        idx = C.c_int32()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_add_sample(self.h, n, W, float(inv_h), _p(x, C.c_float),
                                                             _p(X0, C.c_float), _p(deg, C.c_int32),
                                                             _p(nidx, C.c_int32), _p(y, C.c_float), C.byref(idx)),
                   "add_sample")
        self.n_samples += 1
        return idx.value

    def add_engine_sample(self, engine):
        idx = C.c_int32()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_add_engine_sample(self.h, engine.h, C.byref(idx)),
                   "add_engine_sample")
        self.n_samples += 1
        return idx.value

    def record(self, engine, every):
        """Record a training graph every `every` plain SDPD steps of engine."""
        _abi.check(self.lib, self.lib.ssasdpd_mpn_record(self.h, engine.h, int(every)), "mpn_record")
        engine._surrogate_ref = self     # the engine holds the network's address

    def stop_recording(self, engine):
        self.lib.ssasdpd_mpn_record(None, engine.h, 1)
        self.n_samples = int(self.lib.ssasdpd_mpn_num_samples(self.h))

    def num_samples(self):
        return int(self.lib.ssasdpd_mpn_num_samples(self.h))

    def clear_samples(self):
        self.lib.ssasdpd_mpn_clear_samples(self.h)
        self.n_samples = 0

    # compute ---------------------------------------------------------------
    def forward(self, sample, n):
        out = np.zeros((n, 3), np.float32)
        _abi.check(self.lib, self.lib.ssasdpd_mpn_forward(self.h, int(sample), _p(out, C.c_float)), "mpn_forward")
        return out

    def time_forward(self, sample, reps=10):
        ms = C.c_double()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_time_forward(self.h, int(sample), int(reps), C.byref(ms)),
                   "mpn_time_forward")
        return ms.value

    def grad(self, sample):
        g = np.zeros(self.P, np.float32)
        loss = C.c_double()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_grad(self.h, int(sample), _p(g, C.c_float), C.byref(loss)),
                   "mpn_grad")
        return loss.value, g

    def train(self, epochs, lr=1e-3, seed=0):
        loss = C.c_double()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_train(self.h, int(epochs), float(lr), int(seed), C.byref(loss)),
                   "mpn_train")
        return loss.value

    # engine ----------------------------------------------------------------
    def attach(self, engine, check_every=50, max_rel_err=0.02):
        _abi.check(self.lib, self.lib.ssasdpd_set_surrogate(engine.h, self.h, int(check_every), float(max_rel_err)),
                   "set_surrogate")
        engine._surrogate_ref = self     # the engine holds the network's address

    @staticmethod
    def detach(engine):
        engine.lib.ssasdpd_set_surrogate(engine.h, None, 1, 0.0)

    @staticmethod
    def stats(engine):
        a, b = C.c_uint64(), C.c_uint64()
        e = C.c_double()
        act = C.c_int32()
        engine.lib.ssasdpd_surrogate_stats(engine.h, C.byref(a), C.byref(b), C.byref(e), C.byref(act))
        return {"nn_steps": a.value, "checks": b.value, "last_err": e.value, "active": bool(act.value)}


def fit_scales(engine, sa):
    """Feature and target scales from an engine's current state: RMS speed,
    RMS density deviation, mean nu and RMS pair acceleration of the moving
    particles."""
    st = engine.state()
    mv = sa.fixed == 0
    F, _, _ = engine.eval_forces()
    g = np.asarray(sa.gravity if getattr(sa, "gravity", None) is not None else (0.0, 0.0, 0.0))
    acc = F[mv] - g
    v = st["v"][mv]
    drho = st["rho"][mv] / sa.rho0 - 1.0
    return (float(sa.rho0), float(max(np.sqrt((v * v).sum(1).mean()), 1e-6)),
            float(max(np.sqrt((drho * drho).mean()), 1e-6)), float(max(np.abs(sa.nu).mean(), 1e-12)),
            float(max(np.sqrt((acc * acc).sum(1).mean()), 1e-12)))


def relative_error(engine, sa, surrogate, sample_index):
    """||F_nn - F|| / ||F - g|| over moving particles for a stored sample
    built from the engine's current state."""
    F, _, _ = engine.eval_forces()
    g = np.asarray(sa.gravity if getattr(sa, "gravity", None) is not None else (0.0, 0.0, 0.0))
    mv = sa.fixed == 0
    out = surrogate.forward(sample_index, sa.n) * surrogate.scales[4]
    ref = F[mv] - g
    return float(np.linalg.norm(out[mv] - ref) / max(np.linalg.norm(ref), 1e-300))
