"""The learned surrogate for the SDPD pair accelerations (no counterpart in
the reference): an edge network over the neighbour lists, trained and
evaluated on the GPU by this package's own kernels, attached to an Engine
with periodic checks against the plain SDPD and fallback.

Per listed pair (i, j) a 3-layer MLP of 16 rotation-invariant features
gives the coefficients of dx/h, dv/v_scale, v_i/v_scale and v_j/v_scale
(the four vectors the SDPD pair term is built from), times the envelope
(1-R)^2 R / (R + 0.001); a particle's predicted pair acceleration is the
sum over its neighbours, in units of acc_scale."""
import ctypes as C

import numpy as np

from . import _abi

N_SCALES = 6


def _p(a, t):
    return _abi.ptr(a, t)


def graph(x, v, vt, rho, mass, nu, solid, rp, idx, dist, h, dim=3):
    """A particle graph for the network from per-particle arrays and CSR
    neighbour lists (rp [n+1], idx, dist in list order), as ELL arrays.  The
    arrays are kept alive by the returned dict."""
    n = len(rho)
    deg = np.diff(np.asarray(rp)).astype(np.int32)
    W = max(1, int(deg.max()) if n else 1)
    nidx = np.zeros((W, n), np.int32)
    nr = np.zeros((W, n))
    for k in range(W):
        sel = np.nonzero(deg > k)[0]
        nidx[k, sel] = np.asarray(idx)[np.asarray(rp)[sel] + k]
        nr[k, sel] = np.asarray(dist)[np.asarray(rp)[sel] + k]
    f64 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape))
    return {"n": n, "W": W, "dim": int(dim), "h": float(h), "x": f64(x, (n, 3)), "v": f64(v, (n, 3)),
            "vt": f64(vt, (n, 3)), "rho": f64(rho, (n,)), "mass": f64(mass, (n,)), "nu": f64(nu, (n,)),
            "solid": np.ascontiguousarray(np.asarray(solid, np.int32)), "deg": deg,
            "nidx": np.ascontiguousarray(nidx), "nr": np.ascontiguousarray(nr)}


def engine_graph(engine, sa):
    """The network's input for an engine's current state and lists."""
    st = engine.state()
    rp, idx, dist, _, _ = engine.neighbors()
    return graph(st["x"], st["v"], st["vt"], st["rho"], sa.mass, st["nu"], sa.fixed, rp, idx, dist, sa.h,
                 dim=sa.dimension)


def _cgraph(g):
    keep = {k: g[k] for k in ("x", "v", "vt", "rho", "mass", "nu", "solid", "deg", "nidx", "nr")}
    for k in ("x", "v", "vt", "rho", "mass", "nu", "nr"):
        keep[k] = np.ascontiguousarray(keep[k], np.float64)
    for k in ("solid", "deg", "nidx"):
        keep[k] = np.ascontiguousarray(keep[k], np.int32)
    cg = _abi.MpnGraph(int(g["n"]), int(g["W"]), int(g.get("dim", 3)), float(g["h"]),
                       *[_p(keep[k], C.c_double) for k in ("x", "v", "vt", "rho", "mass", "nu")],
                       *[_p(keep[k], C.c_int32) for k in ("solid", "deg", "nidx")], _p(keep["nr"], C.c_double))
    return cg, keep


class Surrogate:
    def __init__(self, seed=0):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_create(int(seed), C.byref(h)), "ssasdpd_mpn_create")
        self.h = h
        self.P = int(self.lib.ssasdpd_mpn_num_params())
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

    def set_scales(self, rho0, v_scale, rho_scale, nu_scale, acc_scale, tv_scale):
        self.scales = np.array([rho0, v_scale, rho_scale, nu_scale, acc_scale, tv_scale], np.float32)
        _abi.check(self.lib, self.lib.ssasdpd_mpn_set_scales(self.h, _p(self.scales, C.c_float)), "set_scales")

    # data ------------------------------------------------------------------
    def add_sample(self, g, acc, sel=None):
        """A training sample from a host graph (see graph()) with target pair
        accelerations acc [n][3]; sel: particle indices (None: every moving
        particle)."""
        cg, keep = _cgraph(g)
        acc = np.ascontiguousarray(acc, np.float64)
        assert acc.shape == (g["n"], 3)
        idx = C.c_int32()
        if sel is None:
            _abi.check(self.lib, self.lib.ssasdpd_mpn_add_sample(self.h, C.byref(cg), _p(acc, C.c_double), 0, None,
                                                                 C.byref(idx)), "add_sample")
        else:
            s = np.ascontiguousarray(sel, np.int32)
            _abi.check(self.lib, self.lib.ssasdpd_mpn_add_sample(self.h, C.byref(cg), _p(acc, C.c_double), s.size,
                                                                 _p(s, C.c_int32), C.byref(idx)), "add_sample")
        return idx.value

    def add_engine_sample(self, engine, max_particles=32768, seed=0):
        idx = C.c_int32()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_add_engine_sample(self.h, engine.h, int(max_particles), int(seed),
                                                                    C.byref(idx)), "add_engine_sample")
        return idx.value

    def record(self, engine, every, max_particles=32768):
        """Record a sample every `every` plain SDPD steps of engine."""
        _abi.check(self.lib, self.lib.ssasdpd_mpn_record(self.h, engine.h, int(every), int(max_particles)),
                   "mpn_record")
        engine._surrogate_ref = self     # the engine holds the network's address

    def stop_recording(self, engine):
        self.lib.ssasdpd_mpn_record(None, engine.h, 1, 1)

    def num_samples(self):
        return int(self.lib.ssasdpd_mpn_num_samples(self.h))

    def sample_size(self, k):
        a, b = C.c_int32(), C.c_int32()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_sample_size(self.h, int(k), C.byref(a), C.byref(b)), "sample_size")
        return a.value, b.value

    def clear_samples(self):
        self.lib.ssasdpd_mpn_clear_samples(self.h)

    # compute ---------------------------------------------------------------
    def forward(self, k):
        """Training-path (pred, target) [n_particles][3] of sample k, pair
        accelerations."""
        npart, _ = self.sample_size(k)
        out = np.zeros((npart, 3), np.float32)
        tgt = np.zeros((npart, 3), np.float32)
        _abi.check(self.lib, self.lib.ssasdpd_mpn_forward(self.h, int(k), _p(out, C.c_float), _p(tgt, C.c_float)),
                   "mpn_forward")
        return out, tgt

    def relative_error(self, samples=None):
        """sqrt(sum |pred - target|^2 / sum |target|^2) over the samples (all
        by default), training path."""
        ks = range(self.num_samples()) if samples is None else samples
        num = den = 0.0
        for k in ks:
            p, y = self.forward(k)
            num += float(((p.astype(np.float64) - y) ** 2).sum())
            den += float((y.astype(np.float64) ** 2).sum())
        return float(np.sqrt(num / max(den, 1e-300)))

    def predict(self, g):
        """The fused inference kernel on a host graph: pair accelerations
        [n][3] (0 for solid particles)."""
        cg, keep = _cgraph(g)
        out = np.zeros((g["n"], 3))
        _abi.check(self.lib, self.lib.ssasdpd_mpn_predict(self.h, C.byref(cg), _p(out, C.c_double)), "mpn_predict")
        return out

    def time_engine(self, engine, reps=10):
        ms = C.c_double()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_time_engine(self.h, engine.h, int(reps), C.byref(ms)),
                   "mpn_time_engine")
        return ms.value

    def grad(self, k):
        g = np.zeros(self.P, np.float32)
        loss = C.c_double()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_grad(self.h, int(k), _p(g, C.c_float), C.byref(loss)), "mpn_grad")
        return loss.value, g

    def train(self, epochs, lr=1e-3, seed=0):
        loss = C.c_double()
        _abi.check(self.lib, self.lib.ssasdpd_mpn_train(self.h, int(epochs), float(lr), int(seed), C.byref(loss)),
                   "mpn_train")
        return loss.value

    def fit(self, epochs, lr=2e-3, lr_end=2e-5, stages=6, seed=0):
        """Adam over the stored samples with the learning rate stepped down
        geometrically from lr to lr_end over `stages` equal stages; returns
        the last stage's mean loss."""
        per = max(1, epochs // stages)
        loss = None
        for s in range(stages):
            r = lr * (lr_end / lr) ** (s / max(stages - 1, 1))
            loss = self.train(per, lr=r, seed=seed + s)
        return loss

    # engine ----------------------------------------------------------------
    def attach(self, engine, check_every=50, max_rel_err=0.02, refresh_iters=0, refresh_lr=1e-4, keep_samples=32,
               max_particles=32768, rearm=False, cap_factor=20.0):
        """Use the network on engine's steps between checks every
        check_every steps (fallback above max_rel_err).  refresh_iters > 0:
        each check adds that plain step as a sample (keeping the newest
        keep_samples) and takes refresh_iters Adam steps.  rearm: a failed
        check pauses the network only until a later check passes.
        cap_factor: between checks a particle's predicted pair acceleration
        is capped at this multiple of the last check's RMS plain pair
        acceleration (<= 0: no cap)."""
        _abi.check(self.lib, self.lib.ssasdpd_set_surrogate(engine.h, self.h, int(check_every), float(max_rel_err)),
                   "set_surrogate")
        _abi.check(self.lib, self.lib.ssasdpd_set_surrogate_rearm(engine.h, 1 if rearm else 0), "set_surrogate_rearm")
        _abi.check(self.lib, self.lib.ssasdpd_set_surrogate_cap(engine.h, float(cap_factor)), "set_surrogate_cap")
        _abi.check(self.lib, self.lib.ssasdpd_set_surrogate_refresh(engine.h, int(refresh_iters), float(refresh_lr),
                                                                    int(keep_samples), int(max_particles)),
                   "set_surrogate_refresh")
        engine._surrogate_ref = self     # the engine holds the network's address

    @staticmethod
    def detach(engine):
        engine.lib.ssasdpd_set_surrogate(engine.h, None, 1, 0.0)

    @staticmethod
    def stats(engine):
        a, b, r, f = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        e, m = C.c_double(), C.c_double()
        act = C.c_int32()
        engine.lib.ssasdpd_surrogate_stats(engine.h, C.byref(a), C.byref(b), C.byref(e), C.byref(m), C.byref(r),
                                           C.byref(f), C.byref(act))
        nc, cap = C.c_uint64(), C.c_double()
        _abi.check(engine.lib, engine.lib.ssasdpd_surrogate_capped(engine.h, C.byref(nc), C.byref(cap)),
                   "surrogate_capped")
        # capped: network predictions scaled down to the cap (a multiple of
        # the last check's RMS plain pair acceleration) since the attach
        return {"nn_steps": a.value, "checks": b.value, "last_err": e.value, "max_err": m.value,
                "refreshes": r.value, "failed_checks": f.value, "active": bool(act.value),
                "capped": nc.value, "cap": cap.value}


def fit_scales(engine, sa):
    """Feature and target scales from an engine's current state: rho0, RMS
    speed, RMS density deviation, mean nu, RMS pair acceleration and RMS
    |v - vt| of the moving particles."""
    st = engine.state()
    mv = sa.fixed == 0
    F, _, _ = engine.eval_forces()
    g = np.asarray(sa.gravity if getattr(sa, "gravity", None) is not None else (0.0, 0.0, 0.0))
    acc = F[mv] - g
    v = st["v"][mv]
    dvt = (st["v"] - st["vt"])[mv]
    drho = st["rho"][mv] / sa.rho0 - 1.0
    return (float(sa.rho0), float(max(np.sqrt((v * v).sum(1).mean()), 1e-6)),
            float(max(np.sqrt((drho * drho).mean()), 1e-6)), float(max(np.abs(sa.nu).mean(), 1e-12)),
            float(max(np.sqrt((acc * acc).sum(1).mean()), 1e-12)),
            float(max(np.sqrt((dvt * dvt).sum(1).mean()), 1e-9)))


def pair_accelerations(engine, sa):
    """F - g of the plain pair kernel on the engine's current state [n][3]
    (0 for fixed particles)."""
    F, _, _ = engine.eval_forces()
    g = np.asarray(sa.gravity if getattr(sa, "gravity", None) is not None else (0.0, 0.0, 0.0))
    acc = F - g
    acc[sa.fixed != 0] = 0.0
    acc[:, sa.dimension:] = 0.0
    return acc


def relative_error(engine, sa, surrogate):
    """||a_nn - a|| / ||a|| over moving particles on the engine's current
    state, a the plain pair kernel's F - g, a_nn the fused inference
    kernel's prediction."""
    ref = pair_accelerations(engine, sa)
    out = surrogate.predict(engine_graph(engine, sa))
    mv = sa.fixed == 0
    return float(np.linalg.norm(out[mv] - ref[mv]) / max(np.linalg.norm(ref[mv]), 1e-300))
