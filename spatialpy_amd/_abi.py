"""ctypes mirror of include/ssasdpd.h and the loader for libssasdpd.so.

The structures here must match the C header field for field; the loader
fails loudly when the HIP library is missing (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

ABI_VERSION = 3

OK = 0
TIMEOUT = 8
CAPACITY = 10
# limits of ssasdpd_fr_search_noalloc (include/ssasdpd.h)
NOALLOC_MAX_POINTS = 32768
NOALLOC_MAX_QUERIES = 32768
NOALLOC_MAX_WIDTH = 32
STATUS_NAMES = {
    0: "OK", 1: "BAD_DESC", 2: "NAN_STATE", 3: "NEGATIVE_STATE",
    4: "PROPENSITY_OVERFLOW", 5: "DIFFUSION_OVERFLOW", 6: "NAN_DIJ", 7: "HIP",
    8: "TIMEOUT", 9: "REPLAY_EXHAUSTED", 10: "CAPACITY", 11: "UNSUPPORTED",
}

FLAG_TEXTBOOK_SELECTION = 1 << 0
FLAG_ANN_ORDER_MOVING = 1 << 1
FLAG_NO_RDME = 1 << 2
FLAG_TRACE = 1 << 3
FLAG_FP32_FORCES = 1 << 4
FLAG_NO_RENUMBER = 1 << 5

BC_CHEM, BC_V, BC_NU, BC_RHO = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)


class Reaction(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("pad", C.c_int32),
                ("k", C.c_double), ("type_mask", C.c_uint64)]


class BC(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("type", C.c_int32),
                ("target", C.c_int32), ("species", C.c_int32), ("pad", C.c_int32),
                ("value", C.c_double * 3)]


class Desc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("n", C.c_int32), ("dimension", C.c_int32),
        ("n_types", C.c_int32), ("n_species", C.c_int32), ("n_reactions", C.c_int32),
        ("n_chem_species", C.c_int32), ("n_chem_reactions", C.c_int32),
        ("static_domain", C.c_int32), ("n_bcs", C.c_int32), ("n_output_steps", C.c_int32),
        ("flags", C.c_uint32), ("nt", C.c_uint32), ("pad0", C.c_uint32),
        ("dt", C.c_double), ("h", C.c_double), ("rho0", C.c_double), ("c0", C.c_double),
        ("P0", C.c_double), ("bbox", C.c_double * 6), ("gravity", C.c_double * 3),
        ("x", _dp), ("v", _dp), ("type", _ip), ("nu", _dp), ("mass", _dp), ("rho", _dp),
        ("fixed", _ip), ("u0", _up), ("C0", _dp),
        ("diffusion_matrix", _dp), ("reactions", C.POINTER(Reaction)),
        ("jcN", _ip), ("irN", _ip), ("prN", _ip), ("jcG", _ip), ("irG", _ip), ("N_dense", _ip),
        ("bcs", C.POINTER(BC)), ("output_steps", _up),
        ("n_code", C.c_int32), ("n_consts", C.c_int32), ("code", _up), ("consts", _dp),
        ("n_data_fn", C.c_int32), ("pad1", C.c_int32), ("data_fn", _dp),
    ]


class Outputs(C.Structure):
    _fields_ = [("n_snapshots", C.c_int32), ("pad", C.c_int32), ("xx", _up), ("C", _dp),
                ("x", _dp), ("v", _dp), ("rho", _dp), ("bvf_phi", _dp), ("nu", _dp),
                ("steps", _up)]


class Rng(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("seed", C.c_uint64),
                ("words", _u64p), ("expo", _dp), ("n_words", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("particle_steps", C.c_uint64),
                ("reactions", C.c_uint64), ("diffusions", C.c_uint64),
                ("rng_words_used", C.c_uint64),
                ("ms_neighbors", C.c_double), ("ms_sdpd", C.c_double),
                ("ms_rates", C.c_double), ("ms_ssa", C.c_double), ("ms_total", C.c_double),
                ("bad_particle", C.c_int32), ("max_degree", C.c_int32)]


class State(C.Structure):
    _fields_ = [("x", _dp), ("v", _dp), ("vt", _dp), ("F", _dp), ("Fbp", _dp),
                ("rho", _dp), ("Frho", _dp), ("bvf_phi", _dp), ("nu", _dp), ("old_rho", _dp),
                ("C", _dp), ("Q", _dp), ("xx", _up), ("srrate", _dp), ("sdrate", _dp),
                ("rrate", _dp), ("Ddiag", _dp), ("next_time", _dp)]


class MpnGraph(C.Structure):
    _fields_ = [("n", C.c_int32), ("W", C.c_int32), ("dim", C.c_int32), ("h", C.c_double),
                ("x", _dp), ("v", _dp), ("vt", _dp), ("rho", _dp), ("mass", _dp), ("nu", _dp),
                ("solid", _ip), ("deg", _ip), ("nidx", _ip), ("nr", _dp)]


class Event(C.Structure):
    _fields_ = [("t", C.c_double), ("voxel", C.c_int32), ("kind", C.c_int32),
                ("which", C.c_int32), ("dest", C.c_int32)]


EVENT_DTYPE = np.dtype([("t", "<f8"), ("voxel", "<i4"), ("kind", "<i4"),
                        ("which", "<i4"), ("dest", "<i4")])


def ptr(a, ctype):
    """Pointer to a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays handed to the C ABI must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


_LIB = None
LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libssasdpd.so")


def _declare(lib):
    P = C.c_void_p
    lib.ssasdpd_create.argtypes = [C.POINTER(Desc), C.POINTER(P)]
    lib.ssasdpd_destroy.argtypes = [P]
    lib.ssasdpd_destroy.restype = None
    lib.ssasdpd_last_error.restype = C.c_char_p
    lib.ssasdpd_abi_version.restype = C.c_int
    lib.ssasdpd_num_snapshots.argtypes = [C.POINTER(Desc), _ip]
    lib.ssasdpd_reset.argtypes = [P, C.POINTER(Rng)]
    lib.ssasdpd_advance.argtypes = [P, C.c_uint32]
    lib.ssasdpd_run.argtypes = [P, C.POINTER(Rng), C.POINTER(Outputs)]
    lib.ssasdpd_run_batch.argtypes = [P, C.c_int32, _u64p, C.POINTER(Outputs), _up]
    lib.ssasdpd_stats_get.argtypes = [P, C.POINTER(Stats)]
    lib.ssasdpd_get_state.argtypes = [P, C.POINTER(State)]
    lib.ssasdpd_get_neighbors.argtypes = [P, _i64p, _ip, _dp, _dp, _dp]
    lib.ssasdpd_fr_search.argtypes = [_dp, C.c_int32, _dp, C.c_int32, C.c_int32, C.c_double,
                                      _i64p, _ip, _dp]
    lib.ssasdpd_set_trace_capacity.argtypes = [P, C.c_uint64]
    lib.ssasdpd_fr_search_noalloc.argtypes = [_dp, C.c_int32, _dp, C.c_int32, C.c_int32, C.c_double,
                                              _i64p, _ip, _dp]
    lib.ssasdpd_set_timeout.argtypes = [P, C.c_double]
    lib.ssasdpd_get_trace.argtypes = [P, C.POINTER(Event), C.c_uint64, _u64p]
    lib.ssasdpd_mt19937_64_stream.argtypes = [C.c_uint64, C.c_uint64, _u64p, _dp]
    lib.ssasdpd_phase_times.argtypes = [P, _dp, _u64p, C.c_int32]
    lib.ssasdpd_eval_forces.argtypes = [P, C.c_int32, _dp, _dp, _dp]
    lib.ssasdpd_vm_eval.argtypes = [_up, C.c_int32, _dp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_int32, _dp, _dp, _dp, _dp, _dp]
    lib.ssasdpd_ensemble_create.argtypes = [C.POINTER(C.POINTER(Desc)), C.c_int32, C.POINTER(P)]
    lib.ssasdpd_ensemble_destroy.argtypes = [P]
    lib.ssasdpd_ensemble_destroy.restype = None
    lib.ssasdpd_ensemble_num_snapshots.argtypes = [P, _ip]
    lib.ssasdpd_ensemble_set_tree_width.argtypes = [P, C.c_int32]
    lib.ssasdpd_ensemble_autotune.argtypes = [P, _u64p, C.c_int32, _ip, _dp]
    lib.ssasdpd_ensemble_run.argtypes = [P, _u64p, _up, C.POINTER(_up)]
    lib.ssasdpd_ensemble_stats.argtypes = [P, C.POINTER(Stats)]
    _fp = C.POINTER(C.c_float)
    lib.ssasdpd_mpn_num_params.restype = C.c_int32
    lib.ssasdpd_mpn_create.argtypes = [C.c_uint64, C.POINTER(P)]
    lib.ssasdpd_mpn_destroy.argtypes = [P]
    lib.ssasdpd_mpn_destroy.restype = None
    lib.ssasdpd_mpn_get_params.argtypes = [P, _fp]
    lib.ssasdpd_mpn_set_params.argtypes = [P, _fp]
    lib.ssasdpd_mpn_set_scales.argtypes = [P, _fp]
    lib.ssasdpd_mpn_add_sample.argtypes = [P, C.POINTER(MpnGraph), _dp, C.c_int32, _ip, _ip]
    lib.ssasdpd_mpn_add_engine_sample.argtypes = [P, P, C.c_int32, C.c_uint64, _ip]
    lib.ssasdpd_mpn_clear_samples.argtypes = [P]
    lib.ssasdpd_mpn_sample_size.argtypes = [P, C.c_int32, _ip, _ip]
    lib.ssasdpd_mpn_forward.argtypes = [P, C.c_int32, _fp, _fp]
    lib.ssasdpd_mpn_predict.argtypes = [P, C.POINTER(MpnGraph), _dp]
    lib.ssasdpd_mpn_time_engine.argtypes = [P, P, C.c_int32, _dp]
    lib.ssasdpd_mpn_grad.argtypes = [P, C.c_int32, _fp, _dp]
    lib.ssasdpd_mpn_train.argtypes = [P, C.c_int32, C.c_float, C.c_uint64, _dp]
    lib.ssasdpd_set_surrogate.argtypes = [P, P, C.c_int32, C.c_double]
    lib.ssasdpd_mpn_record.argtypes = [P, P, C.c_int32, C.c_int32]
    lib.ssasdpd_mpn_num_samples.argtypes = [P]
    lib.ssasdpd_mpn_num_samples.restype = C.c_int32
    lib.ssasdpd_surrogate_stats.argtypes = [P, _u64p, _u64p, _dp, _dp, _u64p, _u64p, _ip]
    lib.ssasdpd_surrogate_capped.argtypes = [P, _u64p, _dp]
    lib.ssasdpd_set_surrogate_cap.argtypes = [P, C.c_double]
    lib.ssasdpd_grid_shape.argtypes = [_dp, C.c_int32, C.c_int32, C.c_double, _ip, C.POINTER(C.c_int64), _dp]
    lib.ssasdpd_set_surrogate_rearm.argtypes = [P, C.c_int32]
    lib.ssasdpd_set_surrogate_refresh.argtypes = [P, C.c_int32, C.c_float, C.c_int32, C.c_int32]
    for name in ["ssasdpd_create", "ssasdpd_num_snapshots", "ssasdpd_reset", "ssasdpd_advance",
                 "ssasdpd_run", "ssasdpd_run_batch", "ssasdpd_stats_get", "ssasdpd_get_state",
                 "ssasdpd_get_neighbors", "ssasdpd_fr_search", "ssasdpd_set_trace_capacity",
                 "ssasdpd_get_trace", "ssasdpd_mt19937_64_stream", "ssasdpd_phase_times",
                 "ssasdpd_eval_forces", "ssasdpd_ensemble_create", "ssasdpd_ensemble_num_snapshots",
                 "ssasdpd_ensemble_set_tree_width", "ssasdpd_ensemble_autotune", "ssasdpd_ensemble_run",
                 "ssasdpd_ensemble_stats", "ssasdpd_mpn_create", "ssasdpd_mpn_get_params",
                 "ssasdpd_mpn_set_params", "ssasdpd_mpn_set_scales", "ssasdpd_mpn_add_sample",
                 "ssasdpd_mpn_add_engine_sample", "ssasdpd_mpn_clear_samples", "ssasdpd_mpn_forward",
                 "ssasdpd_mpn_grad", "ssasdpd_mpn_train", "ssasdpd_set_surrogate", "ssasdpd_surrogate_stats",
                 "ssasdpd_surrogate_capped", "ssasdpd_set_surrogate_cap", "ssasdpd_grid_shape",
                 "ssasdpd_mpn_record", "ssasdpd_mpn_time_engine", "ssasdpd_mpn_sample_size",
                 "ssasdpd_mpn_predict", "ssasdpd_set_surrogate_refresh", "ssasdpd_set_surrogate_rearm", "ssasdpd_vm_eval", "ssasdpd_set_timeout", "ssasdpd_fr_search_noalloc"]:
        getattr(lib, name).restype = C.c_int
    return lib


def load_library(path=None):
    """Load libssasdpd.so (built by __graft_entry__.build()).  Raises when it
    is missing: the engine has no CPU fallback."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP engine library not built: {p} (run __graft_entry__.build())")
    lib = _declare(C.CDLL(p))
    if lib.ssasdpd_abi_version() != ABI_VERSION:
        raise RuntimeError("libssasdpd.so ABI version mismatch")
    if path is None:
        _LIB = lib
    return lib


def check(lib, rc, what=""):
    if rc != OK:
        msg = lib.ssasdpd_last_error()
        msg = msg.decode() if msg else ""
        from .core.errors import SimulationError
        raise SimulationError(f"{what}: {STATUS_NAMES.get(rc, rc)}: {msg}", code=rc)
