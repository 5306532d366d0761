"""ctypes declarations of the C ABI in include/zeus_c.h (libzeus_amd.so).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libzeus_amd.so")

OK, E_ALLOC, E_LAUNCH, E_ARG, E_UNSUPPORTED, E_NODEVICE, E_IO, E_SCRATCH = \
    0, 3, 4, -10, -11, -12, -13, -14
UNSET, SURRENDERED, CONVERGED, STOPPED, MALLOC, KERNEL, NONFINITE = -1, 0, 1, 2, 3, 4, 5
OBJ_ROSENBROCK, OBJ_RASTRIGIN, OBJ_ACKLEY, OBJ_GAUSSIAN, OBJ_LOGLIK, OBJ_MLP, OBJ_SPHERE = (
    1, 2, 3, 4, 5, 6, 7)
GRAD_ADJOINT, GRAD_DUAL = 0, 1
FORM_AUTO, FORM_THREAD, FORM_GROUP, FORM_TILE, FORM_TEAM = 0, 1, 2, 3, 4
FLAG_METRICS = 1
PRNG_XORWOW, PRNG_PHILOX, PRNG_SOBOL = 0, 1, 2

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


class Trajectory(C.Structure):
    _fields_ = [("coords", C.c_void_p), ("fval", C.c_void_p), ("grad", C.c_void_p),
                ("status", C.c_void_p), ("alpha", C.c_void_p), ("steps", C.c_int32),
                ("starts", C.c_int32)]


class Swarm(C.Structure):
    _fields_ = [("x", C.c_void_p), ("v", C.c_void_p), ("pbest_x", C.c_void_p),
                ("pbest_val", C.c_void_p), ("gbest_x", C.c_void_p), ("gbest_val", C.c_void_p),
                ("rng", C.c_void_p)]


class Problem(C.Structure):
    _fields_ = [("objective", C.c_int32), ("dim", C.c_int32), ("lower", C.c_double),
                ("upper", C.c_double), ("n_starts", C.c_int64), ("max_iter", C.c_int32),
                ("pso_iter", C.c_int32), ("required", C.c_int32), ("seed", C.c_uint32),
                ("stream", C.c_uint32), ("prng", C.c_int32), ("tol", C.c_double),
                ("params", C.c_void_p), ("n_params", C.c_int64), ("x0", C.c_void_p),
                ("name", C.c_char_p), ("trajectory", C.POINTER(Trajectory))]


class Options(C.Structure):
    _fields_ = [("gradient", C.c_int32), ("form", C.c_int32), ("group_threads", C.c_int32),
                ("nzerosteps", C.c_int32), ("use_graph", C.c_int32), ("flags", C.c_int32)]


class Metrics(C.Structure):
    _fields_ = [("ms_per_call", C.c_double), ("calls_per_thread_mean", C.c_double),
                ("fraction_of_kernel", C.c_double), ("block95", C.c_double),
                ("serialized", C.c_double), ("fraction_of_thread", C.c_double)]


class Result(C.Structure):
    _fields_ = [("idx", C.c_int32), ("status", C.c_int32), ("iter", C.c_int32),
                ("nonfinite", C.c_int32), ("fval", C.c_double), ("gradient_norm", C.c_double),
                ("claimed", C.c_int32), ("actual", C.c_int32), ("surrendered", C.c_int32),
                ("stopped", C.c_int32), ("ms_pso", C.c_double), ("ms_opt", C.c_double),
                ("ms_total", C.c_double), ("sum_iter", C.c_int64),
                ("sum_ls_evals", C.c_int64), ("sum_grads", C.c_int64), ("ad", Metrics),
                ("bfgs", Metrics)]


class StartOutputs(C.Structure):
    _fields_ = [("status", _i32p), ("iter", _i32p), ("fval", _dp), ("gnorm", _dp),
                ("x", _dp), ("ls_evals", _i32p), ("grads", _i32p)]


# exported symbol -> (restype, argtypes); mirrors include/zeus_c.h
SIGNATURES = {
    "zeus_run_batch": (C.c_int32, [C.POINTER(Problem), C.c_int32, C.POINTER(Options),
                                   C.POINTER(Result), _dp, C.POINTER(StartOutputs),
                                   C.c_void_p]),
    "zeus_metrics_from_ticks": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                            C.c_int32, C.c_double, C.c_double,
                                            C.POINTER(Metrics)]),
    "zeus_euclidean_error": (C.c_double, [C.c_char_p, _dp, C.c_int32]),
    "zeus_append_results_tsv": (C.c_int32, [C.c_char_p, C.c_int32, C.c_int64, C.c_char_p,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                            C.POINTER(Result), _dp]),
    "zeus_pso_init": (C.c_int32, [C.POINTER(Problem), C.POINTER(Swarm), C.c_void_p]),
    "zeus_pso_step": (C.c_int32, [C.POINTER(Problem), C.c_int32, C.POINTER(Swarm), C.c_void_p]),
    "zeus_bfgs_update": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                     C.c_void_p]),
    "zeus_stream_fill": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, C.c_uint32, C.c_uint32,
                                     C.c_uint64, C.c_uint32, C.c_void_p]),
    "zeus_write_trajectory": (C.c_int32, [C.c_char_p, _dp, _dp, _dp, C.c_void_p, _dp,
                                          C.c_int32, C.c_int32, C.c_int32]),
    "zeus_pso_run": (C.c_int32, [C.POINTER(Problem), C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p]),
    "zeus_pso_run_noalloc": (C.c_int32, [C.POINTER(Problem), C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p]),
    "zeus_eval": (C.c_int32, [C.POINTER(Problem), C.c_void_p, C.c_int64, C.c_void_p,
                              C.c_void_p, C.c_int32, C.c_void_p]),
    "zeus_argmin": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_void_p]),
    "zeus_philox_uniform": (C.c_double, [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64]),
    "zeus_philox_fill": (C.c_int32, [C.c_void_p, C.c_int64, C.c_uint32, C.c_uint32,
                                     C.c_uint64, C.c_uint64, C.c_void_p]),
    "zeus_alloc": (C.c_int32, [C.POINTER(C.c_void_p), C.c_uint64]),
    "zeus_free": (C.c_int32, [C.c_void_p]),
    "zeus_memcpy_h2d": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "zeus_memcpy_d2h": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "zeus_memcpy_d2d": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "zeus_synchronize": (C.c_int32, []),
    "zeus_status_string": (C.c_char_p, [C.c_int32]),
    "zeus_set_fault_hooks": (C.c_int32, [C.c_int32, C.c_int32]),
    "zeus_tuned_config": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, _i32p, _i32p]),
    "zeus_make_dijet": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_int32, _dp]),
    "zeus_make_mlp_data": (C.c_int32, [C.c_uint32, C.c_int32, _dp]),
    "zeus_make_spd": (C.c_int32, [C.c_uint32, C.c_uint32, C.c_int32, _dp]),
}

_lib = None


class ZeusError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: status {code} ({status_string(code)})")
        self.code = code


def lib():
    """Load libzeus_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m zeus_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(code: int) -> str:
    if _lib is None and not os.path.exists(LIB_PATH):
        return "?"
    return lib().zeus_status_string(code).decode()


def check(code: int, what: str) -> None:
    if code != OK:
        raise ZeusError(code, what)
