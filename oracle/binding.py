"""ctypes binding of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, never by the product package.  The oracle is a
C restatement of the reference method; see zeus_oracle.h for what pins each
function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

ROSENBROCK, RASTRIGIN, ACKLEY, GAUSSIAN, LOGLIK, MLP, SPHERE = 1, 2, 3, 4, 5, 6, 7
GRAD_ADJOINT, GRAD_DUAL = 0, 1
PRNG_XORWOW, PRNG_PHILOX, PRNG_SOBOL = 0, 1, 2
UPDATE_RANK2, UPDATE_LITERAL = 0, 1

D_ADD, D_SUB, D_MUL, D_DIV, D_NEG, D_ABS, D_SIN, D_COS, D_EXP, D_SQRT, \
    D_ATAN2, D_LOG, D_POW, D_POWD, D_LGAMMA, D_TANH = range(16)

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Dual(C.Structure):
    _fields_ = [("re", C.c_double), ("du", C.c_double)]


class Objective(C.Structure):
    _fields_ = [("kind", C.c_int), ("dim", C.c_int), ("params", _dp),
                ("n_params", C.c_int64)]


class Problem(C.Structure):
    _fields_ = [("objective", C.c_int), ("dim", C.c_int), ("lower", C.c_double),
                ("upper", C.c_double), ("n_starts", C.c_int64),
                ("required", C.c_int), ("max_iter", C.c_int),
                ("pso_iter", C.c_int), ("seed", C.c_uint32),
                ("stream", C.c_uint32), ("tol", C.c_double), ("params", _dp),
                ("n_params", C.c_int64), ("x0", _dp), ("prng", C.c_int)]


class Traj(C.Structure):
    _fields_ = [("coords", _dp), ("fval", _dp), ("grad", _dp),
                ("status", C.POINTER(C.c_int8)), ("alpha", _dp), ("steps", C.c_int32),
                ("starts", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("grad_mode", C.c_int), ("update_mode", C.c_int),
                ("n_threads", C.c_int), ("early_stop", C.c_int),
                ("traj", C.POINTER(Traj))]


class Best(C.Structure):
    _fields_ = [("idx", C.c_int32), ("status", C.c_int32), ("iter", C.c_int32),
                ("fval", C.c_double), ("gnorm", C.c_double),
                ("claimed", C.c_int32), ("stopped", C.c_int32),
                ("surrendered", C.c_int32)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_uniform.restype = C.c_double
        L.orc_uniform.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64]
        L.orc_digamma.restype = C.c_double
        L.orc_digamma.argtypes = [C.c_double]
        L.orc_dual_apply.restype = Dual
        L.orc_dual_apply.argtypes = [C.c_int, Dual, Dual, C.c_double]
        L.orc_value.restype = C.c_double
        L.orc_value.argtypes = [C.POINTER(Objective), _dp]
        L.orc_value_dual.restype = Dual
        L.orc_value_dual.argtypes = [C.POINTER(Objective), C.POINTER(Dual)]
        L.orc_grad_dual.argtypes = [C.POINTER(Objective), _dp, _dp]
        L.orc_value_grad_adjoint.restype = C.c_double
        L.orc_value_grad_adjoint.argtypes = [C.POINTER(Objective), _dp, _dp]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [_dp, _dp, C.c_int]
        L.orc_norm.restype = C.c_double
        L.orc_norm.argtypes = [_dp, C.c_int]
        L.orc_search_direction.argtypes = [_dp, _dp, _dp, C.c_int]
        L.orc_update_literal.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int]
        L.orc_update_rank2.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int]
        L.orc_line_search.restype = C.c_double
        L.orc_line_search.argtypes = [C.POINTER(Objective), C.c_double, _dp, _dp,
                                      _dp, C.c_int, C.POINTER(C.c_int), _dp]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.POINTER(Problem), C.POINTER(Options), _i32p, _i32p,
                              _dp, _dp, _dp, _i64p, _i64p, _dp, C.POINTER(Best)]
        L.orc_pso.restype = C.c_int
        L.orc_pso.argtypes = [C.POINTER(Problem), _dp, _dp, _dp, _dp]
        L.orc_pso_swarm.restype = C.c_int
        L.orc_pso_swarm.argtypes = [C.POINTER(Problem), _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_pso_step.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_xorwow_seed.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                      C.POINTER(C.c_uint32)]
        L.orc_xorwow_next.restype = C.c_uint32
        L.orc_xorwow_next.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_sobol_uniform.restype = C.c_double
        L.orc_sobol_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_stream_uniforms.restype = C.c_int
        L.orc_stream_uniforms.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint64,
                                          C.c_uint32, C.c_int64, _dp]
        L.orc_argmin.restype = C.c_int64
        L.orc_argmin.argtypes = [_dp, C.c_int64]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def uniform(seed, stream, start, draw):
    return lib().orc_uniform(seed, stream, start, draw)


def digamma(x):
    return lib().orc_digamma(float(x))


def dual_op(op, a, b=(0.0, 0.0), e=0.0):
    r = lib().orc_dual_apply(op, Dual(*a), Dual(*b), float(e))
    return (r.re, r.du)


class Obj:
    """An objective descriptor; keeps its parameter array alive."""

    def __init__(self, kind, dim, params=None):
        self.kind, self.dim = kind, dim
        self.params = None if params is None else _f64(params)
        np_ = 0 if self.params is None else self.params.size
        self.c = Objective(kind, dim, _ptr(self.params), np_)

    def value(self, x):
        x = _f64(x)
        return lib().orc_value(C.byref(self.c), _ptr(x))

    def grad_dual(self, x):
        x = _f64(x)
        g = np.zeros(self.dim)
        lib().orc_grad_dual(C.byref(self.c), _ptr(x), _ptr(g))
        return g

    def value_grad_adjoint(self, x):
        x = _f64(x)
        g = np.zeros(self.dim)
        f = lib().orc_value_grad_adjoint(C.byref(self.c), _ptr(x), _ptr(g))
        return f, g

    def value_dual(self, xd):
        arr = (Dual * self.dim)(*[Dual(a, b) for a, b in xd])
        r = lib().orc_value_dual(C.byref(self.c), arr)
        return (r.re, r.du)


def dot(a, b):
    a, b = _f64(a), _f64(b)
    return lib().orc_dot(_ptr(a), _ptr(b), a.size)


def norm(g):
    g = _f64(g)
    return lib().orc_norm(_ptr(g), g.size)


def search_direction(H, g):
    H, g = _f64(H), _f64(g)
    p = np.zeros(g.size)
    lib().orc_search_direction(_ptr(p), _ptr(H), _ptr(g), g.size)
    return p


def update(H, s, y, sTy, mode=UPDATE_RANK2):
    H = _f64(H).copy()
    s, y = _f64(s), _f64(y)
    fn = lib().orc_update_literal if mode == UPDATE_LITERAL else lib().orc_update_rank2
    fn(_ptr(H), _ptr(s), _ptr(y), float(sTy), s.size)
    return H


def line_search(obj, f0, x, p, g):
    x, p, g = _f64(x), _f64(p), _f64(g)
    ne = C.c_int(0)
    facc = C.c_double(0.0)
    a = lib().orc_line_search(C.byref(obj.c), float(f0), _ptr(x), _ptr(p), _ptr(g),
                              x.size, C.byref(ne), C.byref(facc))
    return a, ne.value, facc.value


def argmin(v):
    v = _f64(v)
    return int(lib().orc_argmin(_ptr(v), v.size))


def pso_step(w, c1, c2, r, X, V, pbx, gbx):
    X, V = _f64(X).copy(), _f64(V).copy()
    n, dim = X.shape
    r, pbx, gbx = _f64(r), _f64(pbx), _f64(gbx)
    lib().orc_pso_step(n, dim, w, c1, c2, _ptr(r), _ptr(X), _ptr(V), _ptr(pbx), _ptr(gbx))
    return X, V


@dataclass
class RunResult:
    status: np.ndarray
    iter: np.ndarray
    fval: np.ndarray
    gnorm: np.ndarray
    x: np.ndarray
    evals: np.ndarray
    grads: np.ndarray
    best: dict
    best_x: np.ndarray
    trajectory: dict | None = None


def _problem(objective, dim, lower, upper, n_starts, required, max_iter, pso_iter,
             seed, stream, tol, params, x0, prng=PRNG_XORWOW):
    return Problem(objective, dim, lower, upper, n_starts, required, max_iter,
                   pso_iter, seed & 0xFFFFFFFF, stream & 0xFFFFFFFF, tol,
                   _ptr(params), 0 if params is None else params.size, _ptr(x0), prng)


def xorwow_seed(seed, stream, start, coord=0):
    st = (C.c_uint32 * 6)()
    lib().orc_xorwow_seed(seed & 0xFFFFFFFF, stream & 0xFFFFFFFF, start, coord, st)
    return list(st)


def xorwow_next(state):
    """One xorwow step on a 6-word state list (updated in place)."""
    st = (C.c_uint32 * 6)(*state)
    v = lib().orc_xorwow_next(st)
    state[:] = list(st)
    return int(v)


def sobol_uniform(start, draw):
    return lib().orc_sobol_uniform(start, draw)


def stream_uniforms(prng, seed, stream, start, n, coord=0):
    out = np.zeros(n)
    rc = lib().orc_stream_uniforms(prng, seed & 0xFFFFFFFF, stream & 0xFFFFFFFF, start, coord,
                                   n, _ptr(out))
    if rc != 0:
        raise ValueError("bad generator arguments")
    return out


def run(objective, dim, lower, upper, n_starts, max_iter, pso_iter, required, tol,
        seed, *, stream=0, params=None, x0=None, grad=GRAD_ADJOINT,
        update_mode=UPDATE_RANK2, threads=0, early_stop=True, trajectory=None,
        trajectory_alpha=True, prng=PRNG_XORWOW) -> RunResult:
    """trajectory=(steps, starts) records the BFGS steps (orc_traj layout)."""
    params = None if params is None else _f64(params)
    x0 = None if x0 is None else _f64(x0).reshape(n_starts, dim)
    P = _problem(objective, dim, lower, upper, n_starts, required, max_iter,
                 pso_iter, seed, stream, tol, params, x0, prng)
    tarr = None
    tptr = None
    if trajectory is not None:
        T_, S_ = int(trajectory[0]), int(trajectory[1])
        tarr = dict(coords=np.zeros((T_, dim, S_)), fval=np.zeros((T_, S_)),
                    grad=np.zeros((T_, S_)), status=np.zeros((T_, S_), np.int8),
                    alpha=np.zeros((T_, S_)))
        tr = Traj(_ptr(tarr["coords"]), _ptr(tarr["fval"]), _ptr(tarr["grad"]),
                  tarr["status"].ctypes.data_as(C.POINTER(C.c_int8)),
                  _ptr(tarr["alpha"]) if trajectory_alpha else None, T_, S_)
        tptr = C.pointer(tr)
    O = Options(grad, update_mode, threads, 1 if early_stop else 0, tptr)
    N = n_starts
    st = np.zeros(N, np.int32)
    it = np.zeros(N, np.int32)
    fv = np.zeros(N)
    gn = np.zeros(N)
    xs = np.zeros((N, dim))
    ev = np.zeros(N, np.int64)
    gr = np.zeros(N, np.int64)
    bx = np.zeros(dim)
    B = Best()
    rc = lib().orc_run(C.byref(P), C.byref(O), st.ctypes.data_as(_i32p),
                       it.ctypes.data_as(_i32p), _ptr(fv), _ptr(gn), _ptr(xs),
                       ev.ctypes.data_as(_i64p), gr.ctypes.data_as(_i64p), _ptr(bx),
                       C.byref(B))
    if rc != 0:
        raise RuntimeError(f"orc_run failed: {rc}")
    best = dict(idx=B.idx, status=B.status, iter=B.iter, fval=B.fval, gnorm=B.gnorm,
                claimed=B.claimed, stopped=B.stopped, surrendered=B.surrendered)
    return RunResult(st, it, fv, gn, xs, ev, gr, best, bx, tarr)


def pso(objective, dim, lower, upper, n_starts, pso_iter, seed, *, stream=0,
        params=None, prng=PRNG_XORWOW):
    params = None if params is None else _f64(params)
    P = _problem(objective, dim, lower, upper, n_starts, 0, 0, pso_iter, seed,
                 stream, 0.0, params, None, prng)
    pbx = np.zeros((n_starts, dim))
    pbv = np.zeros(n_starts)
    gbx = np.zeros(dim)
    gbv = C.c_double(0.0)
    rc = lib().orc_pso(C.byref(P), _ptr(pbx), _ptr(pbv), _ptr(gbx), C.byref(gbv))
    if rc != 0:
        raise RuntimeError(f"orc_pso failed: {rc}")
    return pbx, pbv, gbx, gbv.value


def pso_swarm(objective, dim, lower, upper, n_starts, pso_iter, seed, *, stream=0,
              params=None, prng=PRNG_XORWOW):
    """The swarm after pso_iter iterations: dict of x, v (final positions and
    velocities), pbest_x, pbest_val, gbest_x, gbest_val."""
    params = None if params is None else _f64(params)
    P = _problem(objective, dim, lower, upper, n_starts, 0, 0, pso_iter, seed,
                 stream, 0.0, params, None, prng)
    out = dict(x=np.zeros((n_starts, dim)), v=np.zeros((n_starts, dim)),
               pbest_x=np.zeros((n_starts, dim)), pbest_val=np.zeros(n_starts),
               gbest_x=np.zeros(dim))
    gbv = C.c_double(0.0)
    rc = lib().orc_pso_swarm(C.byref(P), _ptr(out["x"]), _ptr(out["v"]), _ptr(out["pbest_x"]),
                             _ptr(out["pbest_val"]), _ptr(out["gbest_x"]), C.byref(gbv))
    if rc != 0:
        raise RuntimeError(f"orc_pso_swarm failed: {rc}")
    out["gbest_val"] = gbv.value
    return out
