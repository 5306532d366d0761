"""Python mirror of the reference's host API, over the C ABI.

``Zeus(f, lower, upper, N, MAX_ITER, PSO_ITER, requiredConverged, fun_name,
tolerance, seed, run, parallel=True, prng=PRNGType.XORWOW,
trajectory_file="", nzerosteps=0)`` keeps the argument order, meaning and
result fields of ``zeus::Zeus`` (reference include/zeus.cuh:341-375,
include/util.hpp:36-54) for the built-in objectives.  Everything runs in
libzeus_amd.so on the GPU; this module only marshals arguments.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
import warnings
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi
from ._capi import check, lib

__all__ = [
    "PRNGType", "Convergence", "Result", "DeviceArray", "Objective", "Rosenbrock",
    "Rastrigin", "Ackley", "Sphere", "Gaussian", "LogLikelihood", "Mlp", "Zeus",
    "Problem", "run_batch", "pso", "evaluate", "argmin", "philox_uniform", "philox_fill",
    "make_dijet", "make_mlp_data", "make_spd", "tuned_config", "set_fault_hooks",
    "write_trajectory", "euclidean_error", "append_results_tsv", "stream_uniforms",
    "Swarm", "bfgs_update",
]


class PRNGType(enum.IntEnum):
    """Reference traits.hpp:67-71 (values = ZEUS_PRNG_*): XORWOW (Marsaglia's
    xorwow, a state per start), PHILOX (counter-based Philox4x32-10), SOBOL
    (dimension 1 of Sobol's sequence from index start on).  Own seeding and
    conversions: none reproduces cuRAND's streams."""
    XORWOW = 0
    PHILOX = 1
    SOBOL = 2


@dataclass
class Convergence:
    actual: int = 0
    claimed: int = 0
    surrendered: int = 0
    stopped: int = 0


@dataclass
class Result:
    """Fields of zeus::Result<D> plus the call's timing and work counters."""
    idx: int = -1
    status: int = -1
    fval: float = 37.0
    gradientNorm: float = 0.0
    coordinates: np.ndarray = field(default_factory=lambda: np.zeros(0))
    iter: int = 0
    c: Convergence = field(default_factory=Convergence)
    ms_opt: float = 0.0
    ms_pso: float = 0.0
    ms_total: float = 0.0
    nonfinite: int = 0
    sum_iter: int = 0
    sum_ls_evals: int = 0
    sum_grads: int = 0
    per_start: dict | None = None
    trajectory: dict | None = None
    # timing of gradient evaluations (ad) and H updates (bfgs) per start,
    # per-thread form only (reference Metrics; zeros otherwise)
    ad: dict = field(default_factory=dict)
    bfgs: dict = field(default_factory=dict)


class DeviceArray:
    """Owning device buffer (the role of CudaBuffer<T>, cuda_buffer.hpp:14-209)."""

    def __init__(self, host: np.ndarray | None = None, nbytes: int = 0, dtype=np.float64):
        self.ptr = C.c_void_p(0)
        self.dtype = np.dtype(dtype)
        if host is not None:
            host = np.ascontiguousarray(host, dtype=self.dtype)
            nbytes = host.nbytes
            self.shape = host.shape
        else:
            self.shape = (nbytes // self.dtype.itemsize,)
        self.nbytes = nbytes
        check(lib().zeus_alloc(C.byref(self.ptr), nbytes), "zeus_alloc")
        if host is not None and nbytes:
            check(lib().zeus_memcpy_h2d(self.ptr, host.ctypes.data, nbytes), "zeus_memcpy_h2d")

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().zeus_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes),
                  "zeus_memcpy_d2h")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value and capi._lib is not None:
            capi._lib.zeus_free(self.ptr)
            self.ptr = C.c_void_p(0)


# --- objectives -------------------------------------------------------------
class Objective:
    kind = 0
    name = ""

    def __init__(self, dim: int, params: np.ndarray | None = None):
        self.dim = int(dim)
        self.params = None if params is None else np.ascontiguousarray(params, np.float64)
        self._dev = None

    @property
    def arity(self) -> int:
        return self.dim

    def device_params(self) -> tuple[int, int]:
        if self.params is None:
            return 0, 0
        if self._dev is None:
            self._dev = DeviceArray(self.params)
        return self._dev.addr, int(self.params.size)


class Rosenbrock(Objective):
    """examples/rosenbrock.hpp:32-41."""
    kind, name = capi.OBJ_ROSENBROCK, "rosenbrock"


class Rastrigin(Objective):
    """examples/rastrigin.hpp:72-81."""
    kind, name = capi.OBJ_RASTRIGIN, "rastrigin"


class Ackley(Objective):
    """examples/ackley.hpp:124-133."""
    kind, name = capi.OBJ_ACKLEY, "ackley"


class Sphere(Objective):
    """sum x_i^2 (the reference tests' Quad / GoodObjective)."""
    kind, name = capi.OBJ_SPHERE, "sphere"


class Gaussian(Objective):
    """0.5 x^T C x (examples/gaussian.hpp:11-34); C uploaded once."""
    kind, name = capi.OBJ_GAUSSIAN, "gaussian"

    def __init__(self, C_host):
        C_host = np.ascontiguousarray(C_host, np.float64)
        if C_host.ndim != 2 or C_host.shape[0] != C_host.shape[1] or C_host.shape[0] == 0:
            raise ValueError("Gaussian needs a non-empty square matrix")
        super().__init__(C_host.shape[0], C_host.reshape(-1))


_libm = C.CDLL("libm.so.6")
_libm.lgamma.restype = C.c_double
_libm.lgamma.argtypes = [C.c_double]


class LogLikelihood(Objective):
    """Poisson NLL of binned counts (simulation/fitting/LogLikelihood.h:11-82).

    params = [x (n) | k (n) | lgamma(k + 1) (n)]; the lgamma constants are
    taken from the C library, as the reference's device lgamma would give.
    """
    kind, name = capi.OBJ_LOGLIK, "poisson"

    def __init__(self, bin_counts, bin_centers):
        k = np.asarray(bin_counts, np.float64)
        x = np.asarray(bin_centers, np.float64)
        if k.shape != x.shape or k.ndim != 1 or k.size == 0:
            raise ValueError("counts and centers must be equal-length 1-D arrays")
        lgk = np.array([_libm.lgamma(float(v) + 1.0) for v in k])
        super().__init__(4, np.concatenate([x, k, lgk]))

    @classmethod
    def from_params(cls, params):
        obj = cls.__new__(cls)
        Objective.__init__(obj, 4, params)
        return obj


class Mlp(Objective):
    """1-16-16-1 tanh network, MSE over (x, y) points; 321 parameters."""
    kind, name = capi.OBJ_MLP, "mlp"

    def __init__(self, x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        if x.shape != y.shape or x.ndim != 1 or x.size == 0:
            raise ValueError("x and y must be equal-length 1-D arrays")
        super().__init__(321, np.concatenate([x, y]))


# --- problems and batches -----------------------------------------------------
@dataclass
class Problem:
    f: Objective
    lower: float
    upper: float
    N: int
    MAX_ITER: int
    PSO_ITER: int
    requiredConverged: int
    tolerance: float
    seed: int
    fun_name: str | None = None
    stream: int = 0
    x0: np.ndarray | None = None
    # (steps, starts): record the first `steps` iterations of the first
    # `starts` starts (zeus_trajectory); alpha: also the line-search steps
    trajectory: tuple | None = None
    trajectory_alpha: bool = True
    prng: int = PRNGType.XORWOW


class _TrajBuffers:
    """Device buffers of one problem's trajectory and their C descriptor."""

    def __init__(self, steps: int, starts: int, dim: int, alpha: bool):
        if steps < 0 or starts < 0:
            raise ValueError("trajectory steps and starts must be >= 0")
        self.steps, self.starts, self.dim = int(steps), int(starts), int(dim)
        n = self.steps * self.starts
        self.coords = DeviceArray(nbytes=8 * n * dim)
        self.fval = DeviceArray(nbytes=8 * n)
        self.grad = DeviceArray(nbytes=8 * n)
        self.status = DeviceArray(nbytes=n, dtype=np.int8)
        self.alpha = DeviceArray(nbytes=8 * n) if alpha else None
        self.c = capi.Trajectory(self.coords.addr or None, self.fval.addr or None,
                                 self.grad.addr or None, self.status.addr or None,
                                 (self.alpha.addr or None) if alpha else None,
                                 self.steps, self.starts)

    def to_host(self) -> dict:
        T, S, D = self.steps, self.starts, self.dim
        return dict(coords=self.coords.to_host().reshape(T, D, S),
                    fval=self.fval.to_host().reshape(T, S),
                    grad=self.grad.to_host().reshape(T, S),
                    status=self.status.to_host().reshape(T, S),
                    alpha=(self.alpha.to_host().reshape(T, S) if self.alpha is not None
                           else np.zeros((T, S))))


def _options(gradient="adjoint", form=None, parallel=True, group_threads=0, nzerosteps=0,
             use_graph=False, metrics=False):
    g = {"adjoint": capi.GRAD_ADJOINT, "dual": capi.GRAD_DUAL}[gradient]
    if form is None:
        form = capi.FORM_GROUP if parallel else capi.FORM_THREAD
    elif isinstance(form, str):
        form = {"auto": capi.FORM_AUTO, "thread": capi.FORM_THREAD,
                "group": capi.FORM_GROUP, "tile": capi.FORM_TILE,
                "team": capi.FORM_TEAM}[form]
    return capi.Options(g, int(form), int(group_threads), int(nzerosteps),
                        1 if use_graph else 0, capi.FLAG_METRICS if metrics else 0)


def _metrics_dict(m) -> dict:
    return {k: getattr(m, k) for k, _ in capi.Metrics._fields_}


# what the library returns for a call without ZEUS_FLAG_METRICS
_METRICS_ZERO = _metrics_dict(capi.Metrics())


def run_batch(problems: list[Problem], *, gradient="adjoint", form="auto", group_threads=0,
              per_start=False, use_graph=False, metrics=False) -> list[Result]:
    """Many independent problems in one launch sequence (zeus_run_batch)."""
    n = len(problems)
    keep = []
    trajs = {}
    P = (capi.Problem * max(n, 1))()
    for i, pr in enumerate(problems):
        addr, nparam = pr.f.device_params()
        x0addr = 0
        if pr.x0 is not None:
            x0 = np.ascontiguousarray(pr.x0, np.float64).reshape(pr.N, pr.f.dim)
            dx = DeviceArray(x0)
            keep.append(dx)
            x0addr = dx.addr
        name = (pr.fun_name if pr.fun_name is not None else pr.f.name).encode()
        keep.append(name)
        tptr = None
        if pr.trajectory is not None:
            tb = _TrajBuffers(pr.trajectory[0], pr.trajectory[1], pr.f.dim,
                              pr.trajectory_alpha)
            trajs[i] = tb
            tptr = C.pointer(tb.c)
            keep.append(tptr)
        P[i] = capi.Problem(pr.f.kind, pr.f.dim, float(pr.lower), float(pr.upper), int(pr.N),
                            int(pr.MAX_ITER), int(pr.PSO_ITER), int(pr.requiredConverged),
                            int(pr.seed) & 0xFFFFFFFF, int(pr.stream) & 0xFFFFFFFF,
                            int(pr.prng),
                            float(pr.tolerance), addr or None, nparam, x0addr or None, name,
                            tptr)
    opt = _options(gradient, form, True, group_threads, 0, use_graph, metrics)
    R = (capi.Result * max(n, 1))()
    dims = sum(pr.f.dim for pr in problems)
    bx = (C.c_double * max(dims, 1))()
    so_ptr = None
    arrays = None
    if per_start:
        tot = sum(pr.N for pr in problems)
        totx = sum(pr.N * pr.f.dim for pr in problems)
        arrays = dict(status=np.zeros(tot, np.int32), iter=np.zeros(tot, np.int32),
                      fval=np.zeros(tot), gnorm=np.zeros(tot), x=np.zeros(max(totx, 1)),
                      ls_evals=np.zeros(tot, np.int32), grads=np.zeros(tot, np.int32))
        so = capi.StartOutputs(*[arrays[k].ctypes.data_as(t) for k, t in (
            ("status", capi._i32p), ("iter", capi._i32p), ("fval", capi._dp),
            ("gnorm", capi._dp), ("x", capi._dp), ("ls_evals", capi._i32p),
            ("grads", capi._i32p))])
        so_ptr = C.byref(so)
    rc = lib().zeus_run_batch(P, n, C.byref(opt), R, bx, so_ptr, None)
    if rc < 0:
        raise capi.ZeusError(rc, "zeus_run_batch")
    out = []
    ox = os_ = oxx = 0
    for i, pr in enumerate(problems):
        r = R[i]
        res = Result(idx=r.idx, status=r.status, fval=r.fval, gradientNorm=r.gradient_norm,
                     coordinates=np.frombuffer(bx, np.float64, pr.f.dim, 8 * ox).copy(),
                     iter=r.iter,
                     c=Convergence(r.actual, r.claimed, r.surrendered, r.stopped),
                     ms_opt=r.ms_opt, ms_pso=r.ms_pso, ms_total=r.ms_total,
                     nonfinite=r.nonfinite, sum_iter=r.sum_iter, sum_ls_evals=r.sum_ls_evals,
                     sum_grads=r.sum_grads,
                     ad=_metrics_dict(r.ad) if metrics else dict(_METRICS_ZERO),
                     bfgs=_metrics_dict(r.bfgs) if metrics else dict(_METRICS_ZERO))
        if rc in (capi.E_ALLOC, capi.E_LAUNCH):
            res.status = capi.MALLOC if rc == capi.E_ALLOC else capi.KERNEL
        if i in trajs and rc == capi.OK:
            res.trajectory = trajs[i].to_host()
        if arrays is not None:
            res.per_start = {k: (v[os_:os_ + pr.N] if k != "x" else
                                 v[oxx:oxx + pr.N * pr.f.dim].reshape(pr.N, pr.f.dim))
                             for k, v in arrays.items()}
        ox += pr.f.dim
        os_ += pr.N
        oxx += pr.N * pr.f.dim
        out.append(res)
    return out


def Zeus(f: Objective, lower: float, upper: float, N: int, MAX_ITER: int, PSO_ITER: int,
         requiredConverged: int, fun_name: str, tolerance: float, seed: int, run: int,
         parallel: bool = True, prng: PRNGType = PRNGType.XORWOW, trajectory_file: str = "",
         nzerosteps: int = 0, *, gradient: str = "adjoint", form=None, group_threads: int = 0,
         x0=None, stream: int = 0, per_start: bool = False, use_graph: bool = False,
         trajectory_steps: int | None = None, trajectory_starts: int | None = None,
         results_tsv: bool | str = False) -> Result:
    """Multistart PSO -> BFGS; returns the best start's Result.

    Argument order and meaning follow zeus::Zeus (include/zeus.cuh:341-358).
    ``parallel`` selects the per-workgroup form (True) or the per-thread form
    (False) unless ``form`` overrides it; ``run`` only labels the call;
    ``prng`` selects the generator of the starts' streams (PRNGType).
    ``trajectory_file`` (".tsv") records every start's iterations and writes
    them as reference zeus.cuh:189-239 does, the line-search steps only for
    ``parallel`` (the reference's sequential form writes zeros); the arrays
    are also returned in ``Result.trajectory``.  ``trajectory_steps`` /
    ``trajectory_starts`` bound the recording (default MAX_ITER and N).
    With ``parallel=False`` the per-thread form times its gradients and
    updates (``Result.ad`` / ``Result.bfgs``), as the reference's sequential
    form does.  ``results_tsv`` (True or a path) appends the call's row to
    zeus_<D>d_results.tsv (reference zeus.cuh:172-183).
    """
    if N < 0:
        raise ValueError("N must be >= 0")
    if N == 0:
        return Result(coordinates=np.zeros(f.dim))  # zeus.cuh:58-60
    traj = None
    if trajectory_file:
        steps = MAX_ITER if trajectory_steps is None else min(int(trajectory_steps), MAX_ITER)
        starts = N if trajectory_starts is None else min(int(trajectory_starts), N)
        traj = (steps, starts)
    pr = Problem(f, lower, upper, N, MAX_ITER, PSO_ITER, requiredConverged, tolerance, seed,
                 fun_name, stream, x0, traj, bool(parallel), int(prng))
    if form is None:
        form = "group" if parallel else "thread"
    res = run_batch([pr], gradient=gradient, form=form, group_threads=group_threads,
                    per_start=per_start, use_graph=use_graph, metrics=not parallel)[0]
    if results_tsv:
        append_results_tsv(res, f.dim, N, fun_name, MAX_ITER, PSO_ITER, run,
                           None if results_tsv is True else str(results_tsv))
    if traj is not None and res.trajectory is not None:
        # like the reference (zeus.cuh:229-238), a file that cannot be written
        # is reported and the result still returned
        try:
            write_trajectory(trajectory_file, res.trajectory)
        except capi.ZeusError as e:
            warnings.warn(f"trajectory not written to {trajectory_file}: {e}")
    return res


def stream_uniforms(prng, seed: int, stream: int, start: int, n: int,
                    coord: int = 0) -> np.ndarray:
    """Draws 0..n-1 of one stream, generated on the device (zeus_stream_fill):
    XORWOW, that of (start, coord); PHILOX / SOBOL, that of start."""
    out = DeviceArray(nbytes=8 * n)
    check(lib().zeus_stream_fill(out.ptr, int(n), int(prng), int(seed) & 0xFFFFFFFF,
                                 int(stream) & 0xFFFFFFFF, int(start), int(coord), None),
          "zeus_stream_fill")
    return out.to_host()


def euclidean_error(fun_name: str, x) -> float:
    """Distance to the known minimum implied by fun_name (reference
    util::calculateEuclideanError, src/util.cpp:11-41)."""
    x = np.ascontiguousarray(x, np.float64)
    return lib().zeus_euclidean_error(fun_name.encode(), x.ctypes.data_as(capi._dp), x.size)


def append_results_tsv(res: Result, dim: int, N: int, fun_name: str, MAX_ITER: int,
                       PSO_ITER: int, run: int, path: str | None = None) -> None:
    """Append one call's row to the results table (zeus_append_results_tsv;
    reference util::appendResultsToTsv, include/util.hpp:78-149)."""
    r = capi.Result()
    r.idx, r.status, r.iter = res.idx, res.status, res.iter
    r.fval, r.gradient_norm = res.fval, res.gradientNorm
    r.claimed, r.actual = res.c.claimed, res.c.actual
    r.surrendered, r.stopped = res.c.surrendered, res.c.stopped
    for name in ("ad", "bfgs"):
        m = getattr(res, name) or {}
        setattr(r, name, capi.Metrics(*[float(m.get(k, 0.0)) for k, _ in capi.Metrics._fields_]))
    x = np.ascontiguousarray(res.coordinates, np.float64)
    err = euclidean_error(fun_name, x)
    check(lib().zeus_append_results_tsv(path.encode() if path else None, int(dim), int(N),
                                        fun_name.encode(), 0.0, float(res.ms_pso),
                                        float(res.ms_opt), 0.0, int(MAX_ITER), int(PSO_ITER),
                                        err, int(run), C.byref(r),
                                        x.ctypes.data_as(capi._dp)),
          "zeus_append_results_tsv")


def write_trajectory(filename: str, traj: dict) -> None:
    """Write Result.trajectory arrays (zeus_write_trajectory; reference
    util::writeTrajectoryData, src/utils.cu:126-286)."""
    coords = np.ascontiguousarray(traj["coords"], np.float64)
    steps, dim, starts = coords.shape
    fval = np.ascontiguousarray(traj["fval"], np.float64)
    grad = np.ascontiguousarray(traj["grad"], np.float64)
    status = np.ascontiguousarray(traj["status"], np.int8)
    alpha = np.ascontiguousarray(traj["alpha"], np.float64)
    check(lib().zeus_write_trajectory(filename.encode(), coords.ctypes.data_as(capi._dp),
                                      fval.ctypes.data_as(capi._dp),
                                      grad.ctypes.data_as(capi._dp), status.ctypes.data,
                                      alpha.ctypes.data_as(capi._dp), starts, steps, dim),
          "zeus_write_trajectory")


# --- phases and utilities -------------------------------------------------------
def pso(f: Objective, lower, upper, N, PSO_ITER, seed, stream=0, prng=PRNGType.XORWOW):
    """PSO alone (zeus_pso_run): returns pBestX, pBestVal, gBestX, gBestVal."""
    addr, nparam = f.device_params()
    P = capi.Problem(f.kind, f.dim, float(lower), float(upper), int(N), 0, int(PSO_ITER), 0,
                     int(seed) & 0xFFFFFFFF, int(stream) & 0xFFFFFFFF, int(prng), 0.0,
                     addr or None,
                     nparam, None, None, None)
    px = DeviceArray(nbytes=8 * N * f.dim)
    pv = DeviceArray(nbytes=8 * N)
    gx = DeviceArray(nbytes=8 * f.dim)
    gv = DeviceArray(nbytes=8)
    check(lib().zeus_pso_run(C.byref(P), px.ptr, pv.ptr, gx.ptr, gv.ptr, None), "zeus_pso_run")
    return (px.to_host().reshape(N, f.dim), pv.to_host(), gx.to_host(), float(gv.to_host()[0]))


class Swarm:
    """Device swarm of one problem for the PSO phases (zeus_pso_init /
    zeus_pso_step; reference pso::initKernel / iterKernel)."""

    def __init__(self, f: Objective, lower, upper, N, seed, stream=0,
                 prng=PRNGType.XORWOW):
        self.f, self.N, self.D = f, int(N), f.dim
        addr, nparam = f.device_params()
        self.problem = capi.Problem(f.kind, f.dim, float(lower), float(upper), int(N), 0, 0, 0,
                                    int(seed) & 0xFFFFFFFF, int(stream) & 0xFFFFFFFF, int(prng),
                                    0.0, addr or None, nparam, None, None, None)
        n, d = self.N, self.D
        self.x = DeviceArray(nbytes=8 * n * d)
        self.v = DeviceArray(nbytes=8 * n * d)
        self.pbx = DeviceArray(nbytes=8 * n * d)
        self.pbv = DeviceArray(nbytes=8 * n)
        self.gbx = DeviceArray(nbytes=8 * d)
        self.gbv = DeviceArray(nbytes=8)
        self.rng = DeviceArray(nbytes=4 * 6 * n * d, dtype=np.uint32) \
            if int(prng) == PRNGType.XORWOW else None
        self.c = capi.Swarm(self.x.addr or None, self.v.addr or None, self.pbx.addr or None,
                            self.pbv.addr or None, self.gbx.addr or None,
                            self.gbv.addr or None,
                            (self.rng.addr or None) if self.rng is not None else None)

    def init(self):
        check(lib().zeus_pso_init(C.byref(self.problem), C.byref(self.c), None),
              "zeus_pso_init")
        return self

    def step(self, it: int):
        check(lib().zeus_pso_step(C.byref(self.problem), int(it), C.byref(self.c), None),
              "zeus_pso_step")
        return self

    def state(self):
        """pBestX, pBestVal, gBestX, gBestVal (and X, V) on the host."""
        n, d = self.N, self.D
        return dict(pbest_x=self.pbx.to_host().reshape(n, d), pbest_val=self.pbv.to_host(),
                    gbest_x=self.gbx.to_host(), gbest_val=float(self.gbv.to_host()[0]),
                    x=self.x.to_host().reshape(n, d), v=self.v.to_host().reshape(n, d))


def bfgs_update(H, s, y) -> np.ndarray:
    """Batched BFGS update on the device (zeus_bfgs_update): H (B x n x n),
    s and y (B x n); returns the updated matrices."""
    H = np.ascontiguousarray(H, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    B, n = s.shape
    dH, ds, dy = DeviceArray(H), DeviceArray(s), DeviceArray(y)
    check(lib().zeus_bfgs_update(dH.ptr, ds.ptr, dy.ptr, B, n, None), "zeus_bfgs_update")
    return dH.to_host().reshape(B, n, n)


def evaluate(f: Objective, X, gradient: str | None = None):
    """Values (and gradients) at the rows of X, one lane per point."""
    X = np.ascontiguousarray(X, np.float64).reshape(-1, f.dim)
    n = X.shape[0]
    addr, nparam = f.device_params()
    P = capi.Problem(f.kind, f.dim, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0, 0.0, addr or None, nparam,
                     None, None, None)
    dx = DeviceArray(X)
    df = DeviceArray(nbytes=8 * n)
    dg = DeviceArray(nbytes=8 * n * f.dim) if gradient else None
    g = {"adjoint": capi.GRAD_ADJOINT, "dual": capi.GRAD_DUAL, None: capi.GRAD_ADJOINT}[gradient]
    check(lib().zeus_eval(C.byref(P), dx.ptr, n, df.ptr, dg.ptr if dg else None, g, None),
          "zeus_eval")
    fv = df.to_host()
    if dg is None:
        return fv
    return fv, dg.to_host().reshape(n, f.dim)


def argmin(values) -> int:
    v = DeviceArray(np.asarray(values, np.float64))
    out = C.c_int64(0)
    check(lib().zeus_argmin(v.ptr, v.shape[0], C.byref(out), None), "zeus_argmin")
    return out.value


def philox_uniform(seed, stream, start, draw) -> float:
    return lib().zeus_philox_uniform(seed, stream, start, draw)


def philox_fill(n, seed, stream, start, draw0=0) -> np.ndarray:
    d = DeviceArray(nbytes=8 * n)
    check(lib().zeus_philox_fill(d.ptr, n, seed, stream, start, draw0, None), "zeus_philox_fill")
    return d.to_host()


def make_dijet(seed: int, experiment: int, n_bins: int = 100) -> np.ndarray:
    out = np.zeros(3 * n_bins)
    check(lib().zeus_make_dijet(seed, experiment, n_bins, out.ctypes.data_as(capi._dp)),
          "zeus_make_dijet")
    return out


def make_mlp_data(seed: int, m: int = 256) -> np.ndarray:
    out = np.zeros(2 * m)
    check(lib().zeus_make_mlp_data(seed, m, out.ctypes.data_as(capi._dp)), "zeus_make_mlp_data")
    return out


def make_spd(seed: int, stream: int, dim: int) -> np.ndarray:
    out = np.zeros(dim * dim)
    check(lib().zeus_make_spd(seed, stream, dim, out.ctypes.data_as(capi._dp)), "zeus_make_spd")
    return out.reshape(dim, dim)


def tuned_config(objective: int, dim: int, gradient: int = 0):
    form = C.c_int32(0)
    nt = C.c_int32(0)
    rc = lib().zeus_tuned_config(objective, dim, gradient, C.byref(form), C.byref(nt))
    return None if rc != 0 else (form.value, nt.value)


def set_fault_hooks(fail_alloc_at: int = 0, fail_launches: int = 0) -> None:
    check(lib().zeus_set_fault_hooks(fail_alloc_at, fail_launches), "zeus_set_fault_hooks")


def gaussian_150_matrix(D: int = 150, off: float = 0.5) -> np.ndarray:
    """The benchmark's equicorrelation matrix: C_ii = 1 + (D-1) off, C_ij = off."""
    Cm = np.full((D, D), off)
    np.fill_diagonal(Cm, 1.0 + (D - 1) * off)
    return Cm


def dijet_p0() -> float:
    """Normalisation of the dijet truth shape for 1e6 expected events."""
    w = 6000.0 / 100
    s = 0.0
    for i in range(100):
        x = (1000.0 + (i + 0.5) * w) / 13000.0
        s += (1 - x) ** 10 / x ** (5 + 0.1 * math.log(x))
    return 1e6 / s
