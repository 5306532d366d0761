"""The secondary measured configurations besides the headline bench line:
build one workload, run it warm, print one JSON line with its timing and
counts.

    python -m zeus_amd.workloads rosenbrock|dijet|mlp|mixed [--calls K] [--max-iter M]
                                 [--gradient G] [--graph] [--form F [--group-threads T]]

rosenbrock  Rosenbrock D = 2, box [-5, 5], 1024 starts, PSO 10, MAX_ITER 10000,
    required 100, tol 1e-8, seed 42 (the reference's CPU-runnable case).
dijet  1024 dijet fits x 256 starts: Poisson NLL over 100 bins, D = 4, box
    [0, 10], PSO 10, required 100 per fit, tol 1e-8; counts drawn per fit
    from the seed (make_dijet).  MAX_ITER defaults to 100: at tol 1e-8 the
    NLL gradient sits on its roundoff floor and starts surrender.
mlp  MLP 1-16-16-1 tanh (321 parameters) least squares over 256 points, box
    [-2, 2], 2048 starts, PSO 10, tol 1e-8; MAX_ITER defaults to 200.
mixed  64 problems in one batched call: Rosenbrock D in {2, 10, 50}, Rastrigin
    D = 20, Ackley D = 10, Gaussian D in {32, 150} (C = A A^T / D + I), start
    counts log-uniform in [128, 4096], required min(100, starts); kinds and
    counts drawn from a numpy generator seeded with 42; PSO 10, MAX_ITER
    10000, tol 1e-8.  Reports each class's device times and rates
    (class_table); --graph replays the settled batch from a captured
    hipGraph (the per-class times are then not separated).

The first call of a workload is a warm-up (code objects, pools, autotuner);
the line reports the fastest of the timed calls.  With --form the warm-up
call runs 10 iterations (at most): the way to time the full MAX_ITER 10000 of
dijet and mlp, whose tuning call would otherwise run every shape to the cap.
--form thread|tile, or group with --group-threads, fixes the shape; --form
group alone still lets the tuner choose among the parallel shapes, on that
short warm-up call.  Synthetic data, FP64.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

import zeus_amd as z


def mixed_specs():
    """The 64 problems of the mixed batch as plain records: kind (objective
    name), dim, box half-width, starts, required, seed and stream.
    Kinds and start counts come from a numpy generator seeded with 42."""
    rng = np.random.default_rng(42)
    specs = []
    for i in range(64):
        k = int(rng.integers(0, 7))
        n = int(round(math.exp(rng.uniform(math.log(128), math.log(4096)))))
        if k <= 2:
            kind, dim, box = "rosenbrock", (2, 10, 50)[k], 5.0
        elif k == 3:
            kind, dim, box = "rastrigin", 20, 5.12
        elif k == 4:
            kind, dim, box = "ackley", 10, 32.768
        else:
            kind, dim, box = "gaussian", (32, 150)[k - 5], 5.0
        specs.append(dict(kind=kind, dim=dim, box=box, n=n, required=min(100, n), seed=42 + i,
                          stream=i, name=f"{kind}{dim}"))
    return specs


def mixed_objective(spec):
    """The objective of one mixed-batch record (Gaussian: C = A A^T / D + I,
    A ~ N(0, 1) drawn from (42, stream))."""
    kind, dim = spec["kind"], spec["dim"]
    if kind == "rosenbrock":
        return z.Rosenbrock(dim)
    if kind == "rastrigin":
        return z.Rastrigin(dim)
    if kind == "ackley":
        return z.Ackley(dim)
    return z.Gaussian(z.make_spd(42, spec["stream"], dim))


def _mixed_problems(max_starts=None, max_iter=10000, early_stop=True):
    """The mixed batch as Problems; max_starts caps every start count and
    early_stop=False asks for every start (required = N + 1)."""
    probs, names = [], []
    for sp in mixed_specs():
        n = sp["n"] if max_starts is None else min(sp["n"], max_starts)
        req = min(sp["required"], n) if early_stop else n + 1
        probs.append(z.Problem(mixed_objective(sp), -sp["box"], sp["box"], n, max_iter, 10, req,
                               1e-8, sp["seed"], stream=sp["stream"]))
        names.append(sp["name"])
    return probs, names


def run(name: str, calls: int, max_iter: int | None, gradient: str, graph: bool = False,
        form: str = "auto", group_threads: int = 0) -> dict:
    out = {"workload": name, "gradient": gradient, "data": "synthetic", "dtype": "f64"}
    fixed = form != "auto"
    kw = dict(form=form, group_threads=group_threads)
    if fixed:
        out.update(form=form, group_threads=group_threads)
    if name == "rosenbrock":
        it = max_iter or 10000

        def make(n_it):
            return lambda: [z.Zeus(z.Rosenbrock(2), -5, 5, 1024, n_it, 10, 100, "rosenbrock",
                                   1e-8, 42, 0, gradient=gradient, **kw)]
        out["starts"] = 1024
    elif name == "dijet":
        it = max_iter or 100

        def make(n_it):
            probs = [z.Problem(z.LogLikelihood.from_params(z.make_dijet(42, e)), 0.0, 10.0, 256,
                               n_it, 10, 100, 1e-8, 42, stream=e) for e in range(1024)]
            return lambda: z.run_batch(probs, gradient=gradient, **kw)
        out.update(starts=1024 * 256, max_iter=it)
    elif name == "mlp":
        it = max_iter or 200
        d = z.make_mlp_data(42, 256)
        f = z.Mlp(d[:256], d[256:])

        def make(n_it):
            # without --form: Zeus's own default for parallel=True (the group form)
            return lambda: [z.Zeus(f, -2, 2, 2048, n_it, 10, 100, "mlp", 1e-8, 42, 0,
                                   gradient=gradient, **(kw if fixed else {}))]
        out.update(starts=2048, max_iter=it)
    elif name == "mixed":
        it = max_iter or 10000
        probs, names = _mixed_problems(max_iter=it)

        def make(n_it):
            return lambda: z.run_batch(probs, gradient=gradient, use_graph=graph, **kw)
        out.update(starts=sum(p.N for p in probs), problems=len(probs), graph=graph)
    else:
        raise SystemExit(f"unknown workload {name!r} (rosenbrock, dijet, mlp, mixed)")
    call = make(it)
    # warm-up: with a fixed form only code objects and pools need it
    (make(min(it, 10)) if fixed and name != "mixed" else call)()
    best, res = float("inf"), None
    for _ in range(calls):
        t = time.perf_counter()
        r = call()
        dt = time.perf_counter() - t
        if dt < best:
            best, res = dt, r
    conv = sum(r.c.claimed for r in res)
    evals = sum(r.sum_ls_evals + r.sum_grads for r in res)
    out.update(ms_per_call=best * 1e3, converged=conv, converged_per_s=conv / best,
               sum_iter=sum(r.sum_iter for r in res), sum_ls_evals=sum(r.sum_ls_evals for r in res),
               sum_grads=sum(r.sum_grads for r in res), ms_pso=res[0].ms_pso,
               ms_opt=res[0].ms_opt)
    if name == "dijet":  # the work unit is one bin evaluation (100 per objective evaluation)
        out["bin_evals_per_s"] = evals * 100 / best
    if name == "mlp":
        out["start_iterations_per_s"] = out["sum_iter"] / best
    if name == "mixed" and not graph:
        out["classes"] = class_table(probs, names, res)
    return out


def class_table(probs, names, res) -> dict:
    """Per class (objective and dimension) of a mixed batch: its problems and
    starts, its PSO and BFGS device times, and its work rates.  A class's
    problems share one PSO and one BFGS launch sequence; side by side with
    the other classes once the shapes are settled.  For the Gaussian
    classes on the tile forms the BFGS is FP64-MFMA bound: its algorithmic
    rate is 2 D^2 flops per objective evaluation (first point and every
    line-search trial; the gradient is the product's by-product) over the
    class's BFGS time, against the 78.6 TFLOP/s FP64 peak.  The other classes
    are bound by their objective's VALU work (transcendentals) or latency,
    and report evaluations per second."""
    by = {}
    for pr, name, r in zip(probs, names, res):
        c = by.setdefault(name, dict(problems=0, starts=0, converged=0, evals=0, iters=0,
                                     ms_pso=r.ms_pso, ms_bfgs=r.ms_opt, dim=pr.f.dim))
        c["problems"] += 1
        c["starts"] += pr.N
        c["converged"] += r.c.claimed
        c["evals"] += r.sum_ls_evals + pr.N
        c["iters"] += r.sum_iter
    for name, c in by.items():
        t = c["ms_bfgs"] * 1e-3
        c["evals_per_s"] = c["evals"] / t if t > 0 else None
        c["start_iterations_per_s"] = c["iters"] / t if t > 0 else None
        if name.startswith("gaussian") and t > 0:
            tf = 2.0 * c["dim"] ** 2 * c["evals"] / t / 1e12
            c["roofline"] = {"bound": "mfma", "achieved_tflops": tf, "peak": 78.6,
                             "frac": tf / 78.6}
    return by


def main(argv: list[str]) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("workload", choices=["rosenbrock", "dijet", "mlp", "mixed"])
    ap.add_argument("--calls", type=int, default=2)
    ap.add_argument("--max-iter", type=int, default=None)
    ap.add_argument("--gradient", default="adjoint", choices=["adjoint", "dual"])
    ap.add_argument("--graph", action="store_true",
                    help="mixed: replay the batch from a captured hipGraph once settled")
    ap.add_argument("--form", default="auto", choices=["auto", "thread", "group", "team", "tile"],
                    help="fix the BFGS form (no autotuning)")
    ap.add_argument("--group-threads", type=int, default=0,
                    help="workgroup size of the group form (with --form group)")
    a = ap.parse_args(argv)
    print(json.dumps(run(a.workload, a.calls, a.max_iter, a.gradient, a.graph, a.form,
                         a.group_threads)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
