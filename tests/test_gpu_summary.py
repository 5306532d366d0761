"""The summary of a call on the GPU: the best start and the counters.

The reference picks the best start with an argmin over the per-start values
(include/bfgs_common.cuh:280-319: NaN never wins, ties go to the lowest
index) and returns that start's status, iteration, value, gradient norm and
coordinates; the convergence counts are sums over the starts
(include/util.hpp:151-200).  Each check compares the Result of a call with
the per-start outputs of the same call, which is what the summary reduces.
The start counts fall on both sides of the summary's 1024-thread stride.
"""
import numpy as np
import pytest

import zeus_amd as z

pytestmark = pytest.mark.gpu


def best_start(f):
    """Argmin with NaN never winning and ties to the lowest index; 0 when
    every value is NaN."""
    f = np.asarray(f)
    ok = ~np.isnan(f)
    if not ok.any():
        return 0
    return int(np.flatnonzero(ok & (f == f[ok].min()))[0])


def same(a, b):
    return (np.isnan(a) and np.isnan(b)) or a == b


def check_summary(r):
    ps = r.per_start
    b = best_start(ps["fval"])
    assert r.idx == b
    assert r.status == ps["status"][b]
    assert r.iter == ps["iter"][b]
    assert same(r.fval, ps["fval"][b])
    assert same(r.gradientNorm, ps["gnorm"][b])
    assert np.array_equal(r.coordinates, ps["x"][b], equal_nan=True)
    st = ps["status"]
    assert r.c.claimed == np.sum(st == 1)
    assert r.c.stopped == np.sum(st == 2)
    assert r.c.surrendered == np.sum((st != 1) & (st != 2))
    assert r.nonfinite == np.sum(st == 5)
    assert r.sum_iter == ps["iter"].sum()
    assert r.sum_ls_evals == ps["ls_evals"].sum()
    assert r.sum_grads == ps["grads"].sum()


def test_summary_matches_per_start_outputs():
    """Mixed statuses (MAX_ITER 20 leaves most starts surrendered, the last
    problem stops early after 10 converged), a batch of start counts from 1
    to beyond two strides."""
    counts = [1, 5, 1023, 1024, 1025, 2500]
    required = [0, 0, 0, 0, 0, 10]
    probs = [z.Problem(z.Rosenbrock(2), -5, 5, n, 20, 1, q, 1e-8, 7 + i, stream=i)
             for i, (n, q) in enumerate(zip(counts, required))]
    rs = z.run_batch(probs, per_start=True)
    statuses = set()
    for r in rs:
        check_summary(r)
        statuses |= set(r.per_start["status"].tolist())
    assert {0, 1, 2} <= statuses


def test_summary_with_non_finite_starts():
    """Starts at a point where the Poisson NLL is not finite end with status
    5; every other start keeps a finite value, so the winner is a finite
    start.  With every start there, every value is non-finite and the winner
    is start 0 by the tie rule."""
    f = z.LogLikelihood.from_params(z.make_dijet(42, 0))
    bad = np.array([0.0, 10.0, 5.0, 0.1])
    rng = np.random.default_rng(11)
    n = 1500
    x0 = rng.uniform(0.5, 9.5, (n, 4))
    x0[::2] = bad
    r = z.Zeus(f, 0, 10, n, 30, 0, 0, "poisson", 1e-8, 3, 0, form="thread", x0=x0,
               per_start=True)
    assert np.all(r.per_start["status"][::2] == 5)
    check_summary(r)
    assert r.idx % 2 == 1

    x0[:] = bad
    r = z.Zeus(f, 0, 10, n, 30, 0, 0, "poisson", 1e-8, 3, 0, form="thread", x0=x0,
               per_start=True)
    assert r.nonfinite == n
    check_summary(r)
    assert r.idx == 0
