"""GPU parity of the tile form (16 starts per workgroup, Gaussian objective
evaluated for the tile on FP64 MFMA, H kept as B + rank-2 pairs) against the
CPU oracle, through the C ABI.

The tile form sums in another order than the serial oracle, so parity is to
tolerance: equal statuses, |df| <= 1e-8 max(1, |f|), ||dx||_inf <= 1e-6.
Iteration counts may differ once f is at roundoff level, where line-search
decisions are decided by the last bits.
"""
import numpy as np
import pytest

import zeus_amd as z

pytestmark = pytest.mark.gpu

F_TOL = 1e-8
X_TOL = 1e-6


def gauss_matrix(D, off=0.5):
    C = np.full((D, D), off)
    np.fill_diagonal(C, 1.0 + (D - 1) * off)
    return C


def ill_conditioned(D, kappa, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((D, D)))
    C = (Q * np.logspace(0.0, np.log10(kappa), D)) @ Q.T
    return 0.5 * (C + C.T)


def close_f(a, b):
    return abs(a - b) <= F_TOL * max(1.0, abs(a), abs(b))


def check_against_oracle(ps, o, *, x_tol=X_TOL):
    assert np.array_equal(ps["status"], o.status)
    assert all(close_f(a, b) for a, b in zip(ps["fval"], o.fval))
    assert np.abs(ps["x"] - o.x).max() <= x_tol


FORMS = ["tile"]


@pytest.mark.parametrize("form", FORMS)
def test_tile_gaussian150_matches_oracle(orc, form):
    """C2 problem, 64 starts (4 tiles), no early stop."""
    C = gauss_matrix(150)
    N = 64
    r = z.Zeus(z.Gaussian(C), -5, 5, N, 10000, 10, N + 1, "gaussian", 1e-8, 42, 0,
               form=form, per_start=True)
    o = orc.run(orc.GAUSSIAN, 150, -5, 5, N, 10000, 10, N + 1, 1e-8, 42, params=C.reshape(-1))
    ps = r.per_start
    assert set(o.status) == {1}
    check_against_oracle(ps, o)
    assert r.status == 1 and close_f(r.fval, o.best["fval"])
    assert np.abs(r.coordinates - o.best_x).max() <= X_TOL
    assert r.c.claimed == N and r.c.actual == N
    # iteration counts follow the oracle's; a start may take a few more or
    # fewer once f is at roundoff level (line-search decisions on last bits)
    assert np.median(np.abs(ps["iter"] - o.iter)) <= 1
    assert r.sum_grads == int(ps["grads"].sum()) and r.sum_iter == int(ps["iter"].sum())


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("D,kappa,N", [(32, 1e3, 40), (150, 1e3, 24)])
def test_tile_folds_pair_list(orc, D, kappa, N, form):
    """Ill-conditioned quadratics take more than 32 iterations, so the rank-2
    pair list is folded into the explicit B at least once (4 times at
    D = 150); the minimizer and statuses still match the oracle."""
    C = ill_conditioned(D, kappa, 5)
    r = z.Zeus(z.Gaussian(C), -5, 5, N, 10000, 5, N + 1, "gaussian", 1e-8, 42, 0,
               form=form, per_start=True)
    o = orc.run(orc.GAUSSIAN, D, -5, 5, N, 10000, 5, N + 1, 1e-8, 42, params=C.reshape(-1))
    ps = r.per_start
    assert set(o.status) == {1} and o.iter.min() > 32
    assert ps["iter"].min() > 32
    check_against_oracle(ps, o)
    # iteration counts stay close to the oracle's (same method, other rounding)
    assert np.abs(np.median(ps["iter"]) - np.median(o.iter)) <= 3


@pytest.mark.parametrize("form", FORMS)
def test_tile_batch_of_problems(orc, form):
    """Two Gaussian-32 problems with partial last tiles (20 and 37 starts) and
    a Gaussian-150 problem in one call: tiles map to their own problem."""
    C1 = ill_conditioned(32, 10.0, 1)
    C2 = gauss_matrix(32, 0.25)
    C3 = gauss_matrix(150)
    probs = [z.Problem(z.Gaussian(C1), -5, 5, 20, 10000, 3, 21, 1e-8, 7),
             z.Problem(z.Gaussian(C2), -3, 3, 37, 10000, 4, 38, 1e-8, 8, stream=3),
             z.Problem(z.Gaussian(C3), -5, 5, 17, 10000, 2, 18, 1e-8, 9, stream=1)]
    res = z.run_batch(probs, form=form, per_start=True)
    for pr, r, C in zip(probs, res, (C1, C2, C3)):
        D = C.shape[0]
        o = orc.run(orc.GAUSSIAN, D, pr.lower, pr.upper, pr.N, pr.MAX_ITER, pr.PSO_ITER,
                    pr.requiredConverged, pr.tolerance, pr.seed, stream=pr.stream,
                    params=C.reshape(-1))
        check_against_oracle(r.per_start, o)
        assert r.idx == o.best["idx"] or close_f(r.fval, o.best["fval"])


@pytest.mark.parametrize("form", FORMS)
def test_tile_surrender_and_random_starts(orc, form):
    """MAX_ITER = 3 ends every start with status 0 after 3 iterations; with
    PSO_ITER < 0 the starts are drawn in the BFGS from the Philox stream."""
    C = gauss_matrix(150)
    N = 32
    r = z.Zeus(z.Gaussian(C), -5, 5, N, 3, -1, N + 1, "gaussian", 1e-8, 11, 0,
               form=form, per_start=True)
    o = orc.run(orc.GAUSSIAN, 150, -5, 5, N, 3, -1, N + 1, 1e-8, 11, params=C.reshape(-1))
    ps = r.per_start
    assert set(ps["status"]) == {0} and set(ps["iter"]) == {3}
    check_against_oracle(ps, o, x_tol=1e-6 * max(1.0, np.abs(o.x).max()))
    assert r.c.surrendered == N


@pytest.mark.parametrize("form", FORMS)
def test_tile_max_iter_zero(orc, form):
    C = gauss_matrix(32)
    r = z.Zeus(z.Gaussian(C), -5, 5, 16, 0, 2, 17, "gaussian", 1e-8, 3, 0, form=form,
               per_start=True)
    o = orc.run(orc.GAUSSIAN, 32, -5, 5, 16, 0, 2, 17, 1e-8, 3, params=C.reshape(-1))
    check_against_oracle(r.per_start, o)
    assert set(r.per_start["iter"]) == {0}


@pytest.mark.parametrize("form", FORMS)
def test_tile_early_stop(orc, form):
    """required = 5 of 160 starts: the call stops once 5 have converged, every
    start is converged or stopped, and stopped starts stopped no later than
    the oracle's unstopped run converges them."""
    C = gauss_matrix(150)
    N = 160
    r = z.Zeus(z.Gaussian(C), -5, 5, N, 10000, 10, 5, "gaussian", 1e-8, 42, 0, form=form,
               per_start=True)
    ps = r.per_start
    assert r.c.claimed >= 5 and r.c.claimed + r.c.stopped == N
    # the converged count (added per tile and round) is the per-start count
    assert r.c.claimed == int((ps["status"] == 1).sum())
    assert r.status == 1 and abs(r.fval) <= F_TOL
    o = orc.run(orc.GAUSSIAN, 150, -5, 5, N, 10000, 10, N + 1, 1e-8, 42, params=C.reshape(-1))
    conv = ps["status"] == 1
    assert np.abs(ps["x"][conv] - o.x[conv]).max() <= X_TOL
    st = ps["status"] == 2
    assert np.all(ps["iter"][st] <= o.iter[st] + 1)


def test_tile_early_stop_per_problem():
    """Three problems of one tile class with their own required counts (one
    never reached): each problem's converged counters count its own starts
    only (the tiles of all problems add theirs in the same rounds), its
    converged count is its status-1 starts, and a problem stops only on its
    own count."""
    C = gauss_matrix(150)
    probs = [z.Problem(z.Gaussian(C), -5, 5, 300, 10000, 10, 10, 1e-8, 42),
             z.Problem(z.Gaussian(C), -5, 5, 64, 10000, 10, 65, 1e-8, 43, stream=1),
             z.Problem(z.Gaussian(C), -5, 5, 120, 10000, 10, 2, 1e-8, 44, stream=2)]
    res = z.run_batch(probs, form="tile", per_start=True)
    for pr, r in zip(probs, res):
        st = r.per_start["status"]
        assert st.shape == (pr.N,)
        assert r.c.claimed == int((st == 1).sum())
        assert r.c.stopped == int((st == 2).sum())
        assert r.c.claimed + r.c.stopped + r.c.surrendered == pr.N
        if pr.requiredConverged > pr.N:
            assert r.c.stopped == 0 and r.c.claimed == pr.N
        else:
            assert r.c.claimed >= pr.requiredConverged
        assert r.status == 1 and abs(r.fval) <= F_TOL


def test_tile_is_tuned_for_the_headline():
    """The per-workgroup forms of a Gaussian-150 adjoint call include the tile
    form; the autotuner keeps the fastest shape for the size bucket."""
    C = gauss_matrix(150)
    r = z.Zeus(z.Gaussian(C), -5, 5, 512, 10000, 10, 513, "gaussian", 1e-8, 42, 0)
    assert r.status == 1 and r.c.claimed == 512
    form, threads = z.tuned_config(4, 150, 0)
    assert form in (2, 3) and threads in (256, 512, 1024)


def test_graph_replays_match_eager():
    """use_graph: the launch sequence is captured once per batch shape and
    replayed; every replay (with eager calls in between) gives the eager
    results bit for bit (the kernels are deterministic)."""
    C = gauss_matrix(32)
    probs = [z.Problem(z.Gaussian(C), -5, 5, 40, 10000, 3, 41, 1e-8, 7),
             z.Problem(z.Rosenbrock(2), -5, 5, 300, 10000, 5, 301, 1e-8, 8)]
    eager = z.run_batch(probs, per_start=True)
    for k in range(4):  # tune, capture, replay, replay
        g = z.run_batch(probs, per_start=True, use_graph=True)
        for a, b in zip(eager, g):
            for key in ("status", "iter", "fval", "x"):
                assert np.array_equal(a.per_start[key], b.per_start[key]), (k, key)
            assert a.idx == b.idx and a.c.claimed == b.c.claimed
        z.run_batch(probs[1:], per_start=False)  # eager work between replays


def test_graph_replay_follows_a_new_matrix():
    """A replayed graph prepares the matrix of the call that replays it: a
    second Gaussian matrix of the same shape (another device buffer) gives
    the eager results of that matrix bit for bit, in the tile form and the
    per-workgroup form, and the first matrix's results again after it."""
    C1 = gauss_matrix(32)
    C2 = ill_conditioned(32, 50.0, 11)
    for form, threads in (("tile", 0), ("group", 256)):
        def probs(C):
            return [z.Problem(z.Gaussian(C), -5, 5, 48, 10000, 3, 49, 1e-8, 7)]
        kw = dict(per_start=True, form=form, group_threads=threads)
        e1 = z.run_batch(probs(C1), **kw)[0]
        e2 = z.run_batch(probs(C2), **kw)[0]
        assert not np.array_equal(e1.per_start["x"], e2.per_start["x"])
        for k in range(3):  # capture with C1, replay with C2, replay with C1
            C = (C1, C2, C1)[k]
            want = (e1, e2, e1)[k]
            g = z.run_batch(probs(C), use_graph=True, **kw)[0]
            for key in ("status", "iter", "fval", "x"):
                assert np.array_equal(want.per_start[key], g.per_start[key]), (form, k, key)


@pytest.mark.parametrize("prng", [z.PRNGType.PHILOX, z.PRNGType.XORWOW], ids=["philox", "xorwow"])
@pytest.mark.parametrize("N", [512, 333])
def test_tile_repeatable_without_early_stop(prng, N):
    """Without an early stop the tile form's per-start outputs depend on
    nothing but the inputs: two calls agree bit for bit (also with a ragged
    last tile), and the converged count equals the starts with status 1."""
    C = gauss_matrix(150)
    f = z.Gaussian(C)
    a, b = (z.Zeus(f, -5, 5, N, 10000, 10, N + 1, "gaussian", 1e-8, 42, 0, prng=prng,
                   form="tile", per_start=True) for _ in range(2))
    for key in ("status", "iter", "ls_evals", "grads", "fval", "gnorm", "x"):
        assert np.array_equal(a.per_start[key], b.per_start[key]), key
    assert a.idx == b.idx and a.fval == b.fval and a.sum_ls_evals == b.sum_ls_evals
    assert a.c.claimed == int((a.per_start["status"] == 1).sum()) == N


def test_tile_more_tiles_than_workgroups():
    """4500 starts make 282 tiles: with one workgroup per CU some workgroups
    run two tiles one after the other (slots, counts and the stop flag are
    reset per tile).  Without early stop two calls agree bit for bit and
    every start converges to the minimum at 0; with required = 100 the
    converged count (added per tile and round) equals the starts with status
    1, and every start is converged or stopped."""
    C = gauss_matrix(150)
    f = z.Gaussian(C)
    N = 4500
    a, b = (z.Zeus(f, -5, 5, N, 10000, 10, N + 1, "gaussian", 1e-8, 42, 0, prng=z.PRNGType.PHILOX,
                   form="tile", per_start=True) for _ in range(2))
    for key in ("status", "iter", "fval", "x"):
        assert np.array_equal(a.per_start[key], b.per_start[key]), key
    assert set(a.per_start["status"]) == {1} and a.c.claimed == N
    assert np.abs(a.per_start["x"]).max() <= 1e-6
    r = z.Zeus(f, -5, 5, N, 10000, 10, 100, "gaussian", 1e-8, 42, 0, prng=z.PRNGType.PHILOX,
               form="tile", per_start=True)
    st = r.per_start["status"]
    assert r.c.claimed >= 100 and r.c.claimed == int((st == 1).sum())
    assert set(st) <= {1, 2} and r.c.claimed + r.c.stopped == N
