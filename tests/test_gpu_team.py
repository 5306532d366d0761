"""GPU parity of the team form (a few lanes per start, several starts per
wavefront; separable objectives of small dimension) against the CPU oracle
and against the per-workgroup form, through the C ABI.

Sums run in another fixed order than the serial oracle, so parity is to
tolerance, as for the other cooperative forms: statuses agree (a line-search
decision taken at roundoff level may send a start elsewhere, so a small
share may differ), and where both converge the minima agree to
|df| <= 1e-8 max(1, |f|) and the minimizers to 1e-6."""
import numpy as np
import pytest

import zeus_amd as z

pytestmark = pytest.mark.gpu
F_TOL, X_TOL = 1e-8, 1e-6

CASES = [  # objective, oracle id, dim, box, tol, MAX_ITER, team sizes offered
    ("rosenbrock", lambda: z.Rosenbrock(10), 10, 5.0, 1e-8, 10000, (4, 8)),
    ("rastrigin", lambda: z.Rastrigin(20), 20, 5.12, 1e-5, 500, (8, 16)),
    ("ackley", lambda: z.Ackley(10), 10, 32.768, 1e-5, 500, (4, 8)),
]


def close_f(a, b):
    return abs(a - b) <= F_TOL * max(1.0, abs(a), abs(b))


def _oracle_id(orc, name):
    return {"rosenbrock": orc.ROSENBROCK, "rastrigin": orc.RASTRIGIN, "ackley": orc.ACKLEY}[name]


@pytest.mark.parametrize("grad", ["adjoint", "dual"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_team_form_matches_oracle(orc, case, grad):
    name, make, dim, box, tol, max_iter, sizes = case
    N = 256
    o = orc.run(_oracle_id(orc, name), dim, -box, box, N, max_iter, 10, N + 1, tol, 42)
    for ts in sizes:
        r = z.Zeus(make(), -box, box, N, max_iter, 10, N + 1, name, tol, 42, 0, form="team",
                   group_threads=ts, gradient=grad, per_start=True)
        ps = r.per_start
        assert (ps["status"] == o.status).mean() >= 0.95, (name, ts)
        both = (ps["status"] == 1) & (o.status == 1)
        same = np.array([close_f(a, b) for a, b in zip(ps["fval"], o.fval)])
        assert (both & same).sum() >= 0.9 * both.sum() and both.sum() > 0, (name, ts)
        assert np.abs(ps["x"][both & same] - o.x[both & same]).max() <= X_TOL, (name, ts)
        assert r.sum_grads == int(ps["grads"].sum()) and r.sum_iter == int(ps["iter"].sum())


def test_team_form_agrees_with_group_form():
    """Same PSO result, same method: the team and per-workgroup forms end in the
    same place for almost every start."""
    N = 512
    f = z.Rastrigin(20)
    a = z.Zeus(f, -5.12, 5.12, N, 300, 10, N + 1, "r", 1e-5, 7, 0, form="team",
               group_threads=16, per_start=True)
    b = z.Zeus(f, -5.12, 5.12, N, 300, 10, N + 1, "r", 1e-5, 7, 0, form="group",
               group_threads=64, per_start=True)
    pa, pb = a.per_start, b.per_start
    assert (pa["status"] == pb["status"]).mean() >= 0.95
    same = np.array([close_f(x, y) for x, y in zip(pa["fval"], pb["fval"])])
    assert same.mean() >= 0.9


def test_team_form_early_stop_and_counts():
    N = 1000
    r = z.Zeus(z.Rosenbrock(10), -5, 5, N, 10000, 10, 20, "r", 1e-8, 42, 0, form="team",
               group_threads=8, per_start=True)
    st = r.per_start["status"]
    assert r.c.claimed == int((st == 1).sum()) >= 20
    assert r.c.stopped == int((st == 2).sum())
    assert r.c.claimed + r.c.stopped + r.c.surrendered == N
    assert r.status == 1 and abs(r.fval) <= 1e-12


def test_autotuner_offers_the_team_form():
    f = z.Ackley(10)
    for _ in range(2):
        r = z.Zeus(f, -32.768, 32.768, 2048, 200, 10, 2049, "a", 1e-5, 3, 0)
    form, threads = z.tuned_config(3, 10, 0)
    assert form in (z.capi.FORM_THREAD, z.capi.FORM_GROUP, z.capi.FORM_TEAM)
    assert r.status in (0, 1)


def test_team_form_batch_equals_single_calls_and_edge_shapes():
    """Starts do not interact without an early stop, so a batch of problems on
    the team form gives every start what its own call gives, bit for bit;
    ragged start counts (not a multiple of the teams of a workgroup) and
    MAX_ITER = 0 take the same paths."""
    probs = [z.Problem(z.Rastrigin(20), -5.12, 5.12, n, 150, 5, n + 1, 1e-5, 11, stream=i)
             for i, n in enumerate((37, 129, 5))]
    probs.append(z.Problem(z.Ackley(10), -32.768, 32.768, 300, 150, 5, 301, 1e-5, 11, stream=9))
    batch = z.run_batch(probs, form="team", group_threads=8, per_start=True)
    for p, rb in zip(probs, batch):
        (rs,) = z.run_batch([p], form="team", group_threads=8, per_start=True)
        for key in ("status", "iter", "fval", "x", "ls_evals", "grads"):
            assert np.array_equal(rb.per_start[key], rs.per_start[key]), key
    (r0,) = z.run_batch([z.Problem(z.Rosenbrock(10), -5, 5, 33, 0, 3, 34, 1e-8, 1)],
                        form="team", group_threads=4, per_start=True)
    assert set(r0.per_start["status"]) == {0} and set(r0.per_start["iter"]) == {0}
    assert r0.c.surrendered == 33


def test_settled_batch_runs_classes_side_by_side():
    """After the first call settles every class's shape, a batch's classes run
    on side streams (forked from and joined into the caller's stream), eagerly
    or as a captured graph: the results are those of the first call, bit for
    bit (classes share nothing but the stream order)."""
    C = z.make_spd(42, 3, 32)
    probs = [
        z.Problem(z.Rastrigin(20), -5.12, 5.12, 200, 300, 5, 201, 1e-5, 3, stream=1),
        z.Problem(z.Ackley(10), -32.768, 32.768, 150, 300, 5, 151, 1e-5, 3, stream=2),
        z.Problem(z.Gaussian(C), -5, 5, 100, 1000, 5, 101, 1e-8, 3, stream=3),
        z.Problem(z.Rosenbrock(2), -5, 5, 64, 1000, 5, 65, 1e-8, 3, stream=4),
    ]
    first = z.run_batch(probs, per_start=True)
    runs = [z.run_batch(probs, per_start=True),
            z.run_batch(probs, per_start=True, use_graph=True),
            z.run_batch(probs, per_start=True, use_graph=True)]
    for later in runs:
        for a, b in zip(first, later):
            for key in ("status", "iter", "fval", "x"):
                assert np.array_equal(a.per_start[key], b.per_start[key]), key
            assert a.idx == b.idx and a.fval == b.fval
    assert runs[0][0].ms_pso > 0 and runs[0][0].ms_opt > 0
