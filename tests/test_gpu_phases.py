"""The PSO phases and the update primitive of the C ABI (zeus_pso_init,
zeus_pso_step, zeus_bfgs_update): the phases reproduce the one-call PSO and
the oracle; the batched update reproduces the oracle's update and the
reference's update KAT (test/bfgs_tests.cu:228-264)."""
import numpy as np
import pytest

import zeus_amd as z

pytestmark = pytest.mark.gpu
GENS = [z.PRNGType.XORWOW, z.PRNGType.PHILOX, z.PRNGType.SOBOL]


def _gauss32():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(32, 32))
    return A @ A.T / 32 + np.eye(32)


@pytest.mark.parametrize("case", ["rosenbrock4", "gaussian32"])
def test_pso_phases_equal_one_call(orc, case):
    N, T = 300, 5
    if case == "rosenbrock4":
        f, kind, params = z.Rosenbrock(4), orc.ROSENBROCK, None
    else:
        C = _gauss32()
        f, kind, params = z.Gaussian(C), orc.GAUSSIAN, C.reshape(-1)
    for g in GENS:
        sw = z.Swarm(f, -5.0, 5.0, N, 42, prng=g).init()
        for t in range(1, T + 1):
            sw.step(t)
        got = sw.state()
        pbx, pbv, gbx, gbv = z.pso(f, -5.0, 5.0, N, T, 42, prng=g)
        assert np.array_equal(got["pbest_x"], pbx) and np.array_equal(got["pbest_val"], pbv)
        assert np.array_equal(got["gbest_x"], gbx) and got["gbest_val"] == gbv
        o = orc.pso(kind, f.dim, -5.0, 5.0, N, T, 42, params=params, prng=int(g))
        if case == "rosenbrock4":  # lane form: bitwise with the oracle
            assert np.array_equal(got["pbest_x"], o[0]) and got["gbest_val"] == o[3]
        else:  # MFMA evaluation: values to rounding
            assert np.allclose(got["pbest_val"], o[1], rtol=1e-12, atol=0)
            assert abs(got["gbest_val"] - o[3]) <= 1e-12 * abs(o[3])


@pytest.mark.parametrize("g", GENS, ids=[x.name for x in GENS])
def test_pso_positions_and_velocities_equal_oracle(orc, g):
    """Positions and velocities after init and after two steps equal the
    oracle's swarm bit for bit (lane form, Rosenbrock-4 on [-3, 7]); the
    oracle's PSO is pinned to the reference rule by a numpy restatement."""
    N, D, lo, hi = 37, 4, -3.0, 7.0
    sw = z.Swarm(z.Rosenbrock(D), lo, hi, N, 42, prng=g).init()
    for T in (0, 1, 2):
        if T:
            sw.step(T)
        got = sw.state()
        o = orc.pso_swarm(orc.ROSENBROCK, D, lo, hi, N, T, 42, prng=int(g))
        for k in ("x", "v", "pbest_x", "pbest_val", "gbest_x"):
            assert np.array_equal(got[k], o[k]), (T, k)
        assert got["gbest_val"] == o["gbest_val"]


def test_pso_step_keeps_invariants():
    sw = z.Swarm(z.Rastrigin(2), -5.12, 5.12, 64, 7).init()
    prev = sw.state()
    assert prev["gbest_val"] == prev["pbest_val"].min()
    for t in range(1, 8):
        cur = sw.step(t).state()
        assert np.all(cur["pbest_val"] <= prev["pbest_val"])
        assert cur["gbest_val"] == cur["pbest_val"].min() <= prev["gbest_val"]
        assert np.array_equal(cur["gbest_x"], cur["pbest_x"][np.argmin(cur["pbest_val"])])
        prev = cur
    with pytest.raises(z.capi.ZeusError):
        sw.step(0)


def test_bfgs_update_kat_and_oracle(orc):
    # reference KAT: H = I, s = (1, 0), y = (2, 0), s.y = 2 -> [[0.5, 0], [0, 1]]
    H = z.bfgs_update(np.eye(2)[None], np.array([[1.0, 0.0]]), np.array([[2.0, 0.0]]))
    assert np.array_equal(H[0], [[0.5, 0.0], [0.0, 1.0]])
    rng = np.random.default_rng(3)
    for n in (1, 7, 150, 321):
        B = 5
        A = rng.normal(size=(B, n, n))
        H0 = np.einsum("bij,bkj->bik", A, A) / n + np.eye(n)
        s = rng.normal(size=(B, n))
        y = rng.normal(size=(B, n))
        y[1] = 1e-20 * s[1]  # |s.y| < 1e-14: skipped
        got = z.bfgs_update(H0, s, y)
        for b in range(B):
            sty = orc.dot(s[b], y[b])
            assert np.array_equal(got[b], orc.update(H0[b], s[b], y[b], sty)), (n, b)
        assert np.array_equal(got[1], H0[1])


def test_pso_init_single_particle():
    # reference test/pso_tests.cu:27-94: with N = 1 the initial pBest is the
    # particle, pBestVal = gBestVal = f(pBestX) and gBestX = pBestX
    sw = z.Swarm(z.Rastrigin(2), -5.0, 5.0, 1, 42).init()
    st = sw.state()
    x = st["pbest_x"][0]
    want = 20.0 + sum(v * v - 10.0 * np.cos(2 * np.pi * v) for v in x)
    assert np.all((x >= -5.0) & (x <= 5.0)) and np.array_equal(st["x"][0], x)
    assert st["pbest_val"][0] == pytest.approx(want, abs=1e-12)
    assert st["gbest_val"] == st["pbest_val"][0]
    assert np.array_equal(st["gbest_x"], x)
