"""Pins the CPU oracle before it is trusted as the checker.

Every function of oracle/zeus_oracle.c is checked against a golden vector,
a reference-test known answer, a closed form or brute force, each chosen so
that a plausible slip (a dropped term, a wrong sign or index, a transposed
operand) fails at least one assertion.
"""
import math

import numpy as np
import pytest


# --- random numbers ---------------------------------------------------------
def test_philox_random123_kats(orc, golden):
    for v in golden["philox4x32_10"]:
        assert orc.philox(v["ctr"], v["key"]) == v["out"]


def test_uniform_conversion_closed_form(orc):
    # draw d of start s = words (2j, 2j+1) of block d>>1, top 53 bits, (m+1)/2^53
    for seed, stream, start, draw in [(42, 0, 0, 0), (42, 0, 0, 1), (7, 3, 12345, 77),
                                      (0xFFFFFFFF, 0x80000001, 2**40 + 3, 2**33 + 1)]:
        blk = draw >> 1
        o = orc.philox([blk & 0xFFFFFFFF, blk >> 32, start & 0xFFFFFFFF, start >> 32],
                       [seed, stream])
        j = 2 * (draw & 1)
        m = ((o[j] << 32) | o[j + 1]) >> 11
        assert orc.uniform(seed, stream, start, draw) == (m + 1) / 2.0**53


def test_uniform_range_and_moments(orc):
    u = np.array([orc.uniform(42, 0, s, d) for s in range(64) for d in range(64)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01


# --- dual numbers and digamma -------------------------------------------------
def test_digamma_scipy_vectors(orc, golden):
    for x, psi in zip(golden["digamma"]["x"], golden["digamma"]["psi"]):
        got = orc.digamma(x)
        assert got == pytest.approx(psi, rel=2e-13, abs=1e-14), x


def test_digamma_reference_kats(orc, golden):
    g = golden["reference_kats"]["digamma_exact"]["gamma"]
    tol = golden["reference_kats"]["digamma_exact"]["tol"]
    for x, want in [(1.0, -g), (2.0, 1 - g), (3.0, 1.5 - g), (0.5, -g - 2 * math.log(2))]:
        assert abs(orc.digamma(x) - want) <= tol * max(1.0, abs(want))
    eps = 1e-9
    for pole in (0.0, -1.0, -2.0):  # test/dual_tests.cu:534-571
        assert math.isinf(orc.digamma(pole)) and orc.digamma(pole) > 0
        r, l = orc.digamma(pole + eps), orc.digamma(pole - eps)
        assert r < 0 < l
        assert r == pytest.approx(-1 / eps, rel=1.2e-7)
        assert l == pytest.approx(1 / eps, rel=1.2e-7)


OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "neg": 4, "abs": 5, "sin": 6, "cos": 7,
       "exp": 8, "sqrt": 9, "atan2": 10, "log": 11, "pow": 13}


def test_dual_reference_kats(orc, golden):
    for name, a, b, want in golden["reference_kats"]["dual_ops"]["cases"]:
        op = OPS[name]
        if name == "pow":
            got = orc.dual_op(op, a, (0, 0), b)
        elif name == "atan2":
            got = orc.dual_op(op, a, b)
        else:
            got = orc.dual_op(op, a, b or (0, 0))
        assert got[0] == pytest.approx(want[0], abs=1e-12), name
        assert got[1] == pytest.approx(want[1], abs=1e-12), name


def test_dual_abs_at_zero_is_nan(orc):
    r = orc.dual_op(OPS["abs"], (0.0, 1.5))  # duals.cuh:144-146
    assert r[0] == 0.0 and math.isnan(r[1])


def test_dual_closed_form_derivatives(orc):
    # log, tanh, lgamma, pow(dual, dual) against analytic derivatives
    x = 1.7
    assert orc.dual_op(11, (x, 1.0))[1] == pytest.approx(1 / x, rel=1e-15)
    assert orc.dual_op(15, (x, 1.0))[1] == pytest.approx(1 - math.tanh(x) ** 2, rel=1e-14)
    assert orc.dual_op(14, (x, 1.0))[1] == pytest.approx(orc.digamma(x), rel=1e-15)
    b, e = 1.3, 2.6  # d/db b^e = e b^(e-1);  d/de b^e = b^e ln b
    assert orc.dual_op(12, (b, 1.0), (e, 0.0))[1] == pytest.approx(e * b ** (e - 1), rel=1e-14)
    assert orc.dual_op(12, (b, 0.0), (e, 1.0))[1] == pytest.approx(b ** e * math.log(b), rel=1e-14)
    # pow(dual, double) at a zero base takes the e * 0^(e-1) branch (duals.cuh:216-218)
    assert orc.dual_op(13, (0.0, 1.0), (0, 0), 3.0) == (0.0, 0.0)


# --- objectives -----------------------------------------------------------------
def _objs(orc, rng):
    C = rng.standard_normal((6, 6))
    kx = np.linspace(0.1, 0.5, 12)
    kk = np.round(3.2 * (1 - kx) ** 10 / kx ** (5 + 0.1 * np.log(kx)))
    lg = np.array([math.lgamma(v + 1) for v in kk])
    xs = rng.uniform(-1, 1, 8)
    return [
        (orc.Obj(orc.ROSENBROCK, 5), lambda: rng.uniform(-2, 2, 5)),
        (orc.Obj(orc.RASTRIGIN, 4), lambda: rng.uniform(-5, 5, 4)),
        (orc.Obj(orc.ACKLEY, 3), lambda: rng.uniform(-3, 3, 3)),
        (orc.Obj(orc.GAUSSIAN, 6, C.reshape(-1)), lambda: rng.uniform(-5, 5, 6)),
        (orc.Obj(orc.LOGLIK, 4, np.concatenate([kx, kk, lg])),
         lambda: np.array([3.0, 10.0, 5.0, 0.1]) + rng.uniform(-0.3, 0.3, 4)),
        (orc.Obj(orc.MLP, 321, np.concatenate([xs, np.sin(3 * xs)])),
         lambda: rng.uniform(-1, 1, 321)),
        (orc.Obj(orc.SPHERE, 3), lambda: rng.uniform(-5, 5, 3)),
    ]


def test_reference_function_kats(orc, golden):
    k = golden["reference_kats"]["fun_values"]
    assert orc.Obj(orc.RASTRIGIN, 2).value([0, 0]) == k["rastrigin_origin"]
    assert abs(orc.Obj(orc.ACKLEY, 2).value([0, 0])) <= k["ackley_origin_abs_tol"]
    assert orc.Obj(orc.ROSENBROCK, 2).value([1, 1]) == k["rosenbrock_min"]
    assert list(orc.Obj(orc.RASTRIGIN, 2).grad_dual([0, 0])) == [0.0, 0.0]
    assert list(orc.Obj(orc.ROSENBROCK, 2).grad_dual([1, 1])) == [0.0, 0.0]
    ack = orc.Obj(orc.ACKLEY, 2)
    g = ack.grad_dual(k["ackley_grad_point"])
    assert g == pytest.approx(k["ackley_grad"], rel=1e-12)
    assert ack.value_grad_adjoint(k["ackley_grad_point"])[1] == pytest.approx(k["ackley_grad"],
                                                                            rel=1e-12)
    assert np.isnan(ack.grad_dual([0.0, 0.0])).all()  # sqrt'(0): dual gives NaN


def test_closed_form_gradients(orc):
    rng = np.random.default_rng(1)
    x = rng.uniform(-2, 2, 4)
    # Rosenbrock closed form
    g = np.zeros(4)
    for i in range(3):
        g[i] += -2 * (1 - x[i]) - 400 * x[i] * (x[i + 1] - x[i] ** 2)
        g[i + 1] += 200 * (x[i + 1] - x[i] ** 2)
    r = orc.Obj(orc.ROSENBROCK, 4)
    assert r.value_grad_adjoint(x)[1] == pytest.approx(g, rel=1e-13)
    # Rastrigin closed form
    ra = orc.Obj(orc.RASTRIGIN, 4)
    assert ra.value_grad_adjoint(x)[1] == pytest.approx(2 * x + 20 * math.pi *
                                                        np.sin(2 * math.pi * x), rel=1e-13)
    # Gaussian with a non-symmetric C: grad = 0.5 (C + C^T) x
    C = rng.standard_normal((4, 4))
    ga = orc.Obj(orc.GAUSSIAN, 4, C.reshape(-1))
    f, gg = ga.value_grad_adjoint(x)
    assert f == pytest.approx(0.5 * x @ C @ x, rel=1e-13)
    assert gg == pytest.approx(0.5 * (C + C.T) @ x, rel=1e-13)
    assert ga.grad_dual(x) == pytest.approx(0.5 * (C + C.T) @ x, rel=1e-13)


def test_adjoint_value_is_bit_identical_and_matches_dual(orc):
    rng = np.random.default_rng(7)
    for obj, draw in _objs(orc, rng):
        for _ in range(3):
            x = draw()
            f, ga = obj.value_grad_adjoint(x)
            assert f == obj.value(x) or (math.isnan(f) and math.isnan(obj.value(x)))
            gd = obj.grad_dual(x)
            scale = max(1.0, np.abs(gd).max())
            assert np.abs(ga - gd).max() <= 1e-12 * scale, obj.kind


def test_adjoint_central_differences(orc):
    rng = np.random.default_rng(11)
    for obj, draw in _objs(orc, rng):
        x = draw()
        _, g = obj.value_grad_adjoint(x)
        idx = range(obj.dim) if obj.dim < 20 else rng.choice(obj.dim, 12, replace=False)
        for i in idx:
            h = 1e-6 * max(1.0, abs(x[i]))
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            fd = (obj.value(xp) - obj.value(xm)) / (2 * h)
            assert fd == pytest.approx(g[i], rel=2e-6, abs=2e-6 * max(1.0, abs(obj.value(x))))


def test_mlp_and_loglik_against_torch_autograd(orc, golden):
    m = golden["mlp_grad"]
    mlp = orc.Obj(orc.MLP, 321, np.concatenate([m["x"], m["y"]]))
    for c in m["cases"]:
        f, g = mlp.value_grad_adjoint(c["theta"])
        assert f == pytest.approx(c["f"], rel=1e-13)
        assert g == pytest.approx(c["g"], rel=1e-11, abs=1e-13)
    ll = golden["loglik_grad"]
    k = np.array(ll["k"])
    lg = np.array([math.lgamma(v + 1) for v in k])
    obj = orc.Obj(orc.LOGLIK, 4, np.concatenate([ll["x"], k, lg]))
    # the NLL is a sum of terms up to ~k ln(mu) ~ 1e4 that cancel to O(10^2):
    # compare to 1e-13 of the largest term rather than of the total
    big = float(np.max(k * np.abs(np.log(np.maximum(k, 1.0))))) + float(k.sum())
    for c in ll["cases"]:
        f, g = obj.value_grad_adjoint(c["theta"])
        assert abs(f - c["f"]) <= 1e-13 * big
        assert g == pytest.approx(c["g"], rel=1e-10, abs=1e-13 * big)
        assert obj.grad_dual(c["theta"]) == pytest.approx(c["g"], rel=1e-10, abs=1e-13 * big)


# --- BFGS building blocks ------------------------------------------------------
def test_line_search_kat(orc, golden):
    k = golden["reference_kats"]["line_search"]
    a, ne, facc = orc.line_search(orc.Obj(orc.SPHERE, 1), k["f0"], [k["x"]], [k["p"]], [k["g"]])
    assert a == k["alpha"] and ne == 2 and facc == 0.0


def test_line_search_never_returns_zero(orc):
    # an ascent direction: no trial is accepted, alpha ends at 2^-20
    a, ne, facc = orc.line_search(orc.Obj(orc.SPHERE, 1), 1.0, [1.0], [1.0], [2.0])
    assert a == 2.0**-20 and ne == 20 and math.isnan(facc)


def test_norm_identity_direction_kats(orc, golden):
    k = golden["reference_kats"]
    assert orc.norm(k["norm"]["g"]) == k["norm"]["norm"]
    assert list(orc.search_direction(np.eye(3), k["search_direction"]["g"])) == \
        k["search_direction"]["p"]
    H = np.array([[2.0, 1.0], [0.5, 3.0]])  # row-major, j inner: p = -H g
    assert list(orc.search_direction(H, [1.0, 2.0])) == [-4.0, -6.5]


@pytest.mark.parametrize("mode", [0, 1])
def test_bfgs_update_kat(orc, golden, mode):
    k = golden["reference_kats"]["bfgs_update"]
    H = orc.update(np.eye(2), k["s"], k["y"], k["sTy"], mode)
    assert H == pytest.approx(np.array(k["H"]), abs=1e-12)


def test_update_literal_equals_rank2_and_textbook(orc):
    rng = np.random.default_rng(3)
    for n in (2, 3, 5, 8):
        A = rng.standard_normal((n, n))
        H = A @ A.T + n * np.eye(n)
        s, y = rng.standard_normal(n), rng.standard_normal(n)
        sTy = float(s @ y)
        lit = orc.update(H, s, y, sTy, orc.UPDATE_LITERAL)
        r2 = orc.update(H, s, y, sTy, orc.UPDATE_RANK2)
        rho = 1 / sTy
        I = np.eye(n)
        ref = (I - rho * np.outer(s, y)) @ H @ (I - rho * np.outer(y, s)) + rho * np.outer(s, s)
        assert np.abs(lit - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.abs(r2 - lit).max() <= 1e-13 * np.abs(lit).max() * n
        assert np.array_equal(r2, r2.T)  # rank-2 form keeps symmetry bit for bit
        assert r2 @ y == pytest.approx(s, rel=1e-10, abs=1e-10)  # secant condition


def test_update_skipped_for_tiny_curvature(orc):
    H = np.eye(2) * 3
    assert np.array_equal(orc.update(H, [1e-8, 0], [1e-7, 0], 1e-15), H)


def test_argmin_brute_force(orc):
    rng = np.random.default_rng(5)
    for _ in range(20):
        v = rng.integers(0, 5, 40).astype(float)
        v[rng.random(40) < 0.2] = np.nan
        finite = np.where(~np.isnan(v))[0]
        want = finite[np.argmin(v[finite])] if finite.size else 0
        assert orc.argmin(v) == want
    assert orc.argmin([np.nan, np.nan]) == 0


# --- PSO -----------------------------------------------------------------------------
def test_pso_inertia_only_kat(orc, golden):
    k = golden["reference_kats"]["pso_inertia"]
    X = np.array(k["x"], float).reshape(4, 1)
    V = np.array(k["v"], float).reshape(4, 1)
    r = np.random.default_rng(0).random(8)
    X1, V1 = orc.pso_step(k["w"], 0.0, 0.0, r, X, V, X, np.zeros(1))
    assert np.array_equal(V1, V / 2) and np.array_equal(X1, X + V / 2)
    X2, V2 = orc.pso_step(0.0, 0.0, 0.0, r, X, V, X, np.zeros(1))
    assert np.array_equal(V2, np.zeros_like(V)) and np.array_equal(X2, X)


def test_pso_init_and_invariants(orc):
    for obj, D in ((orc.RASTRIGIN, 2), (orc.ROSENBROCK, 3)):
        pbx0, pbv0, gbx0, gbv0 = orc.pso(obj, D, -5, 5, 37, 0, 42)
        f = orc.Obj(obj, D)
        assert all(pbv0[i] == f.value(pbx0[i]) for i in range(37))
        assert gbv0 == pbv0.min() and np.array_equal(gbx0, pbx0[np.argmin(pbv0)])
        assert pbx0.min() > -5 and pbx0.max() <= 5
        # the starting position of coordinate 1 of start 3, in (lo, hi]: the
        # first draw of its XORWOW stream, draw 2 of the start for PHILOX
        # and SOBOL
        assert pbx0[3, 1] == -5 + (5 + 5) * orc.stream_uniforms(orc.PRNG_XORWOW, 42, 0, 3, 1,
                                                                1)[0]
        px = orc.pso(obj, D, -5, 5, 37, 0, 42, prng=orc.PRNG_PHILOX)[0]
        assert px[3, 1] == -5 + (5 + 5) * orc.uniform(42, 0, 3, 2)
        ps = orc.pso(obj, D, -5, 5, 37, 0, 42, prng=orc.PRNG_SOBOL)[0]
        assert ps[3, 1] == -5 + (5 + 5) * orc.sobol_uniform(3, 2)
        prev = pbv0
        for it in (1, 2, 5):
            pbx, pbv, gbx, gbv = orc.pso(obj, D, -5, 5, 37, it, 42)
            assert np.all(pbv <= prev)
            assert gbv == pbv.min()
            assert all(pbv[i] == f.value(pbx[i]) for i in range(37))
            prev = pbv


def _pso_numpy(orc, prng, N, D, lo, hi, T, seed, f, w=0.5, c1=1.2, c2=1.5, vfrac=0.1,
               swap_r=False):
    """The reference PSO rule restated in numpy from include/pso.cuh: init
    per coordinate a position U(lo, hi] then a velocity U(-vr, vr], vr =
    0.1 (hi - lo) (pso.cuh:39-44); iteration t draws r1 then r2 per
    coordinate (pso.cuh:105-106), v' = w v + c1 r1 (pb - x) + c2 r2 (gb - x),
    x' = x + v' with w = 0.5, c1 = 1.2, c2 = 1.5 (pso.cuh:113-118, 302);
    pBest on strict improvement; gBest = the lowest-index minimum of pBest
    when it is strictly below the old gBest.  The uniforms come from the
    generator streams alone (orc.stream_uniforms)."""
    K = 2 * (T + 1)
    U = np.zeros((N, D, K))
    for i in range(N):
        if prng == orc.PRNG_XORWOW:  # one stream per (start, coordinate)
            for d in range(D):
                U[i, d] = orc.stream_uniforms(prng, seed, 0, i, K, coord=d)
        else:  # one stream per start; coordinate d of iteration t at 2Dt + 2d
            s = orc.stream_uniforms(prng, seed, 0, i, K * D)
            for t in range(T + 1):
                for d in range(D):
                    U[i, d, 2 * t:2 * t + 2] = s[2 * D * t + 2 * d:2 * D * t + 2 * d + 2]
    vr = (hi - lo) * vfrac
    X = lo + (hi + (-lo)) * U[:, :, 0]
    V = -vr + (vr + vr) * U[:, :, 1]
    pbx, pbv = X.copy(), np.array([f(x) for x in X])
    gbv, gbx = np.inf, np.zeros(D)
    m = int(np.argmin(pbv))
    if pbv[m] < gbv:
        gbv, gbx = pbv[m], pbx[m].copy()
    for t in range(1, T + 1):
        r1, r2 = U[:, :, 2 * t], U[:, :, 2 * t + 1]
        if swap_r:
            r1, r2 = r2, r1
        V = w * V + c1 * r1 * (pbx - X) + c2 * r2 * (gbx[None, :] - X)
        X = X + V
        fx = np.array([f(x) for x in X])
        better = fx < pbv
        pbv = np.where(better, fx, pbv)
        pbx = np.where(better[:, None], X, pbx)
        m = int(np.argmin(pbv))
        if pbv[m] < gbv:
            gbv, gbx = pbv[m], pbx[m].copy()
    return dict(x=X, v=V, pbest_x=pbx, pbest_val=pbv, gbest_x=gbx, gbest_val=gbv)


@pytest.mark.parametrize("prng", ["xorwow", "philox"])
def test_pso_rule_restated_from_streams(orc, prng):
    """The oracle's PSO equals the reference rule written out in numpy from
    the generator streams, bit for bit, on init and two iterations (N = 5,
    D = 3, an asymmetric box so that a velocity range of 0.1 hi or 0.1 (hi +
    lo) differs from 0.1 (hi - lo)).  A swap of c1 and c2, of r1 and r2, or
    another velocity range gives another swarm, so each of those mistakes in
    the oracle would fail this test."""
    g = orc.PRNG_XORWOW if prng == "xorwow" else orc.PRNG_PHILOX
    N, D, lo, hi, T, seed = 5, 3, -3.0, 7.0, 2, 42
    obj = orc.Obj(orc.ROSENBROCK, D)
    got = orc.pso_swarm(orc.ROSENBROCK, D, lo, hi, N, T, seed, prng=g)
    want = _pso_numpy(orc, g, N, D, lo, hi, T, seed, obj.value)
    for k in ("x", "v", "pbest_x", "pbest_val", "gbest_x"):
        assert np.array_equal(got[k], want[k]), k
    assert got["gbest_val"] == want["gbest_val"]
    # T = 0: the initial velocities lie in (-0.1 R, 0.1 R]
    init = orc.pso_swarm(orc.ROSENBROCK, D, lo, hi, N, 0, seed, prng=g)
    R = hi - lo
    assert np.all(init["v"] > -0.1 * R) and np.all(init["v"] <= 0.1 * R)
    assert np.all(init["x"] > lo) and np.all(init["x"] <= hi)
    assert np.array_equal(init["x"], init["pbest_x"])
    # the comparison above has the power to see each plausible mistake
    for bad in (dict(c1=1.5, c2=1.2), dict(swap_r=True), dict(vfrac=0.1 * hi / R)):
        alt = _pso_numpy(orc, g, N, D, lo, hi, T, seed, obj.value, **bad)
        assert not np.array_equal(alt["x"], got["x"]), bad


# --- whole runs -----------------------------------------------------------------------
def test_bfgs_reference_launch_kats(orc):
    # test/bfgs_tests.cu:266-329 -- Rastrigin from its minimum
    r = orc.run(orc.RASTRIGIN, 2, -5, 5, 1, 10, 0, 1, 1e-6, 123, x0=[0, 0], grad=orc.GRAD_DUAL)
    assert r.best["status"] == 1 and r.best["iter"] == 0 and r.best["fval"] == 0.0
    # test/bfgs_tests.cu:340-389 -- Quad from (4.5, 4.5)
    r = orc.run(orc.SPHERE, 2, -10, 10, 1, 100, 0, 1, 1e-6, 123, x0=[4.5, 4.5],
                grad=orc.GRAD_DUAL)
    assert r.best["status"] == 1 and abs(r.best["fval"]) < 1e-6
    assert np.abs(r.best_x).max() < 1e-6
    # test/bfgs_tests.cu:391-453 -- Rosenbrock from (4.5, 4.5) at tol 1e-17
    for grad in (orc.GRAD_DUAL, orc.GRAD_ADJOINT):
        for mode in (orc.UPDATE_LITERAL, orc.UPDATE_RANK2):
            r = orc.run(orc.ROSENBROCK, 2, -5, 5, 1, 1000, 0, 1, 1e-17, 123, x0=[4.5, 4.5],
                        grad=grad, update_mode=mode)
            assert r.best["status"] == 1 and r.best["iter"] < 1000
            assert abs(r.best["fval"]) < 1e-4
            assert np.abs(r.best_x - 1.0).max() < 1e-6
    # test/bfgs_tests.cu:554-671 -- sum x^2 from (1, 1): converges before slot 49
    r = orc.run(orc.SPHERE, 2, -5, 5, 1, 50, 0, 1, 1e-6, 42, x0=[1, 1])
    assert r.best["status"] == 1 and r.best["iter"] < 49


def test_status_codes(orc):
    # 0: surrendered at a tiny MAX_ITER (iter == MAX_ITER)
    r = orc.run(orc.ROSENBROCK, 2, -5, 5, 4, 2, 0, 10, 1e-12, 42)
    assert set(r.status) == {0} and set(r.iter) == {2}
    # MAX_ITER = 0: no iteration, status 0 at iter 0
    r = orc.run(orc.ROSENBROCK, 2, -5, 5, 3, 0, 0, 10, 1e-12, 42)
    assert set(r.status) == {0} and set(r.iter) == {0}
    # 5: an objective that is +inf at the start (NLL with a = 0)
    kx = np.linspace(0.1, 0.5, 5)
    kk = np.array([5.0, 3.0, 2.0, 1.0, 1.0])
    lg = np.array([math.lgamma(v + 1) for v in kk])
    r = orc.run(orc.LOGLIK, 4, 0, 10, 1, 50, 0, 1, 1e-8, 1,
                params=np.concatenate([kx, kk, lg]), x0=[0.0, 10.0, 5.0, 0.1])
    assert r.best["status"] == 5
    # -1 / defaults for N == 0 (zeus.cuh:58-60)
    r = orc.run(orc.SPHERE, 2, -5, 5, 0, 10, 0, 1, 1e-6, 1)
    assert r.best["status"] == -1 and r.best["fval"] == 37.0 and r.best["idx"] == -1


def test_early_stop_semantics(orc):
    n = 48
    full = orc.run(orc.SPHERE, 3, -5, 5, n, 100, 2, n + 1, 1e-10, 9, threads=1)
    assert set(full.status) == {1}
    r = orc.run(orc.SPHERE, 3, -5, 5, n, 100, 2, 5, 1e-10, 9, threads=1)
    # serial order: starts 0..4 converge, the flag stops every later start at iter 0
    assert list(r.status[:5]) == [1] * 5 and set(r.status[5:]) == {2}
    assert set(r.iter[5:]) == {0}
    assert r.best["claimed"] == 5 and r.best["stopped"] == n - 5
    # converged starts are the ones of the unstopped run, bit for bit
    assert np.array_equal(r.fval[:5], full.fval[:5]) and np.array_equal(r.x[:5], full.x[:5])


def test_quadratics_and_rosenbrock_reach_their_minima(orc):
    D = 20
    C = np.full((D, D), 0.5)
    np.fill_diagonal(C, 1 + (D - 1) * 0.5)
    r = orc.run(orc.GAUSSIAN, D, -5, 5, 16, 1000, 3, 17, 1e-8, 42, params=C)
    assert set(r.status) == {1} and np.abs(r.x).max() < 1e-8
    r = orc.run(orc.ROSENBROCK, 2, -5, 5, 64, 10000, 10, 65, 1e-8, 42)
    conv = r.status == 1
    assert conv.sum() >= 60
    assert np.abs(r.x[conv] - 1.0).max() < 1e-6
