"""The selectable generators (reference PRNGType, include/traits.hpp:67-71):
XORWOW against Marsaglia's published recurrence, its per-(start,
coordinate) seeding against the Philox blocks it is defined from, SOBOL
against the closed form of its first dimension; then the GPU streams and
whole runs against the oracle."""
import numpy as np
import pytest

import zeus_amd as z

M32 = 0xFFFFFFFF


def _marsaglia_xorwow(state, n):
    """The xorwow routine of Marsaglia, 'Xorshift RNGs' (2003), in Python."""
    x, y, zz, w, v, d = state
    out = []
    for _ in range(n):
        t = (x ^ (x >> 2)) & M32
        x, y, zz, w = y, zz, w, v
        v = ((v ^ ((v << 4) & M32)) ^ (t ^ ((t << 1) & M32))) & M32
        d = (d + 362437) & M32
        out.append((v + d) & M32)
    return out


def test_xorwow_is_marsaglias_recurrence(orc):
    # the paper's initial values
    st = [123456789, 362436069, 521288629, 88675123, 5783321, 6615241]
    want = _marsaglia_xorwow(st, 2000)
    s = list(st)
    got = [orc.xorwow_next(s) for _ in range(2000)]
    assert got == want


def test_xorwow_seeding_and_conversion(orc):
    for seed, stream, start, coord in ((42, 0, 0, 0), (7, 3, 12345, 149),
                                       (0xDEADBEEF, 9, 2**40 + 3, 320)):
        st = orc.xorwow_seed(seed, stream, start, coord)
        lo, hi = start & M32, start >> 32
        a = orc.philox([2 * coord, M32, lo, hi], [seed, stream])
        b = orc.philox([2 * coord + 1, M32, lo, hi], [seed, stream])
        assert st == a + b[:2]
        # a uniform is two outputs: top 32 + next 21 bits, (m + 1) 2^-53
        u = orc.stream_uniforms(orc.PRNG_XORWOW, seed, stream, start, 50, coord)
        raw = _marsaglia_xorwow(st, 100)
        for k in range(50):
            m = (raw[2 * k] << 21) | (raw[2 * k + 1] >> 11)
            assert u[k] == (m + 1) * 2.0**-53


def test_xorwow_statistics(orc):
    u = np.concatenate([orc.stream_uniforms(orc.PRNG_XORWOW, 42, 0, s, 2000, c)
                        for s in range(20) for c in range(5)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.005
    h = np.histogram(u, bins=20, range=(0, 1))[0]
    chi2 = ((h - u.size / 20) ** 2 / (u.size / 20)).sum()
    assert chi2 < 45.0  # 19 degrees of freedom, p ~ 1e-3
    # distinct starts and coordinates get distinct streams
    a = orc.stream_uniforms(orc.PRNG_XORWOW, 42, 0, 5, 100, 0)
    b = orc.stream_uniforms(orc.PRNG_XORWOW, 42, 0, 6, 100, 0)
    c = orc.stream_uniforms(orc.PRNG_XORWOW, 42, 0, 5, 100, 1)
    assert not np.any(a == b) and not np.any(a == c)


def test_sobol_first_dimension(orc):
    # Gray-code order: points 0, 1, 2, 3 -> 0, 1/2, 3/4, 1/4 (plus 2^-32)
    e = 2.0**-32
    assert [orc.sobol_uniform(0, k) for k in range(4)] == [e, 0.5 + e, 0.75 + e, 0.25 + e]
    # the first 2^k points are the multiples of 2^-k
    for k in (4, 10):
        u = np.sort(orc.stream_uniforms(orc.PRNG_SOBOL, 0, 0, 0, 2**k))
        assert np.array_equal(u, np.arange(2**k) * 2.0**-k + e)
    # start i reads the sequence from point i on (streams overlap by design)
    a = orc.stream_uniforms(orc.PRNG_SOBOL, 1, 2, 5, 40)
    b = orc.stream_uniforms(orc.PRNG_SOBOL, 99, 0, 6, 39)
    assert np.array_equal(a[1:], b)


def test_philox_stream_is_the_counter_draws(orc):
    u = orc.stream_uniforms(orc.PRNG_PHILOX, 42, 3, 17, 30)
    assert all(u[k] == orc.uniform(42, 3, 17, k) for k in range(30))


def test_oracle_runs_with_each_generator(orc):
    outs = []
    for g in (orc.PRNG_XORWOW, orc.PRNG_PHILOX, orc.PRNG_SOBOL):
        for pso in (3, -1):
            r = orc.run(orc.ROSENBROCK, 2, -5, 5, 32, 2000, pso, 33, 1e-8, 42, prng=g)
            assert r.best["status"] == 1 and abs(r.best["fval"]) < 1e-12
            outs.append(r.x.copy())
    for i in range(len(outs)):
        for j in range(i + 1, len(outs)):
            assert not np.array_equal(outs[i], outs[j])


# ---------------------------------------------------------------- GPU -------
GENS = [z.PRNGType.XORWOW, z.PRNGType.PHILOX, z.PRNGType.SOBOL]


@pytest.mark.gpu
def test_device_streams_match_oracle(orc):
    for g in GENS:
        for start, coord in ((0, 0), (77, 3), (2**33 + 5, 149)):
            got = z.stream_uniforms(g, 42, 5, start, 3000, coord)
            want = orc.stream_uniforms(int(g), 42, 5, start, 3000, coord)
            assert np.array_equal(got, want), (g, start)


@pytest.mark.gpu
def test_pso_lane_form_each_generator_bitwise(orc):
    for g in GENS:
        got = z.pso(z.Rosenbrock(4), -5.0, 5.0, 300, 6, 42, prng=g)
        want = orc.pso(orc.ROSENBROCK, 4, -5.0, 5.0, 300, 6, 42, prng=int(g))
        for a, b in zip(got, want):
            assert np.array_equal(np.asarray(a), np.asarray(b)), g


@pytest.mark.gpu
def test_thread_form_each_generator_bitwise(orc):
    N = 192
    for g in GENS:
        for pso in (4, -1):
            r = z.Zeus(z.Rosenbrock(2), -5.0, 5.0, N, 3000, pso, N + 1, "rosenbrock", 1e-8, 42,
                       0, parallel=False, prng=g, per_start=True)
            o = orc.run(orc.ROSENBROCK, 2, -5.0, 5.0, N, 3000, pso, N + 1, 1e-8, 42,
                        prng=int(g))
            assert np.array_equal(r.per_start["x"], o.x), (g, pso)
            assert np.array_equal(r.per_start["status"], o.status)


@pytest.mark.gpu
def test_mfma_pso_and_tile_random_starts_each_generator(orc):
    D, N = 32, 64
    rng = np.random.default_rng(3)
    A = rng.normal(size=(D, D))
    C = A @ A.T / D + np.eye(D)
    for g in GENS:
        for pso in (3, -1):
            r = z.Zeus(z.Gaussian(C), -5.0, 5.0, N, 2000, pso, N + 1, "gaussian", 1e-8, 42, 0,
                       form="tile", prng=g, per_start=True)
            o = orc.run(orc.GAUSSIAN, D, -5.0, 5.0, N, 2000, pso, N + 1, 1e-8, 42,
                        params=C.reshape(-1), prng=int(g))
            assert np.array_equal(r.per_start["status"], o.status), (g, pso)
            assert np.abs(r.per_start["x"] - o.x).max() < 1e-6
