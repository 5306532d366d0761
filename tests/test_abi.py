"""The C ABI library without a GPU: it loads, exports what include/zeus_c.h
declares, and its host-side functions and argument checks behave."""
import math
import os
import re

import numpy as np
import pytest

import zeus_amd as z
from zeus_amd import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "zeus_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(zeus_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.SIGNATURES), "ctypes table out of sync with zeus_c.h"


def test_philox_matches_oracle_bitwise(orc, golden):
    rng = np.random.default_rng(0)
    for _ in range(200):
        seed, stream = (int(v) for v in rng.integers(0, 2**32, 2))
        start, draw = int(rng.integers(0, 2**41)), int(rng.integers(0, 2**40))
        assert z.philox_uniform(seed, stream, start, draw) == orc.uniform(seed, stream, start,
                                                                          draw)


def test_status_strings():
    assert capi.status_string(0) == "surrendered"
    assert capi.status_string(1) == "converged"
    assert capi.status_string(2) == "stopped early"
    assert capi.status_string(5) == "non-finite value"
    assert capi.status_string(capi.E_ARG) == "invalid argument"


def test_argument_validation_needs_no_device():
    bad = [
        z.Problem(z.Rosenbrock(0), -5, 5, 4, 10, 1, 1, 1e-8, 1),           # dim 0
        z.Problem(z.Rosenbrock(2), -5, 5, -1, 10, 1, 1, 1e-8, 1),          # N < 0
        z.Problem(z.Rosenbrock(2), -5, 5, 4, -3, 1, 1, 1e-8, 1),           # MAX_ITER < 0
    ]
    for pr in bad:
        with pytest.raises(capi.ZeusError) as e:
            z.run_batch([pr])
        assert e.value.code == capi.E_ARG
    with pytest.raises(capi.ZeusError) as e:  # dimension not instantiated
        z.run_batch([z.Problem(z.Rosenbrock(7), -5, 5, 4, 10, 1, 1, 1e-8, 1)])
    assert e.value.code == capi.E_UNSUPPORTED


def test_zero_starts_returns_default_result():
    r = z.Zeus(z.Rosenbrock(2), -5, 5, 0, 100, 10, 10, "rosenbrock", 1e-8, 42, 0)
    assert (r.idx, r.status, r.fval, r.iter) == (-1, -1, 37.0, 0)
    assert list(r.coordinates) == [0.0, 0.0]


def test_trajectory_argument_validation_needs_no_device():
    pr = z.Problem(z.Rosenbrock(2), -5, 5, 4, 10, 1, 1, 1e-8, 1, trajectory=(-1, 4))
    with pytest.raises(ValueError):
        z.run_batch([pr])


def test_dijet_generator():
    a = z.make_dijet(42, 3)
    b = z.make_dijet(42, 3)
    c = z.make_dijet(42, 4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    x, k, lg = a[:100], a[100:200], a[200:]
    edges = 1000 + 60 * np.arange(101)
    assert x == pytest.approx((edges[:-1] + 30) / 13000, rel=1e-15)
    assert np.all(k >= 0) and np.all(k == np.round(k))
    assert abs(k.sum() - 1e6) < 5 * math.sqrt(1e6)
    assert all(lg[i] == math.lgamma(k[i] + 1) or abs(lg[i] - math.lgamma(k[i] + 1)) < 1e-9
               for i in range(100))
    # per-bin Poisson: standardized residuals are O(1)
    from zeus_amd.api import dijet_p0
    p0 = dijet_p0()
    mu = p0 * (1 - x) ** 10 / x ** (5 + 0.1 * np.log(x))
    big = mu > 50
    zres = (k[big] - mu[big]) / np.sqrt(mu[big])
    assert np.abs(zres).max() < 5 and abs(zres.mean()) < 1


def test_mlp_and_spd_generators():
    d = z.make_mlp_data(42, 256)
    x, y = d[:256], d[256:]
    assert x.min() > -1 and x.max() <= 1
    res = y - np.sin(3 * x)
    assert 0.07 < res.std() < 0.13 and abs(res.mean()) < 0.03
    C = z.make_spd(42, 5, 12)
    assert np.array_equal(C, C.T)
    assert np.linalg.eigvalsh(C).min() > 1.0 - 1e-12  # A A^T / D + I >= I
