"""Reporting of a call: timing metrics (reference report_metrics_and_cleanup,
bfgs_sequential.cuh:240-382), the distance to the known minimum
(src/util.cpp:11-41), the per-dimension results table (include/util.hpp:
78-149) and the example drivers' command-line contract
(examples/optimization_utils.cc:10-124)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import zeus_amd as z
from zeus_amd import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "zeus_amd", "bin", "zeus_optimize")
N_COLS = 26  # results-table columns before the coordinates


def _metrics(ticks, calls, total, tpb, ms, hz):
    m = capi.Metrics()
    t = np.ascontiguousarray(ticks, np.uint64)
    c = np.ascontiguousarray(calls, np.int32)
    tt = None if total is None else np.ascontiguousarray(total, np.uint64)
    rc = capi.lib().zeus_metrics_from_ticks(t.ctypes.data, c.ctypes.data,
                                            None if tt is None else tt.ctypes.data, t.size,
                                            tpb, ms, hz, C.byref(m))
    assert rc == capi.OK
    return m


def test_metrics_closed_forms():
    # every start: 1000 ticks in 4 calls, 4000 ticks in total, 1 MHz clock,
    # kernel 2 ms; 600 starts in blocks of 256 -> blocks of 256, 256, 88
    n, hz, ms = 600, 1e6, 2.0
    m = _metrics(np.full(n, 1000), np.full(n, 4), np.full(n, 4000), 256, ms, hz)
    assert m.calls_per_thread_mean == 4.0
    assert m.ms_per_call == pytest.approx(250 / hz * 1e3)
    assert m.fraction_of_kernel == pytest.approx((1000 / hz * 1e3) / ms)
    assert m.serialized == pytest.approx((n * 1000 / hz * 1e3) / ms)
    assert m.fraction_of_thread == pytest.approx(0.25)
    # p95 of 3 blocks is the largest: 256 * 1 ms / 2 ms
    assert m.block95 == pytest.approx(256 * 1.0 / ms)


def test_metrics_p95_index_and_guards():
    # 40 blocks of one start with 1..40 ms: ceil(0.95 * 40) - 1 = 37 -> 38 ms
    hz = 1e3
    m = _metrics(np.arange(1, 41), np.ones(40), None, 1, 10.0, hz)
    assert m.block95 == pytest.approx(38e3 / hz / 10.0)
    assert m.fraction_of_thread == 0.0  # no totals given
    z0 = _metrics(np.zeros(5), np.zeros(5), np.ones(5), 256, 0.0, hz)  # zero kernel time
    assert (z0.ms_per_call, z0.fraction_of_kernel, z0.serialized, z0.block95) == (0, 0, 0, 0)
    bad = capi.Metrics()
    assert capi.lib().zeus_metrics_from_ticks(None, None, None, 3, 256, 1.0, 1e6,
                                              C.byref(bad)) == capi.E_ARG


def test_euclidean_error():
    assert z.euclidean_error("rosenbrock", [1.0, 1.0, 1.0]) == 0.0
    assert z.euclidean_error("rosenbrock", [1.0, 3.0]) == 2.0
    assert z.euclidean_error("goldstein_price", [0.0, -1.0]) == 0.0
    assert math.isnan(z.euclidean_error("goldstein_price", [0.0, 0.0, 0.0]))
    assert z.euclidean_error("ackley", [3.0, 4.0]) == 5.0


def _row(status=1, fval=1.5e-20, coords=(1.0, 0.999)):
    return z.Result(idx=7, status=status, fval=fval, gradientNorm=3e-9,
                    coordinates=np.array(coords), iter=12,
                    c=z.Convergence(actual=90, claimed=100, surrendered=3, stopped=921),
                    ms_opt=1.25, ms_pso=0.5,
                    ad=dict(ms_per_call=0.001, calls_per_thread_mean=20.5,
                            fraction_of_kernel=0.1, block95=0.75, serialized=12.0,
                            fraction_of_thread=0.3))


def test_results_table_format(tmp_path):
    path = str(tmp_path / "zeus_2d_results.tsv")
    z.append_results_tsv(_row(), 2, 1024, "rosenbrock", 10000, 10, 3, path)
    z.append_results_tsv(_row(status=0), 2, 1024, "rosenbrock", 10000, 0, 4, path)
    lines = open(path).read().splitlines()
    assert len(lines) == 3  # header once
    head = lines[0].split("\t")
    assert head[:4] == ["fun", "run", "N", "kernel"]
    assert head[N_COLS - 4:N_COLS] == ["time", "error", "fval", "norm"]
    assert head[N_COLS:] == ["coord_0", "coord_1"]
    r = lines[1].split("\t")
    assert len(r) == N_COLS + 2
    assert r[:4] == ["rosenbrock", "3", "1024", "1.25"]
    assert r[4:9] == ["0.001", "20.5", "0.3", "0.75", "12"]  # ad_fraction is per-thread
    assert r[9:14] == ["0", "0", "0", "0", "0"]
    assert r[14:22] == ["100", "90", "3", "921", "7", "1", "10000", "10"]
    assert r[22] == "1.75"                       # pso_iter > 0: pso + opt
    assert r[23] == f"{abs(0.999 - 1.0):e}"
    assert r[24:26] == ["1.500000e-20", "3.000000e-09"]
    assert r[26:] == ["1.000000e+00", "9.990000e-01"]
    r2 = lines[2].split("\t")
    assert r2[22] == "1.25" and r2[21] == "0"    # pso_iter 0: opt only


def test_driver_argument_errors():
    if not os.path.exists(DRIVER):
        pytest.fail("zeus_amd/bin/zeus_optimize is missing: run python -m zeus_amd.build")
    p = subprocess.run([DRIVER, "rosenbrock", "-5", "5"], capture_output=True, text=True)
    assert p.returncode == 1 and "Usage:" in p.stderr
    args = ["-5", "5", "100", "5", "10", "64", "1e-8", "42", "0"]
    p = subprocess.run([DRIVER, "rosenbrock", *args, "--bogus"], capture_output=True, text=True)
    assert p.returncode == 1 and "Unknown or incomplete argument: --bogus" in p.stderr
    p = subprocess.run([DRIVER, "nosuchfunction", *args], capture_output=True, text=True)
    assert p.returncode == 1 and "Usage:" in p.stderr
    link = os.path.join(ROOT, "zeus_amd", "bin", "optimize_himmelblau")
    p = subprocess.run([link, "1"], capture_output=True, text=True)
    assert p.returncode == 1 and "Usage:" in p.stderr


# ---------------------------------------------------------------- GPU -------
@pytest.mark.gpu
def test_sequential_form_metrics():
    N = 256
    r = z.Zeus(z.Rosenbrock(2), -5.0, 5.0, N, 1000, 5, N + 1, "rosenbrock", 1e-8, 42, 0,
               parallel=False, per_start=True)
    g = r.per_start["grads"]
    assert r.ad["calls_per_thread_mean"] == pytest.approx(g.mean(), rel=1e-12)
    assert r.bfgs["calls_per_thread_mean"] == pytest.approx(g.mean() - 1.0, rel=1e-12)
    for m in (r.ad, r.bfgs):
        assert m["ms_per_call"] > 0 and 0 < m["fraction_of_thread"] < 1
        assert m["serialized"] > 0 and m["block95"] > 0
    rp = z.Zeus(z.Rosenbrock(2), -5.0, 5.0, N, 1000, 5, N + 1, "rosenbrock", 1e-8, 42, 0,
                parallel=True)
    assert all(v == 0.0 for v in rp.ad.values()) and all(v == 0.0 for v in rp.bfgs.values())


@pytest.mark.gpu
def test_driver_matches_oracle(orc, tmp_path):
    """optimize_rosenbrock with the reference's arguments (per-thread form,
    dual gradient, no early stop) reports the oracle's best start and appends
    its row to zeus_2d_results.tsv in the working directory."""
    N = 256
    args = ["-5", "5", "10000", "10", str(N + 1), str(N), "1e-8", "42", "7"]
    link = os.path.join(ROOT, "zeus_amd", "bin", "optimize_rosenbrock")
    p = subprocess.run([link, *args], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    out = p.stdout
    assert "=== Rosenbrock Function Optimization (2D) ===" in out
    assert "Algorithm: sequential" in out and "[AD] Calls/thread" in out
    o = orc.run(orc.ROSENBROCK, 2, -5, 5, N, 10000, 10, N + 1, 1e-8, 42, grad=orc.GRAD_DUAL,
                early_stop=False)
    b = o.best
    assert f"Status: {b['status']} (Converged)" in out
    assert f"Function value: {b['fval']:e}" in out
    assert f"Iterations: {b['iter']}" in out
    rows = open(tmp_path / "zeus_2d_results.tsv").read().splitlines()
    assert len(rows) == 2
    r = rows[1].split("\t")
    assert r[0] == "rosenbrock" and r[1] == "7" and r[2] == str(N)
    assert int(r[18]) == b["idx"] and int(r[19]) == b["status"]
    assert int(r[14]) == b["claimed"]
    assert float(r[4]) > 0.0  # AD ms per call from the per-thread timing


@pytest.mark.gpu
def test_driver_functions_converge(tmp_path):
    args = ["-5", "5", "2000", "10", "20", "512", "1e-8", "42", "0"]
    want = {"goldstein_price": (3.0, (0.0, -1.0)), "himmelblau": (0.0, None),
            "rastrigin": (0.0, (0.0, 0.0)), "ackley": (0.0, (0.0, 0.0))}
    for fn, (fstar, xstar) in want.items():
        p = subprocess.run([DRIVER, fn, *args, "--parallel"], capture_output=True, text=True,
                           cwd=tmp_path, timeout=120)
        assert p.returncode == 0, p.stderr
        line = [l for l in p.stdout.splitlines() if l.startswith("Function value:")][0]
        assert float(line.split(":")[1]) == pytest.approx(fstar, abs=1e-6), fn
        assert "Algorithm: parallel" in p.stdout
    rows = open(tmp_path / "zeus_2d_results.tsv").read().splitlines()
    assert len(rows) == 1 + len(want)
    gp = rows[1].split("\t")
    assert gp[0] == "goldstein_price" and float(gp[23]) < 1e-6  # error vs (0, -1)
