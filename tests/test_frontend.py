"""The C++ front end (zeus/zeus.hpp) driven by a compiled program,
tests/cpp/frontend_test.hip, that defines objectives the way reference users
do (examples/rosenbrock.hpp, examples/gaussian.hpp) and calls zeus::Zeus.

CPU: the program builds, links the library and passes its host-only checks
(Matrix / NonNull / HipBuffer / Result contracts of test/matrix_tests.cu,
test/non_null_tests.cu, test/cuda_buffer_tests.cu, util.hpp:36-54, and the
objective concept of test/bad_objectives_concepts.cu as static_asserts).
GPU: its zeus::Zeus results against the oracle on the same seeds: bit-exact
for the per-thread form (same IEEE operations in the same order), within
stated tolerances for the per-workgroup form.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bin", "frontend_test")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/bin/frontend_test is missing: run python -m zeus_amd.build")
    return EXE


def test_frontend_host_only():
    p = subprocess.run([_exe(), "--host-only"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failure(s)" in p.stdout


@pytest.fixture(scope="module")
def gpu_cases(tmp_path_factory):
    wd = tmp_path_factory.mktemp("frontend")  # zeus::Zeus appends its results table here
    traj = str(wd / "traj.tsv")
    env = dict(os.environ, ZEUS_FRONTEND_TRAJ=traj)
    p = subprocess.run([_exe()], capture_output=True, text=True, timeout=300, env=env, cwd=wd)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    cases = {}
    for line in p.stdout.splitlines():
        if line.startswith("{"):
            d = json.loads(line)
            cases[d["case"]] = d
    assert cases["checks"]["failures"] == 0, p.stderr
    cases["traj_path"] = traj
    cases["workdir"] = str(wd)
    return cases


def _counts(r):
    st = r.status
    return dict(claimed=int((st == 1).sum()), stopped=int((st == 2).sum()),
                surrendered=int(((st != 1) & (st != 2)).sum()))


@pytest.mark.gpu
def test_frontend_rosenbrock_thread_bitwise(orc, gpu_cases):
    """C1 workload (1024 starts, PSO 10, tol 1e-8, seed 42), dual gradient, no
    early stop: the per-thread form repeats the oracle exactly."""
    got = gpu_cases["rosenbrock2_thread_dual"]
    want = orc.run(orc.ROSENBROCK, 2, -5, 5, 1024, 10000, 10, 1025, 1e-8, 42,
                   grad=orc.GRAD_DUAL, early_stop=False)
    b = want.best
    assert (got["idx"], got["status"], got["iter"]) == (b["idx"], b["status"], b["iter"])
    assert got["fval"] == b["fval"]
    assert np.array_equal(np.array(got["x"]), want.best_x)
    c = _counts(want)
    assert (got["claimed"], got["stopped"], got["surrendered"]) == (
        c["claimed"], c["stopped"], c["surrendered"])
    assert got["actual"] == int(((want.status == 1)
                                 & (np.linalg.norm(want.x - 1.0, axis=1) < 0.5)).sum())


@pytest.mark.gpu
def test_frontend_rosenbrock_group(orc, gpu_cases):
    """Per-workgroup form: reductions run in another order, so each start may
    take a slightly different path; the best point and the counts agree to
    within the stated tolerances."""
    got = gpu_cases["rosenbrock2_group_dual"]
    want = orc.run(orc.ROSENBROCK, 2, -5, 5, 256, 10000, 10, 257, 1e-8, 42,
                   grad=orc.GRAD_DUAL, early_stop=False)
    assert got["status"] == 1
    assert got["fval"] <= 1e-12
    assert np.abs(np.array(got["x"]) - 1.0).max() <= 1e-5
    c = _counts(want)
    assert abs(got["claimed"] - c["claimed"]) <= 256 // 50   # <= 2 % of the starts
    assert got["claimed"] + got["stopped"] + got["surrendered"] == 256


@pytest.mark.gpu
def test_frontend_gaussian_matrix_member(orc, gpu_cases):
    """20-D Gaussian whose functor holds a Matrix: every start converges to
    the unique minimum x = 0, f = 0, as in the oracle."""
    D = 20
    C = np.full((D, D), 0.5)
    np.fill_diagonal(C, 1.0 + (D - 1) * 0.5)
    got = gpu_cases["gaussian20_group_dual"]
    want = orc.run(orc.GAUSSIAN, D, -5, 5, 64, 10000, 10, 65, 1e-8, 42, params=C.reshape(-1),
                   grad=orc.GRAD_DUAL, early_stop=False)
    assert got["status"] == want.best["status"] == 1
    assert got["claimed"] == _counts(want)["claimed"] == 64
    assert abs(got["fval"] - want.best["fval"]) <= 1e-12
    assert np.abs(np.array(got["x"])).max() <= 1e-6


@pytest.mark.gpu
def test_frontend_builtin_gaussian150(orc, gpu_cases):
    """zeus::Zeus(zeus::Gaussian<150>{C}, ...) -- the reference's example
    Gaussian as a built-in -- runs the library's kernels (the tuned form of
    the class is a tile form) and agrees with the oracle: 64 starts without
    early stop converge to the minimum; the headline shape (4096 starts, 100
    required, Philox) reaches its required count with the minimum."""
    D = 150
    C = np.full((D, D), 0.5)
    np.fill_diagonal(C, 1.0 + (D - 1) * 0.5)
    assert gpu_cases["gaussian150_builtin_form"]["form"] == 3  # ZEUS_FORM_TILE
    got = gpu_cases["gaussian150_builtin"]
    want = orc.run(orc.GAUSSIAN, D, -5, 5, 64, 10000, 10, 65, 1e-8, 42, params=C.reshape(-1),
                   early_stop=False)
    assert got["status"] == want.best["status"] == 1
    assert got["claimed"] == _counts(want)["claimed"] == 64
    assert abs(got["fval"] - want.best["fval"]) <= 1e-8
    assert np.abs(np.array(got["x"]) - want.best_x).max() <= 1e-6
    h = gpu_cases["gaussian150_builtin_headline"]
    assert h["status"] == 1 and h["claimed"] >= 100 and abs(h["fval"]) <= 1e-8
    assert h["claimed"] + h["stopped"] == 4096
    assert np.abs(np.array(h["x"])).max() <= 1e-6


@pytest.mark.gpu
def test_frontend_builtin_rosenbrock10(orc, gpu_cases):
    """zeus::obj::Rosenbrock<10> is a built-in: the library's autotuned
    kernels give the oracle's best start (status, value, minimizer)."""
    got = gpu_cases["rosenbrock10_builtin"]
    want = orc.run(orc.ROSENBROCK, 10, -5, 5, 256, 10000, 10, 257, 1e-8, 42, early_stop=False)
    assert got["status"] == want.best["status"] == 1
    assert abs(got["fval"] - want.best["fval"]) <= 1e-8
    assert np.abs(np.array(got["x"]) - want.best_x).max() <= 1e-6
    assert abs(got["claimed"] - _counts(want)["claimed"]) <= 256 // 50


@pytest.mark.gpu
def test_frontend_adjoint_functor_bitwise(orc, gpu_cases):
    """A functor with value_and_gradient runs the adjoint path; its arithmetic
    is the oracle's sphere adjoint, so the per-thread result is bit-exact."""
    got = gpu_cases["sphere3_thread_adjoint"]
    want = orc.run(orc.SPHERE, 3, -5, 5, 128, 1000, 5, 129, 1e-10, 7,
                   grad=orc.GRAD_ADJOINT, early_stop=False)
    b = want.best
    assert (got["idx"], got["status"], got["iter"]) == (b["idx"], b["status"], b["iter"])
    assert got["fval"] == b["fval"]
    assert np.array_equal(np.array(got["x"]), want.best_x)
    assert got["claimed"] == _counts(want)["claimed"]


@pytest.mark.gpu
def test_frontend_early_stop(gpu_cases):
    """required = 10 of 512: the call stops once 10 starts have converged
    (reference bfgs_common.cuh early-stop contract); every start is counted
    once and the reported best is a converged start."""
    got = gpu_cases["rosenbrock2_thread_early"]
    assert got["claimed"] >= 10
    assert got["claimed"] + got["stopped"] + got["surrendered"] == 512
    assert got["status"] == 1
    assert np.abs(np.array(got["x"]) - 1.0).max() <= 1e-4


@pytest.mark.gpu
def test_frontend_trajectory_file(orc, gpu_cases):
    """zeus::Zeus(..., trajectory_file) from a user functor writes the TSV of
    the per-thread run, bounded to 50 steps, equal to the oracle's recording."""
    o = orc.run(orc.ROSENBROCK, 2, -5, 5, 32, 300, 3, 33, 1e-8, 42, grad=orc.GRAD_DUAL,
                trajectory=(50, 32), trajectory_alpha=False)
    got = gpu_cases["rosenbrock2_thread_traj"]
    assert got["idx"] == o.best["idx"] and got["fval"] == o.best["fval"]
    with open(gpu_cases["traj_path"]) as fh:
        head = fh.readline().rstrip("\n").split("\t")
        rows = [line.rstrip("\n").split("\t") for line in fh]
    assert head == ["traj", "step", "fval", "grad", "status", "alpha", "x0", "x1"]
    assert len(rows) == 32 * 50
    t = o.trajectory
    for r in rows:
        i, it = int(r[0]), int(r[1])
        for col, arr in ((2, t["fval"][it, i]), (3, t["grad"][it, i]),
                         (6, t["coords"][it, 0, i]), (7, t["coords"][it, 1, i])):
            assert (r[col] == "NaN") if np.isnan(arr) else float(r[col]) == arr
        assert int(r[4]) == t["status"][it, i] and float(r[5]) == 0.0


@pytest.mark.gpu
def test_frontend_device_matrix(gpu_cases):
    """DeviceMatrix (test/device_matrix_tests.cu): a device-heap matrix built,
    filled and read back by one device thread, int and double, repeatedly;
    its copy and move operations are deleted (static_asserts)."""
    assert gpu_cases["device_matrix"]["ok"] == 1
