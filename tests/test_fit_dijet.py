"""The dijet-spectrum fit driver (reference examples/fit_dijet_spectrum.cu:
19-221): its data loader, its fit through the C++ front end and the
reference-shaped LogLikelihood class, and its post-fit report, checked
against the oracle's fit and a numpy restatement of the diagnostics."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIT = os.path.join(ROOT, "zeus_amd", "bin", "zeus_fit_dijet")


def _exe():
    if not os.path.exists(FIT):
        pytest.fail("zeus_amd/bin/zeus_fit_dijet is missing: run python -m zeus_amd.build")
    return FIT


def _spectrum(path, seed=7):
    """100 bins of m in [1, 7] TeV, counts Poisson around the dijet model at
    p1 = 10, p2 = 5, p3 = 0.1 with about 1e6 events (x = m / 13)."""
    lo = 1.0 + 0.06 * np.arange(100)
    hi = lo + 0.06
    x = 0.5 * (lo + hi) / 13.0
    shape = (1 - x) ** 10 / x ** (5 + 0.1 * np.log(x))
    mu = 1e6 * shape / shape.sum()
    k = np.random.default_rng(seed).poisson(mu)
    with open(path, "w") as fh:
        fh.write("bin_low\tbin_high\tcounts\n")
        for a, b, c in zip(lo, hi, k):
            fh.write(f"{a:.6f}\t{b:.6f}\t{int(c)}\n")
    # the edges as the file holds them
    lo = np.array([float(f"{v:.6f}") for v in lo])
    hi = np.array([float(f"{v:.6f}") for v in hi])
    return lo, hi, k


def test_driver_usage_and_missing_data(tmp_path):
    p = subprocess.run([_exe(), "10"], capture_output=True, text=True)
    assert p.returncode == 1 and "Usage:" in p.stderr
    p = subprocess.run([_exe(), "64", "100", "0", "--bogus"], capture_output=True, text=True)
    assert p.returncode == 1 and "Unknown argument: --bogus" in p.stderr
    p = subprocess.run([_exe(), "64", "100", "0", "--data", str(tmp_path / "none.tsv")],
                       capture_output=True, text=True, cwd=tmp_path)
    assert p.returncode == 1 and "couldn't open" in p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [["--parallel"], ["--adjoint"]])
def test_fit_dijet_driver(orc, tmp_path, mode):
    data = str(tmp_path / "spectrum.tsv")
    lo, hi, k = _spectrum(data)
    N, iters = 1024, 400
    p = subprocess.run([_exe(), str(N), str(iters), "3", "--data", data, "--sqrt-s", "13",
                        *mode], capture_output=True, text=True, cwd=tmp_path, timeout=150)
    assert p.returncode == 0, p.stderr[-2000:]
    out = p.stdout
    assert "Loaded 100 bins; x in [" in out and "#bins = 100" in out
    assert f"total events in the spectrum = {int(k.sum())}" in out
    nll = float(out.split("best NLL: ")[1].split()[0])
    # the oracle's fit of the same data (same starts, same generator)
    x = np.minimum(np.maximum(0.5 * (lo + hi) / 13.0, 1e-12), 1 - 1e-12)
    params = np.concatenate([x, k.astype(float), [math.lgamma(v + 1.0) for v in k]])
    o = orc.run(orc.LOGLIK, 4, 0.0, 10.0, N, iters, 10, 100, 1e-8, 42, params=params)
    assert nll == pytest.approx(o.best["fval"], rel=2e-6)  # printed with 6 digits
    # the per-bin table and the summary line
    rows = [l.split("\t") for l in open(tmp_path / "dijet_fit_vs_data.tsv").read().splitlines()]
    assert rows[0] == ["bin_low", "bin_high", "x_center", "true_counts", "pred_mu",
                       "pred_residual", "pred_pull"]
    t = np.array([[float(v) for v in r] for r in rows[1:]])
    assert t.shape == (100, 7)
    assert np.array_equal(t[:, 3], k) and np.array_equal(t[:, 2], x)
    mu = t[:, 4]
    assert np.array_equal(t[:, 5], k - mu)
    assert np.array_equal(t[:, 6], (k - mu) / np.sqrt(np.maximum(mu, 1e-12)))
    # the same minimum as the oracle's fit: predicted means agree
    a, b, c, d = o.best_x
    mu_o = a * (1 - x) ** b / x ** (c + d * np.log(x))
    assert np.allclose(mu, mu_o, rtol=1e-5, atol=1e-9)
    m = np.maximum(mu, 1e-12)
    chi2 = ((k - m) ** 2 / m).sum() / 96
    with np.errstate(divide="ignore", invalid="ignore"):
        dev = np.where(k > 0, 2 * (k * np.log(k / m) - (k - m)), 2 * m).sum() / 96
    line = [l for l in out.splitlines() if l.startswith("chi2/ndf=")][0]
    c2, dv = (float(f.split("=")[1]) for f in line.split())
    assert c2 == pytest.approx(chi2, rel=1e-5) and dv == pytest.approx(dev, rel=1e-5)
    assert 0.5 < chi2 < 2.0  # data drawn from the model
