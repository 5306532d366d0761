"""Writes tests/golden/golden.json: known-answer vectors that pin the oracle.

    python tests/golden/make_golden.py

Contents and sources:
  philox4x32_10  Random123 known-answer vectors (Salmon et al., SC'11
                 distribution, kat_vectors: philox4x32 with 10 rounds).
  digamma        scipy.special.digamma at a spread of points (SciPy 1.15).
  mlp_grad       gradients of the 1-16-16-1 tanh MSE objective from
                 torch.autograd in float64 on the CPU, at seeded points.
  loglik_grad    gradients of the binned Poisson NLL from torch.autograd in
                 float64 on the CPU.
  reference_kats the reference test suite's known answers, restated as
                 values (file:line of each in the 'source' field).
The MLP / NLL inputs are generated here with numpy's seeded generator so the
fixture does not depend on the product code.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def mlp_torch(theta, xs, ys):
    import torch
    H = 16
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    x = torch.tensor(xs, dtype=torch.float64)
    y = torch.tensor(ys, dtype=torch.float64)
    w1 = th[0:32:2]
    b1 = th[1:32:2]
    l2 = th[32:32 + H * (H + 1)].reshape(H, H + 1)
    w2, b2 = l2[:, :H], l2[:, H]
    w3, b3 = th[32 + H * (H + 1):32 + H * (H + 1) + H], th[-1]
    h1 = torch.tanh(x[:, None] * w1[None, :] + b1[None, :])
    h2 = torch.tanh(h1 @ w2.T + b2[None, :])
    out = h2 @ w3 + b3
    loss = ((out - y) ** 2).mean()
    loss.backward()
    return float(loss.detach()), th.grad.numpy().tolist()


def loglik_torch(theta, x, k):
    import torch
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    x = torch.tensor(x, dtype=torch.float64)
    k = torch.tensor(k, dtype=torch.float64)
    a, b, c, d = th[0], th[1], th[2], th[3]
    mean = a * (1 - x) ** b / x ** (c + d * torch.log(x))
    nll = -(k * torch.log(mean) - mean - torch.lgamma(k + 1)).sum()
    nll.backward()
    return float(nll.detach()), th.grad.numpy().tolist()


def main():
    from scipy.special import digamma

    rng = np.random.default_rng(20251202)
    out = {}
    out["philox4x32_10"] = [
        {"ctr": [0, 0, 0, 0], "key": [0, 0],
         "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
        {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2,
         "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
        {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
         "key": [0xa4093822, 0x299f31d0],
         "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
    ]
    xs = [0.5, 1.0, 1.4616321449683622, 2.0, 3.0, 7.5, 8.0, 10.0, 123.25, 1e4,
          0.25, 0.1, 1e-3, -0.5, -1.5, -2.25, -7.75]
    out["digamma"] = {"x": xs, "psi": [float(digamma(v)) for v in xs]}

    # MLP: 1-16-16-1, 32 points, 3 parameter vectors
    pts = rng.uniform(-1, 1, 32)
    ys = np.sin(3 * pts) + 0.1 * rng.standard_normal(32)
    cases = []
    for _ in range(3):
        th = rng.uniform(-1, 1, 321)
        f, g = mlp_torch(th, pts, ys)
        cases.append({"theta": th.tolist(), "f": f, "g": g})
    out["mlp_grad"] = {"x": pts.tolist(), "y": ys.tolist(), "cases": cases}

    # binned Poisson NLL: 20 bins, counts from a fixed truth, 3 parameter points
    centers = np.linspace(0.08, 0.53, 20)
    truth = 3.2 * (1 - centers) ** 10 / centers ** (5 + 0.1 * np.log(centers))
    counts = rng.poisson(truth).astype(float)
    cases = []
    for th in ([3.2, 10.0, 5.0, 0.1], [2.0, 8.0, 4.5, 0.3], [5.0, 12.0, 5.5, 0.0]):
        f, g = loglik_torch(th, centers, counts)
        cases.append({"theta": th, "f": f, "g": g})
    out["loglik_grad"] = {"x": centers.tolist(), "k": counts.tolist(), "cases": cases}

    out["reference_kats"] = {
        "line_search": {"source": "test/bfgs_tests.cu:121-151", "x": 1.0, "p": -2.0,
                        "g": 2.0, "f0": 1.0, "alpha": 0.5},
        "norm": {"source": "test/bfgs_tests.cu:153-172", "g": [3.0, 4.0], "norm": 5.0},
        "search_direction": {"source": "test/bfgs_tests.cu:195-226", "g": [1.0, -2.0, 0.5],
                             "p": [-1.0, 2.0, -0.5]},
        "bfgs_update": {"source": "test/bfgs_tests.cu:228-264", "s": [1.0, 0.0],
                        "y": [2.0, 0.0], "sTy": 2.0, "H": [[0.5, 0.0], [0.0, 1.0]]},
        "digamma_exact": {"source": "test/dual_tests.cu:511-532",
                          "gamma": 0.5772156649015328606, "tol": 1.1e-13},
        "dual_ops": {"source": "test/dual_tests.cu:192-509", "cases": [
            ["add", [1, 2], [3, 4], [4, 6]],
            ["sub", [5, 2.5], [1.5, 0.5], [3.5, 2.0]],
            ["mul", [2, 3], [4, 5], [8, 22]],
            ["div", [6, 2], [3, 1], [2, 0]],
            ["neg", [1.5, 2.5], None, [-1.5, -2.5]],
            ["abs", [-3, 2], None, [3, -2]],
            ["abs", [3, 2], None, [3, 2]],
            ["sqrt", [4, 1.5], None, [2, 0.375]],
            ["pow", [2, 1], 3.0, [8, 12]],
            ["exp", [1, 3], None, [math.e, 3 * math.e]],
            ["sin", [math.pi / 6, 1], None, [0.5, math.cos(math.pi / 6)]],
            ["cos", [math.pi / 3, 2], None, [0.5, -2 * math.sin(math.pi / 3)]],
            ["atan2", [1, 1], [1, 0], [math.pi / 4, 0.5]],
        ]},
        "fun_values": {"source": "test/fun_tests.cu:75-128",
                       "rastrigin_origin": 0.0, "ackley_origin_abs_tol": 1e-6,
                       "rosenbrock_min": 0.0,
                       "ackley_grad_point": [-1e-17, 1e-17], "ackley_grad": [-2.0, 2.0]},
        "pso_inertia": {"source": "test/pso_tests.cu:96-179", "w": 0.5, "x": [0, 1, 2, 3],
                        "v": [10, 20, 30, 40]},
    }
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
