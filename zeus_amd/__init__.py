"""zeus_amd: MI355X (gfx950) multistart PSO -> BFGS optimizer.

A from-scratch HIP implementation of the hot path of fnal-numerics/zeus:
counter-based PSO seeding, per-thread and per-workgroup BFGS with hand-written
adjoint or forward-mode dual gradients, atomic early stop and a deterministic
argmin.  The compute lives in libzeus_amd.so (C ABI: include/zeus_c.h); this
package is the Python host mirror of the reference's zeus::Zeus API.
"""
from .api import (Ackley, Convergence, DeviceArray, Gaussian, LogLikelihood, Mlp, Objective,
                  PRNGType, Problem, Rastrigin, Result, Rosenbrock, Sphere, Zeus, argmin,
                  evaluate, make_dijet, make_mlp_data, make_spd, philox_fill, philox_uniform,
                  pso, run_batch, set_fault_hooks, tuned_config, write_trajectory,
                  euclidean_error, append_results_tsv, stream_uniforms, Swarm, bfgs_update)
from . import _capi as capi

__all__ = [
    "Ackley", "Convergence", "DeviceArray", "Gaussian", "LogLikelihood", "Mlp", "Objective",
    "PRNGType", "Problem", "Rastrigin", "Result", "Rosenbrock", "Sphere", "Zeus", "argmin",
    "evaluate", "make_dijet", "make_mlp_data", "make_spd", "philox_fill", "philox_uniform",
    "pso", "run_batch", "set_fault_hooks", "tuned_config", "write_trajectory",
    "euclidean_error", "append_results_tsv", "stream_uniforms", "Swarm", "bfgs_update",
    "capi",
]
