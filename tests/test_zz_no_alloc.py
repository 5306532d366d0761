"""The allocation-free PSO entry point (zeus_pso_run_noalloc) against
zeus_pso_run, through the C ABI: the same kernels on the same inputs, so the
swarm's pBest coordinates and values and the gBest agree bit for bit.  Two
calls in a row, the second with a smaller input, so the second runs on the
scratch array as the first left it; with the library's next device
allocation armed to fail, so a call that allocated would fail."""
import ctypes as C

import numpy as np
import pytest

import zeus_amd as z
from zeus_amd import _capi as capi
from zeus_amd.api import gaussian_150_matrix

pytestmark = pytest.mark.gpu


def _problem(f, lo, hi, N, T, seed, prng):
    addr, nparam = f.device_params()
    return capi.Problem(f.kind, f.dim, float(lo), float(hi), int(N), 0, int(T), 0, seed, 0,
                        int(prng), 0.0, addr or None, nparam, None, None, None)


def _run(entry, f, lo, hi, N, T, seed, prng):
    P = _problem(f, lo, hi, N, T, seed, prng)
    px = z.DeviceArray(nbytes=8 * N * f.dim)
    pv = z.DeviceArray(nbytes=8 * N)
    gx = z.DeviceArray(nbytes=8 * f.dim)
    gv = z.DeviceArray(nbytes=8)
    rc = getattr(capi.lib(), entry)(C.byref(P), px.ptr, pv.ptr, gx.ptr, gv.ptr, None)
    return rc, (px.to_host().reshape(N, f.dim), pv.to_host(), gx.to_host(), gv.to_host()[0])


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y))


CASES = [  # objective, box, (N first, N second), PSO iterations, generator
    ("gaussian150", lambda: z.Gaussian(gaussian_150_matrix()), 5.0, (4096, 1000), 10,
     z.PRNGType.XORWOW),
    ("gaussian150_philox", lambda: z.Gaussian(gaussian_150_matrix()), 5.0, (3000, 77), 10,
     z.PRNGType.PHILOX),
    ("rosenbrock2", lambda: z.Rosenbrock(2), 5.0, (1024, 300), 10, z.PRNGType.XORWOW),
    ("rastrigin20", lambda: z.Rastrigin(20), 5.12, (2000, 129), 3, z.PRNGType.SOBOL),
]


@pytest.mark.parametrize("name,make,box,sizes,T,prng", CASES, ids=[c[0] for c in CASES])
def test_noalloc_pso_equals_pso_run(name, make, box, sizes, T, prng):
    f = make()
    f.device_params()  # the caller's own upload, before the allocation watch
    for N in sizes:
        rc, want = _run("zeus_pso_run", f, -box, box, N, T, 42, prng)
        assert rc == capi.OK
        # the caller's outputs exist before the watch; the next library
        # allocation (counted from the arming) would fail
        P = _problem(f, -box, box, N, T, 42, prng)
        bufs = [z.DeviceArray(nbytes=8 * N * f.dim), z.DeviceArray(nbytes=8 * N),
                z.DeviceArray(nbytes=8 * f.dim), z.DeviceArray(nbytes=8)]
        z.set_fault_hooks(fail_alloc_at=1)
        try:
            px, pv, gx, gv = bufs
            rc = capi.lib().zeus_pso_run_noalloc(C.byref(P), px.ptr, pv.ptr, gx.ptr, gv.ptr,
                                                 None)
        finally:
            z.set_fault_hooks(0, 0)
        assert rc == capi.OK, (name, N, rc)
        got = (px.to_host().reshape(N, f.dim), pv.to_host(), gx.to_host(), gv.to_host()[0])
        _same(got, want)


def _call_noalloc(P, bufs, N, dim):
    px, pv, gx, gv = bufs
    rc = capi.lib().zeus_pso_run_noalloc(C.byref(P), px.ptr, pv.ptr, gx.ptr, gv.ptr, None)
    return rc, (px.to_host().reshape(N, dim), pv.to_host(), gx.to_host(), gv.to_host()[0])


def test_noalloc_rejects_requests_beyond_the_scratch():
    f = z.Gaussian(gaussian_150_matrix())
    N = 8192  # 8192 x 150 start-coordinates with XORWOW states: about 50 MB
    P = _problem(f, -5, 5, N, 2, 42, z.PRNGType.XORWOW)
    px, pv = z.DeviceArray(nbytes=8 * N * 150), z.DeviceArray(nbytes=8 * N)
    rc = capi.lib().zeus_pso_run_noalloc(C.byref(P), px.ptr, pv.ptr, None, None, None)
    assert rc == capi.E_SCRATCH
    assert b"scratch" in capi.lib().zeus_status_string(rc)
    # the library still works afterwards, also for a request that fits
    rc, got = _call_noalloc(_problem(f, -5, 5, 64, 2, 42, z.PRNGType.PHILOX),
                            [z.DeviceArray(nbytes=8 * 64 * 150), z.DeviceArray(nbytes=8 * 64),
                             z.DeviceArray(nbytes=8 * 150), z.DeviceArray(nbytes=8)], 64, 150)
    assert rc == capi.OK
    rc, want = _run("zeus_pso_run", f, -5, 5, 64, 2, 42, z.PRNGType.PHILOX)
    _same(got, want)
