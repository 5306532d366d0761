"""Trajectory recording: the oracle against the reference's own trajectory
test (bfgs_tests.cu:554-671) and against what the recorded quantities must
be; the TSV writer (reference src/utils.cu:126-286, util_tests.cu:120-189);
the GPU forms against the oracle (bitwise for the per-thread form)."""
import math
import os

import numpy as np
import pytest

import zeus_amd as z
from zeus_amd import _capi as capi


def _sphere_run(orc, **kw):
    # reference bfgs_tests.cu:554-671: sum x^2 from (1, 1), N = 1, 50 steps
    return orc.run(orc.SPHERE, 2, -5.0, 5.0, 1, 50, -1, 1, 1e-6, 42,
                   x0=np.array([[1.0, 1.0]]), trajectory=(50, 1), **kw)


def test_oracle_trajectory_matches_reference_test(orc):
    o = _sphere_run(orc)
    t = o.trajectory
    fv, st = t["fval"][:, 0], t["status"][:, 0]
    assert not math.isnan(fv[0]) and fv[0] == 2.0          # f(1, 1) = 2 > 0
    conv = np.flatnonzero(st == 1)
    assert conv.size == 1 and conv[0] < 49                 # converged before the last slot
    c = int(conv[0])
    assert np.all(np.isnan(fv[c + 1:]))                    # never written afterwards
    assert np.all(np.isnan(t["coords"][c + 1:]))
    assert np.all(st[:c] == -1) and np.all(st[c + 1:] == -1)
    assert c == o.iter[0] and o.status[0] == 1
    assert np.array_equal(t["coords"][0, :, 0], [1.0, 1.0])
    assert t["grad"][0, 0] == math.sqrt(8.0)               # |(2, 2)|
    assert 0.0 < t["alpha"][0, 0] <= 1.0


def test_oracle_trajectory_is_the_iteration(orc):
    """Each recorded step is (x, f(x), |grad f(x)|) of the accepted point, the
    values are non-increasing, and the status lands where the run ended."""
    N, T = 24, 200
    o = orc.run(orc.ROSENBROCK, 2, -5.0, 5.0, N, T, 3, N + 1, 1e-8, 42, trajectory=(T, N))
    t = o.trajectory
    ob = orc.Obj(orc.ROSENBROCK, 2)
    for s in range(N):
        last = o.iter[s] if o.status[s] != 0 else T - 1
        fv = t["fval"][:last + 1, s]
        assert not np.any(np.isnan(fv)) and np.all(np.isnan(t["fval"][last + 1:, s]))
        assert np.all(np.diff(fv) <= 0.0)
        for it in (0, last // 2, last):
            x = t["coords"][it, :, s]
            f, g = ob.value_grad_adjoint(x)
            assert fv[it] == f
            assert t["grad"][it, s] == orc.norm(g)
        code = t["status"][last, s]
        assert code == o.status[s]
        assert np.count_nonzero(t["status"][:, s] != -1) == 1
        if o.status[s] == 1:
            assert t["grad"][last, s] >= 1e-8  # the top-of-iteration norm, before the step
        a = t["alpha"][:last + (o.status[s] != 2), s]
        assert np.all((a > 0.0) & (a <= 1.0))


def test_oracle_trajectory_bounded_is_a_slice(orc):
    N, T = 12, 300
    full = orc.run(orc.ROSENBROCK, 2, -5.0, 5.0, N, T, 2, N + 1, 1e-8, 7, trajectory=(T, N))
    part = orc.run(orc.ROSENBROCK, 2, -5.0, 5.0, N, T, 2, N + 1, 1e-8, 7, trajectory=(9, 5))
    for k in ("fval", "grad", "status", "alpha"):
        assert np.array_equal(part.trajectory[k], full.trajectory[k][:9, :5], equal_nan=True)
    assert np.array_equal(part.trajectory["coords"], full.trajectory["coords"][:9, :, :5],
                          equal_nan=True)
    none = orc.run(orc.ROSENBROCK, 2, -5.0, 5.0, N, T, 2, N + 1, 1e-8, 7, trajectory=(T, N),
                   trajectory_alpha=False)
    assert np.all(none.trajectory["alpha"] == 0.0)


def _parse_tsv(path):
    with open(path) as fh:
        head = fh.readline().rstrip("\n").split("\t")
        rows = [line.rstrip("\n").split("\t") for line in fh]
    return head, rows


def test_writer_tsv_format(tmp_path):
    N, T, D = 3, 4, 2
    rng = np.random.default_rng(1)
    coords = rng.normal(size=(T, D, N))
    fval = rng.normal(size=(T, N))
    grad = np.abs(rng.normal(size=(T, N)))
    status = np.full((T, N), -1, np.int8)
    alpha = rng.uniform(size=(T, N))
    fval[3, 1] = np.nan
    coords[3, :, 1] = np.nan
    status[2, 1] = 1
    path = str(tmp_path / "t.tsv")
    z.write_trajectory(path, dict(coords=coords, fval=fval, grad=grad, status=status,
                                  alpha=alpha))
    head, rows = _parse_tsv(path)
    assert head == ["traj", "step", "fval", "grad", "status", "alpha", "x0", "x1"]
    assert len(rows) == N * T
    k = 0
    for i in range(N):          # starts outer, steps inner
        for it in range(T):
            r = rows[k]
            k += 1
            assert r[0] == str(i) and r[1] == str(it)
            if np.isnan(fval[it, i]):
                assert r[2] == "NaN"
            else:
                assert float(r[2]) == fval[it, i] and r[2] == f"{fval[it, i]:.17e}"
            assert float(r[3]) == grad[it, i]
            assert int(r[4]) == status[it, i]
            assert float(r[5]) == alpha[it, i]
            for d in range(D):
                v = coords[it, d, i]
                assert (r[6 + d] == "NaN") if np.isnan(v) else float(r[6 + d]) == v


def test_writer_routes_by_extension(tmp_path):
    # reference util_tests.cu:120-189 (built without NetCDF: .nc is an error)
    L = capi.lib()
    N, T, D = 2, 3, 2
    c = np.zeros(N * T * D)
    s = np.zeros(N * T)
    st = np.zeros(N * T, np.int8)
    args = (c.ctypes.data_as(capi._dp), s.ctypes.data_as(capi._dp),
            s.ctypes.data_as(capi._dp), st.ctypes.data, s.ctypes.data_as(capi._dp), N, T, D)
    ok = str(tmp_path / "test_output.tsv")
    assert L.zeus_write_trajectory(ok.encode(), *args) == capi.OK and os.path.exists(ok)
    assert L.zeus_write_trajectory(str(tmp_path / "o.nc").encode(), *args) == capi.E_UNSUPPORTED
    assert L.zeus_write_trajectory(str(tmp_path / "o.txt").encode(), *args) == capi.E_ARG
    bad = str(tmp_path / "missing_dir" / "o.tsv")
    assert L.zeus_write_trajectory(bad.encode(), *args) == capi.E_IO
    # alpha NULL is written as zeros
    assert L.zeus_write_trajectory(ok.encode(), *args[:4], None, N, T, D) == capi.OK
    _, rows = _parse_tsv(ok)
    assert all(float(r[5]) == 0.0 for r in rows)


# ---------------------------------------------------------------- GPU -------
@pytest.mark.gpu
def test_thread_form_trajectory_bitwise(orc):
    # no early stop: which starts a stop flag catches depends on timing
    N, T = 64, 400
    r = z.run_batch([z.Problem(z.Rosenbrock(2), -5.0, 5.0, N, T, 5, N + 1, 1e-8, 42,
                               trajectory=(T, N))], form="thread", per_start=True)[0]
    o = orc.run(orc.ROSENBROCK, 2, -5.0, 5.0, N, T, 5, N + 1, 1e-8, 42, trajectory=(T, N))
    assert np.array_equal(r.per_start["status"], o.status)
    for k in ("coords", "fval", "grad", "status", "alpha"):
        assert np.array_equal(r.trajectory[k], o.trajectory[k], equal_nan=True), k


@pytest.mark.gpu
def test_reference_trajectory_case_on_gpu():
    for form in ("thread", "group"):
        r = z.run_batch([z.Problem(z.Sphere(2), -5.0, 5.0, 1, 50, -1, 1, 1e-6, 42,
                                   x0=np.array([[1.0, 1.0]]), trajectory=(50, 1))],
                        form=form)[0]
        fv, st = r.trajectory["fval"][:, 0], r.trajectory["status"][:, 0]
        assert fv[0] == 2.0
        c = int(np.flatnonzero(st == 1)[0])
        assert c < 49 and np.all(np.isnan(fv[c + 1:]))


@pytest.mark.gpu
def test_group_form_trajectory_to_tolerance(orc):
    # Philox seed 5: no start ends within rounding of the iteration limit or
    # of the tolerance (such a start may end either way under reordered sums)
    N, T, D = 16, 300, 10
    r = z.run_batch([z.Problem(z.Rosenbrock(D), -5.0, 5.0, N, T, 3, N + 1, 1e-8, 5,
                               trajectory=(T, N), prng=z.PRNGType.PHILOX)],
                    form="group", per_start=True)[0]
    o = orc.run(orc.ROSENBROCK, D, -5.0, 5.0, N, T, 3, N + 1, 1e-8, 5, trajectory=(T, N),
                prng=orc.PRNG_PHILOX)
    assert np.array_equal(r.per_start["status"], o.status)
    tg, to = r.trajectory, o.trajectory
    # the first steps agree to rounding (sums in another order), the written
    # extent of each start and its final status too (within one iteration)
    assert np.allclose(tg["fval"][:5], to["fval"][:5], rtol=1e-10, atol=0)
    assert np.allclose(tg["coords"][:5], to["coords"][:5], rtol=1e-9, atol=1e-12)
    for s in range(N):
        ng = np.count_nonzero(~np.isnan(tg["fval"][:, s]))
        no = np.count_nonzero(~np.isnan(to["fval"][:, s]))
        assert abs(ng - no) <= 1
        assert tg["status"][ng - 1, s] == o.status[s]
    assert np.all(tg["alpha"][0] > 0)


@pytest.mark.gpu
def test_zeus_writes_trajectory_file(tmp_path):
    path = str(tmp_path / "traj.tsv")
    N, T = 8, 100
    r = z.Zeus(z.Rosenbrock(2), -5.0, 5.0, N, T, 2, N + 1, "rosenbrock", 1e-8, 42, 0,
               parallel=False, trajectory_file=path)
    head, rows = _parse_tsv(path)
    assert head[:6] == ["traj", "step", "fval", "grad", "status", "alpha"]
    assert len(rows) == N * T
    t = r.trajectory
    for row in rows[::37]:
        i, it = int(row[0]), int(row[1])
        v = t["fval"][it, i]
        assert (row[2] == "NaN") if np.isnan(v) else float(row[2]) == v
        assert float(row[5]) == 0.0  # sequential form: no alpha, as the reference
    # bounded recording is a slice of the full one
    rb = z.Zeus(z.Rosenbrock(2), -5.0, 5.0, N, T, 2, N + 1, "rosenbrock", 1e-8, 42, 0,
                parallel=False, trajectory_file=path, trajectory_steps=10,
                trajectory_starts=3)
    assert np.array_equal(rb.trajectory["fval"], t["fval"][:10, :3], equal_nan=True)
    assert len(_parse_tsv(path)[1]) == 30


@pytest.mark.gpu
def test_trajectory_refuses_tile_form():
    C = np.eye(32) * 2.0
    with pytest.raises(capi.ZeusError) as e:
        z.run_batch([z.Problem(z.Gaussian(C), -5.0, 5.0, 32, 50, 2, 33, 1e-8, 1,
                               trajectory=(5, 4))], form="tile")
    assert e.value.code == capi.E_ARG
    # auto form still runs (per-thread or per-workgroup) and records
    r = z.run_batch([z.Problem(z.Gaussian(C), -5.0, 5.0, 32, 50, 2, 33, 1e-8, 1,
                               trajectory=(5, 4))])[0]
    assert r.trajectory["fval"].shape == (5, 4) and not np.isnan(r.trajectory["fval"][0, 0])
