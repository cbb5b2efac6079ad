"""Checkpoint / restart on the MI355X (savefile, Restart!, readMHDFlows): a bit-exact restart
continues exactly like the uninterrupted run; the reference's restart from real fields sets
vars, R2C(vars) and clock.t; TimeIntegrator! saves at every dump_dt."""

import math
import os
import sys

import numpy as np
import pytest

import mhdflows_amd as M

from test_gpu_parity import pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _same(a, b):
    assert np.array_equal(a.sol, b.sol)
    for f in range(a.nf):
        assert np.array_equal(a.vars.field(f), b.vars.field(f)), f
    assert (a.clock.t, a.clock.dt, a.clock.step) == (b.clock.t, b.clock.dt, b.clock.step)


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_bit_exact_restart_with_cfl(T, tmp_path):
    _, a = pair(32, T=T, nu=1e-3, eta=1e-3, kmax=10.0)
    M.TimeIntegrator(a, math.inf, 3, CFL_Coef=0.3)
    path = M.savefile(a, 7, file_path_and_name=str(tmp_path / "run"), exact=True)
    assert os.path.basename(path) == "run_t_0007.mhdf"
    M.TimeIntegrator(a, math.inf, 4, CFL_Coef=0.3)
    _, b = pair(32, T=T, nu=1e-3, eta=1e-3, kmax=3.0, seed=99)     # another state: all of it is replaced
    M.Restart(b, path, exact=True)
    M.TimeIntegrator(b, math.inf, 4, CFL_Coef=0.3)
    _same(a, b)


def test_bit_exact_restart_keeps_the_a99_draws(tmp_path):
    _, a = pair(32, T=np.float64, kmax=6.0, dt=2e-3)
    M.SetUpFk(a, kf=2.0, P=1.0, seed=5)
    M.stepforward(a, 3)
    path = M.savefile(a, 1, file_path_and_name=str(tmp_path / "a99"), exact=True)
    M.stepforward(a, 4)
    _, b = pair(32, T=np.float64, kmax=6.0, dt=2e-3)
    M.SetUpFk(b, kf=2.0, P=1.0, seed=5)       # the forcing set-up is not part of the file
    M.Restart(b, path, exact=True)
    M.stepforward(b, 4)
    _same(a, b)


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_restart_from_real_fields(T, tmp_path):
    _, a = pair(32, T=T, kmax=8.0)
    M.stepforward(a, 3)
    path = M.savefile(a, 0, file_path_and_name=str(tmp_path / "ref"))
    _, b = pair(32, T=T, kmax=3.0, seed=99)
    M.Restart(b, path)
    fields = [a.vars.field(f) for f in range(6)]
    _, c = pair(32, T=T, kmax=3.0, seed=99)
    M.SetUpProblemIC(c, *fields)               # Move_Data_to_Prob!: vars, then R2C into sol
    assert np.array_equal(b.sol, c.sol)
    for f in range(6):
        assert np.array_equal(b.vars.field(f), fields[f])
    assert b.clock.t == a.clock.t
    with pytest.raises(M.MHDError):
        M.Restart(b, path, exact=True)          # no spectral state in a reference-style file
    iv, jv, kv, ib, jb, kb, t = M.readMHDFlows(path)
    assert t == a.clock.t and iv.dtype == np.float32
    assert np.array_equal(kb, fields[5].astype(np.float32))
    _, d = pair(16, T=T, kmax=3.0)
    with pytest.raises(M.MHDError):
        M.Restart(d, path)                      # another grid
    with pytest.raises(M.MHDError):
        M.Restart(d, str(tmp_path / "missing.mhdf"))


def test_time_integrator_saves_every_dump_dt(tmp_path):
    _, p = pair(16, T=np.float64, kmax=4.0)
    loc = str(tmp_path) + os.sep
    taken = M.TimeIntegrator(p, 1.0, 9, usr_dt=1e-3, save=True, save_loc=loc, filename="tg", dump_dt=2.7e-3)
    assert taken == 10
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("tg_t_"))
    # the start, then the first steps with t >= 2.7e-3, 5.4e-3, 8.1e-3 (integrator.jl:44-51, 136-141)
    assert files == ["tg_t_0000.mhdf", "tg_t_0001.mhdf", "tg_t_0002.mhdf", "tg_t_0003.mhdf"]
    times = [float(M.read_checkpoint(str(tmp_path / f))["time"]) for f in files]
    assert times[0] == 0.0 and all(math.isclose(t, k * 3e-3, rel_tol=1e-12) for k, t in enumerate(times[1:], 1))
