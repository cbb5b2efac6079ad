"""Ensembles and the continuous-batching server: every run gives the same result alone,
in an ensemble and in a served mix (bitwise: the same kernels act on each run's own
buffers and device clock)."""

import math
import sys

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import ensemble as E
from mhdflows_amd import ic, serve
from oracle import OracleProblem
from oracle.integrator import rel_l2, run_steps

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def alone(n, T, nu, eta, seed, steps, stepper="RK4", cfl=0.3, k0=4.0, dt=0.0):
    p = M.Problem("MI355X", nx=n, ν=nu, η=eta, B_field=True, T=T, stepper=stepper, dt=dt)
    kmax = min(16.0, math.floor(n / 3) - 1)
    idx, vals = ic.solenoidal_modes(seed, k0=k0, kmax=kmax).half_entries(n)
    M.SetUpProblemIC_modes(p, idx, vals)
    if cfl > 0:
        M.TimeIntegrator(p, math.inf, steps - 1, CFL_Coef=cfl)
    else:
        M.stepforward(p, steps)
    return p


def test_sparse_mode_ic_matches_oracle():
    say("sparse IC")
    n = 32
    ms = ic.solenoidal_modes(3, k0=4.0, kmax=9.0)
    p = M.Problem("MI355X", nx=n, ν=1e-2, η=1e-2, B_field=True, T=np.float64, dt=1e-3)
    idx, vals = ms.half_entries(n)
    M.SetUpProblemIC_modes(p, idx, vals)
    o = OracleProblem(nx=n, nu=1e-2, eta=1e-2, B_field=True, T=np.float64, dt=1e-3, mode="fused")
    o.sol[...] = ms.spectrum_on_grid(n)
    assert rel_l2(p.sol, o.sol) < 1e-14
    run_steps(o, 3)
    M.stepforward(p, 3)
    assert rel_l2(p.sol, o.sol) < 1e-12


@pytest.mark.parametrize("stepper", ["RK4", "LSRK54"])
def test_ensemble_uneven_grids_equal_alone(stepper):
    say("ensemble", stepper)
    specs = [(16, 3e-3, 1), (24, 5e-3, 2), (16, 1e-3, 3), (32, 2e-3, 4), (24, 8e-3, 5)]
    cfgs = [E.make_config(n, nu=nu, eta=nu, T=np.float32, stepper=stepper) for (n, nu, _) in specs]
    ens = E.Ensemble(cfgs)
    for r, (n, nu, seed) in enumerate(specs):
        kmax = min(16.0, math.floor(n / 3) - 1)
        idx, vals = ic.solenoidal_modes(seed, k0=3.0, kmax=kmax).half_entries(n)
        ens.set_ic_modes(r, idx, vals)
        ens.set_cfl(r, 0.3)
    ens.step(6)
    for r, (n, nu, seed) in enumerate(specs):
        p = alone(n, np.float32, nu, nu, seed, 6, stepper=stepper, k0=3.0)
        assert np.array_equal(ens.get_sol(r), p.sol), r
        assert ens.clock(r)[2] == 6


def test_server_mix_equals_alone():
    say("server mix")
    reqs = [serve.Request(0, 16, np.float32, 30, 3e-3, 3e-3, 11, 0.0, keep_state=True),
            serve.Request(1, 32, np.float32, 12, 1e-3, 1e-3, 12, 0.0, keep_state=True),
            serve.Request(2, 16, np.float64, 20, 5e-3, 5e-3, 13, 0.0, keep_state=True),
            serve.Request(3, 16, np.float32, 7, 2e-3, 2e-3, 14, 0.01, keep_state=True),
            serve.Request(4, 32, np.float32, 25, 4e-3, 4e-3, 15, 0.02, keep_state=True),
            serve.Request(5, 16, np.float32, 40, 6e-3, 6e-3, 16, 0.03, keep_state=True)]
    srv = serve.Server(hbm_budget=8 << 30, slots={16: 2, 32: 1}, chunk=5)
    res = srv.run(reqs, realtime=False)
    srv.close()
    assert sorted(res) == list(range(6))
    for r in reqs:
        p = alone(r.n, r.T, r.nu, r.eta, r.seed, r.steps)
        assert np.array_equal(res[r.rid].sol, p.sol), r.rid
        assert res[r.rid].finished >= res[r.rid].admitted >= 0


def test_server_budget_refuses_oversized_pool():
    need = E.query_bytes(E.make_config(64, T=np.float32))
    srv = serve.Server(hbm_budget=need // 2, slots={64: 1})
    assert srv._pool(64, np.float32) is None
    srv.close()


def test_server_poisson_stream_small():
    say("server stream")
    reqs = serve.poisson_stream(10, rate=200.0, seed=3, steps=(10, 40), sizes=(16, 32), weights=(0.5, 0.5))
    srv = serve.Server(hbm_budget=8 << 30, slots={16: 4, 32: 2}, chunk=10)
    res = srv.run(reqs, realtime=True)
    s = serve.summarize(res, srv.makespan)
    srv.close()
    assert s["requests"] == 10 and s["p99_latency_s"] >= s["p50_latency_s"] > 0
