"""The workloads the benchmark is quoted on, checked on the MI355X (all through the C ABI).

* 256^3 Float32 MHD (nu = eta = 1e-3, CFL 0.3, seeded k^4 exp(-2(k/4)^2) fields): one
  right-hand side and two CFL steps against the Float32 oracle.
* 1024^3 Float32 MHD with the |k| = 2 forcing (nu = eta = 2e-4, dt = 5e-4, 50 RK4 steps)
  against a 256^3 Float64 run of the same seed on the modes a 256^3 grid keeps.
* 64 uneven runs (16 each of 32^3, 48^3, 64^3, 96^3, Float32 LSRK54, CFL 0.3) stepped as one
  ensemble: every member equals its alone run bit for bit; a 48^3 and a 96^3 member match
  the oracle.
* A served mix over 32^3..256^3 with a Float64 request: every result equals its alone run.
* Every large transform length on one axis of an anisotropic grid (256, 384, 768, 1024),
  right-hand side and two RK4 steps in Float32 and Float64, for each launch variant.

Tolerances: Float64 <= 1e-12 relative L2; Float32 right-hand sides <= 2e-5 and Float32 steps
<= 1e-5 against the Float32 oracle; 1024^3 Float32 against 256^3 Float64 <= 1e-6.
"""

import functools
import math
import sys

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import ensemble as E
from mhdflows_amd import ic, serve
from mhdflows_amd.devmem import DeviceArray
from oracle import OracleProblem
from oracle.integrator import rel_l2, run_steps, time_integrator
from oracle.rhs import c2r
from oracle.rhs import calcN as o_calcN

from test_gpu_parity import _resample, pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _calcN_gpu(p, sol_host):
    sol_d = DeviceArray.from_host(sol_host)
    N_d = DeviceArray(sol_host.shape, sol_host.dtype)
    M.calcN(p, N_d.ptr, sol_d.ptr)
    return N_d.download(), sol_d.download()


# ---------------------------------------------------------------------------------------
# 256^3 Float32: the single large run
# ---------------------------------------------------------------------------------------
def test_256_fp32_rhs_and_cfl_steps_match_oracle():
    n = 256
    say("256^3 FP32: building fields")
    fields = ic.solenoidal_modes(2024, k0=4.0, kmax=16.0).real_fields(n, dtype=np.float64)
    o = OracleProblem(nx=n, nu=1e-3, eta=1e-3, B_field=True, T=np.float32, mode="fused")
    o.set_ic(*fields)
    p = M.Problem("MI355X", nx=n, ν=1e-3, η=1e-3, B_field=True, T=np.float32)
    M.SetUpProblemIC(p, *fields)
    del fields
    say("256^3 FP32: one right-hand side")
    N_gpu, sol_after = _calcN_gpu(p, o.sol.copy())
    N_ref = np.zeros_like(o.sol)
    s_ref = o.sol.copy()
    o_calcN(N_ref, s_ref, 0.0, o)
    err = rel_l2(N_gpu, N_ref)
    say(f"  rhs rel L2 {err:.3e}")
    assert err < 2e-5
    assert np.array_equal(sol_after, s_ref)
    del N_gpu, N_ref, sol_after, s_ref
    say("256^3 FP32: two CFL 0.3 steps")
    time_integrator(o, math.inf, 1, CFL_Coef=0.3)
    taken = M.TimeIntegrator(p, math.inf, 1, CFL_Coef=0.3)
    assert taken == 2 and p.clock.step == 2
    assert math.isclose(p.clock.dt, float(o.dt), rel_tol=1e-5)
    err = rel_l2(p.sol, o.sol)
    say(f"  2 steps rel L2 {err:.3e}, dt {p.clock.dt:.6e} vs {float(o.dt):.6e}")
    assert err < 1e-5


# ---------------------------------------------------------------------------------------
# 1024^3 Float32 against 256^3 Float64 (precision and HBM-scale run)
# ---------------------------------------------------------------------------------------
N_REF = 256


def retained_idx(n):
    """(kx, l, m) array indices on an n^3 grid of the modes a 256^3 grid keeps."""
    lo = N_REF // 3
    ks = np.arange(-lo, lo)                       # kept signed y/z wavenumbers of 256^3
    kx = np.arange(0, lo)                         # kept kx
    m, l, x = np.meshgrid(ks % n, ks % n, kx, indexing="ij")
    return np.stack([x.ravel(), l.ravel(), m.ravel()], 1).astype(np.int32)


def _c3_run(n, T, steps):
    p = M.Problem("MI355X", nx=n, ν=2e-4, η=2e-4, B_field=True, T=T, dt=5e-4)
    idx, vals = ic.solenoidal_modes(2024, k0=4.0, kmax=16.0).half_entries(n)
    M.SetUpProblemIC_modes(p, idx, vals)
    M.set_forcing(p, 7, kf=2.0, amp=1.0)
    M.stepforward(p, steps)
    d = M.Diagnostics(p)
    r = M.get_sol_modes(p, retained_idx(n)).astype(np.complex128) * (float(N_REF) / n) ** 3
    del p
    return r, d


def test_1024_fp32_matches_256_fp64_on_retained_modes():
    steps = 50
    say("1024^3 FP32 precision run: 256^3 FP64 reference")
    ref, dref = _c3_run(256, np.float64, steps)
    say("1024^3 FP32 precision run: 1024^3 FP32")
    got, dgot = _c3_run(1024, np.float32, steps)
    assert np.all(np.isfinite(got))
    err = rel_l2(got, ref)
    names = ["ux", "uy", "uz", "bx", "by", "bz"]
    for f in range(6):
        say(f"  {names[f]}: rel L2 {rel_l2(got[:, f], ref[:, f]):.3e}")
    say(f"  all fields rel L2 {err:.3e}")
    assert err < 1e-6
    for k in ("E_kin", "E_mag"):
        rel = abs(dgot[k] - dref[k]) / dref[k]
        say(f"  {k}: {dgot[k]:.9g} vs {dref[k]:.9g} (rel {rel:.2e})")
        assert rel < 1e-6
    say(f"  max|div b|: {dgot['max_div_b']:.3e} (FP32) vs {dref['max_div_b']:.3e} (FP64)")
    assert not dgot["has_nan"]
    # divergence at Float32 round-off of the |k|-weighted field (|k| b ~ 16 * O(1) per mode)
    assert dgot["max_div_b"] < 5e-3 and dref["max_div_b"] < 1e-11


# ---------------------------------------------------------------------------------------
# 64 uneven runs in one ensemble
# ---------------------------------------------------------------------------------------
def _ensemble_specs():
    rng = np.random.Generator(np.random.PCG64(4))
    specs = []
    for n in (32, 48, 64, 96):
        for _ in range(16):
            specs.append((n, float(rng.uniform(2, 6)), float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-2)))),
                          int(rng.integers(1 << 30))))
    return specs


def _modes(n, k0, seed):
    return ic.solenoidal_modes(seed, k0=k0, kmax=min(16.0, n // 3 - 1))


def test_uneven_ensemble_64_runs_equal_alone_and_oracle():
    steps = 4
    specs = _ensemble_specs()
    say("ensemble: 64 runs of 32/48/64/96^3, FP32 LSRK54, CFL 0.3")
    ens = E.Ensemble([E.make_config(n, nu=nu, eta=nu, T=np.float32, stepper="LSRK54") for (n, k0, nu, s) in specs])
    for r, (n, k0, nu, s) in enumerate(specs):
        idx, vals = _modes(n, k0, s).half_entries(n)
        ens.set_ic_modes(r, idx, vals)
        ens.set_cfl(r, 0.3)
    ens.step(steps)
    got = [ens.get_sol(r) for r in range(len(specs))]
    clocks = [ens.clock(r) for r in range(len(specs))]
    ens.close()
    for r, (n, k0, nu, s) in enumerate(specs):
        p = M.Problem("MI355X", nx=n, ν=nu, η=nu, B_field=True, T=np.float32, stepper="LSRK54")
        idx, vals = _modes(n, k0, s).half_entries(n)
        M.SetUpProblemIC_modes(p, idx, vals)
        M.TimeIntegrator(p, math.inf, steps - 1, CFL_Coef=0.3)
        assert np.array_equal(got[r], p.sol), (r, n)
        assert clocks[r][2] == steps and clocks[r][0] == p.clock.t
        del p
    for r in (16, 63):   # the first 48^3 run and the last 96^3 run
        n, k0, nu, s = specs[r]
        say(f"ensemble: member {r} ({n}^3) against the oracle")
        o = OracleProblem(nx=n, nu=nu, eta=nu, B_field=True, T=np.float32, stepper="LSRK54", mode="fused")
        o.sol[...] = _modes(n, k0, s).spectrum_on_grid(n)
        o.vars[...] = c2r(o.sol, (n, n, n)).astype(np.float32)
        time_integrator(o, math.inf, steps - 1, CFL_Coef=0.3)
        err = rel_l2(got[r], o.sol)
        say(f"  rel L2 {err:.3e}")
        assert err < 1e-5


# ---------------------------------------------------------------------------------------
# serving: a mix of grids and precisions
# ---------------------------------------------------------------------------------------
def test_served_mix_of_grids_and_precisions_equals_alone():
    say("serving: mix over 32..256^3 with a Float64 request")
    reqs = [serve.Request(0, 32, np.float32, 30, 3e-3, 3e-3, 21, 0.0, keep_state=True),
            serve.Request(1, 64, np.float32, 20, 1e-3, 1e-3, 22, 0.0, keep_state=True),
            serve.Request(2, 128, np.float32, 12, 2e-4, 2e-4, 23, 0.0, keep_state=True),
            serve.Request(3, 256, np.float32, 8, 1e-4, 1e-4, 24, 0.0, keep_state=True),
            serve.Request(4, 64, np.float64, 15, 5e-3, 5e-3, 25, 0.0, keep_state=True),
            serve.Request(5, 32, np.float32, 9, 8e-3, 8e-3, 26, 0.01, keep_state=True),
            serve.Request(6, 32, np.float32, 17, 2e-3, 2e-3, 27, 0.02, keep_state=True),
            serve.Request(7, 128, np.float32, 6, 6e-4, 6e-4, 28, 0.03, keep_state=True)]
    srv = serve.Server(hbm_budget=64 << 30, slots={32: 2, 64: 2, 128: 1, 256: 1}, chunk=5)
    res = srv.run(reqs, realtime=False)
    srv.close()
    assert sorted(res) == list(range(len(reqs)))
    for r in reqs:
        assert res[r.rid].error is None
        p = M.Problem("MI355X", nx=r.n, ν=r.nu, η=r.eta, B_field=True, T=r.T, stepper="RK4")
        idx, vals = serve.initial_modes(r)
        M.SetUpProblemIC_modes(p, idx, vals)
        M.TimeIntegrator(p, math.inf, r.steps - 1, CFL_Coef=0.3)
        assert np.array_equal(res[r.rid].sol, p.sol), r.rid
        del p


# ---------------------------------------------------------------------------------------
# every large transform length, on each axis
# ---------------------------------------------------------------------------------------
SHAPES = [(256, 16, 16), (16, 256, 16), (16, 16, 256), (1024, 8, 8), (16, 8, 1024), (768, 8, 8), (384, 16, 8),
          (16, 768, 8), (16, 8, 384), (512, 8, 8), (16, 512, 8), (192, 8, 16)]


@functools.lru_cache(maxsize=None)
def _length_case(shape):
    """Fields of a shape and the Float64 / Float32 oracle right-hand side and two-step state."""
    ms = ic.solenoidal_modes(5, k0=2.0, kmax=2.5, mhd=True)
    fields = _resample(ms, shape[0], shape[1], shape[2], True)
    return fields, {T: _oracle_rhs_and_steps(shape, T, fields) for T in (np.float64, np.float32)}


def _oracle_rhs_and_steps(shape, T, fields):
    o = OracleProblem(nx=shape[0], ny=shape[1], nz=shape[2], nu=0.01, eta=0.02, B_field=True, T=T, dt=1e-3,
                      mode="fused")
    o.set_ic(*fields)
    N = np.zeros_like(o.sol)
    o_calcN(N, o.sol.copy(), 0.0, o)
    run_steps(o, 2)
    return N, o.sol


@pytest.mark.parametrize("var", [0, 1, 2, 3])
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("shape", SHAPES)
def test_large_lengths_each_axis(shape, T, var):
    """Float64: <= 1e-12 against the Float64 oracle.  Float32: a right-hand side on a long axis
    is computed on every mode, and its high-|k| modes are Float32 transform rounding amplified
    by k, so it is held to the accuracy of the oracle's own Float32 path (at most twice that
    path's error against the Float64 oracle, plus 1e-6); after two steps the state is within
    1e-5 of the Float32 oracle and 1e-6 of the Float64 one."""
    say("lengths", shape, T.__name__, var)
    fields, ref = _length_case(shape)
    p = M.Problem("MI355X", nx=shape[0], ny=shape[1], nz=shape[2], ν=0.01, η=0.02, B_field=True, T=T, dt=1e-3)
    M.SetUpProblemIC(p, *fields)
    M.variants(p, [var] * 5)
    N64, s64 = ref[np.float64]
    sol0 = p.sol
    N_gpu, _ = _calcN_gpu(p, sol0)
    M.stepforward(p, 2)
    if T == np.float64:
        assert rel_l2(N_gpu, N64) < 1e-12
        assert rel_l2(p.sol, s64) < 1e-12
        return
    N32, s32 = ref[np.float32]
    e_gpu, e_orc = rel_l2(N_gpu, N64), rel_l2(N32, N64)
    say(f"  rhs vs FP64 oracle: GPU {e_gpu:.2e}, FP32 oracle {e_orc:.2e}")
    assert e_gpu < 2 * e_orc + 1e-6
    assert rel_l2(p.sol, s32) < 1e-5 and rel_l2(p.sol, s64) < 1e-6


# ---------------------------------------------------------------------------------------
# diagnostics against closed forms; partial state reads
# ---------------------------------------------------------------------------------------
def _abc(k, A, B, C, n):
    X, Y, Z = ic.grid_coords(n)
    s, c = np.sin, np.cos
    return [A * s(k * Z) + C * c(k * Y), B * s(k * X) + A * c(k * Z), C * s(k * Y) + B * c(k * X)]


def test_gpu_helicities_beltrami_closed_forms():
    """curl u = k_u u gives H_kin = k_u E_kin; curl b = k_b b gives H_mag = E_mag / k_b."""
    n, ku, kb = 32, 2, 3
    p = M.Problem("MI355X", nx=n, B_field=True, T=np.float64)
    M.SetUpProblemIC(p, *_abc(ku, 0.7, -0.4, 1.1, n), *_abc(kb, 0.3, 0.9, -0.5, n))
    d = M.Diagnostics(p)
    assert math.isclose(d["H_kin"], ku * d["E_kin"], rel_tol=1e-12)
    assert math.isclose(d["H_mag"], d["E_mag"] / kb, rel_tol=1e-12)
    assert d["max_div_u"] < 1e-12 and d["max_div_b"] < 1e-12


def test_get_sol_modes_reads_the_state():
    o, p = pair(32, T=np.float64, kmax=8.0)
    M.stepforward(p, 2)
    rng = np.random.default_rng(3)
    idx = np.stack([rng.integers(0, 17, 500), rng.integers(0, 32, 500), rng.integers(0, 32, 500)], 1)
    got = M.get_sol_modes(p, idx)
    s = p.sol
    assert np.array_equal(got, s[:, idx[:, 2], idx[:, 1], idx[:, 0]].T)
    with pytest.raises(M.MHDError):
        M.get_sol_modes(p, np.array([[17, 0, 0]]))
