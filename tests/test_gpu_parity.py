"""Parity of the MI355X library with the CPU oracle (same seeded inputs, through the C ABI).

Tolerances: Float64 <= 1e-12 relative L2 (10 RK4 steps and single right-hand sides);
Float32 <= 1e-4 relative L2 after 100 RK4 steps (Taylor-Green 32^3).
"""

import math
import sys

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import ic
from mhdflows_amd.devmem import DeviceArray
from oracle import OracleProblem
from oracle.diagnostics import full_diagnostics
from oracle.forcing import Forcing
from oracle.integrator import rel_l2, run_steps, time_integrator
from oracle.rhs import calcN as o_calcN
from oracle.rhs import hd_u_update_literal, mhd_b_update_literal, mhd_u_update_literal

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def pair(n, T=np.float64, physics="MHD", stepper="RK4", nu=0.01, eta=0.02, dt=1e-3, fields=None,
         mode="fused", ny=None, nz=None, seed=5, kmax=None):
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    B = physics == "MHD"
    if fields is None:
        km = kmax if kmax is not None else max(2.0, min(n, ny, nz) / 3 - 1)
        ms = ic.solenoidal_modes(seed, k0=2.0, kmax=km, mhd=B)
        assert nx_ok(n, ny, nz)
        full = ms.real_fields(max(n, ny, nz), nf=6 if B else 3, dtype=np.float64)
        fields = [f for f in full] if (n == ny == nz) else _resample(ms, n, ny, nz, B)
    o = OracleProblem(nx=n, ny=ny, nz=nz, nu=nu, eta=eta, B_field=B, T=T, dt=dt, mode=mode, stepper=stepper)
    o.set_ic(*fields)
    p = M.Problem("MI355X", nx=n, ny=ny, nz=nz, ν=nu, η=eta, B_field=B, T=T, dt=dt, stepper=stepper)
    M.SetUpProblemIC(p, *fields)
    return o, p


def nx_ok(*ns):
    return all(n % 2 == 0 for n in ns)


def _resample(ms, nx, ny, nz, B):
    # evaluate the mode set directly on an anisotropic grid
    x = -math.pi + np.arange(nx) * 2 * math.pi / nx
    y = -math.pi + np.arange(ny) * 2 * math.pi / ny
    z = -math.pi + np.arange(nz) * 2 * math.pi / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    out = []
    amps = [ms.au] + ([ms.ab] if B else [])
    for A in amps:
        for c in range(3):
            f = np.zeros_like(X)
            for (p, q, r), a in zip(ms.kvec, A[:, c]):
                f += 2 * np.real(a * np.exp(1j * (p * X + q * Y + r * Z)))
            out.append(f)
    return out


@pytest.mark.parametrize("n", [8, 16, 24, 32, 48, 64, 96])
@pytest.mark.parametrize("T,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
def test_initial_condition_r2c(n, T, tol):
    # Float64 MHD at nx = 24 is refused (its row kernel is disabled): that case transforms
    # the three velocity fields of an HD problem instead
    physics = "HD" if (n == 24 and T == np.float64) else "MHD"
    say("ic r2c", n, T.__name__, physics)
    o, p = pair(n, T=T, physics=physics)
    assert rel_l2(p.sol, o.sol) < tol
    last = 3 if physics == "MHD" else 0
    assert rel_l2(p.vars.field(last), o.vars[last]) == 0.0


@pytest.mark.parametrize("shape", [(32, 16, 24), (16, 48, 32), (64, 32, 16)])
def test_initial_condition_anisotropic(shape):
    say("ic aniso", shape)
    o, p = pair(shape[0], ny=shape[1], nz=shape[2], kmax=2.5)
    assert rel_l2(p.sol, o.sol) < 1e-13


def _calcN_gpu(p, sol_host):
    sol_d = DeviceArray.from_host(sol_host)
    N_d = DeviceArray(sol_host.shape, sol_host.dtype)
    M.calcN(p, N_d.ptr, sol_d.ptr)
    return N_d.download(), sol_d.download()


FAULTED_24_MHD_FP64 = ("24^3 MHD Float64 row kernel faulted the GPU once (memory aperture violation) in "
                       "a spill-heavy build; no cause was shown, so this case stays off the GPU (the "
                       "library refuses the configuration)")


@pytest.mark.parametrize("physics", ["MHD", "HD"])
@pytest.mark.parametrize("n", [16, 24, 32, 48, 64])
def test_rhs_parity_fp64(physics, n):
    if n == 24 and physics == "MHD":
        pytest.skip(FAULTED_24_MHD_FP64)
    say("rhs fp64", physics, n)
    o, p = pair(n, physics=physics, kmax=n / 2.0)  # modes beyond the 1/3 band: the mask matters
    sol = o.sol.copy()
    N_gpu, sol_after = _calcN_gpu(p, sol)
    N_ref = np.zeros_like(o.sol)
    s_ref = o.sol.copy()
    o_calcN(N_ref, s_ref, 0.0, o)
    assert rel_l2(N_gpu, N_ref) < 1e-12
    assert np.array_equal(sol_after, s_ref)  # dealias! in place
    # vars refreshed to C2R of the dealiased input
    assert rel_l2(p.vars.ux, o.vars[0]) < 1e-13


@pytest.mark.parametrize("shape", [(32, 16, 24), (16, 48, 32)])
def test_rhs_parity_anisotropic(shape):
    say("rhs aniso", shape)
    o, p = pair(shape[0], ny=shape[1], nz=shape[2], kmax=2.5)
    N_gpu, _ = _calcN_gpu(p, o.sol.copy())
    N_ref = np.zeros_like(o.sol)
    o_calcN(N_ref, o.sol.copy(), 0.0, o)
    assert rel_l2(N_gpu, N_ref) < 1e-12


def test_advection_no_dealias_fp64():
    say("advection")
    o, p = pair(32, kmax=16.0)
    sol = o.sol.copy()
    sol_d = DeviceArray.from_host(sol)
    N_d = DeviceArray(sol.shape, sol.dtype)
    M.MHDcalcN_advection(p, N_d.ptr, sol_d.ptr)
    from oracle.rhs import mhd_advection_fused
    N_ref = np.zeros_like(sol)
    mhd_advection_fused(N_ref, sol.copy(), 0.0, o)
    assert rel_l2(N_d.download(), N_ref) < 1e-12
    assert np.array_equal(sol_d.download(), sol)  # no dealiasing


@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_component_updates(physics):
    say("component updates", physics)
    o, p = pair(32 if physics == "MHD" else 24, physics=physics, kmax=10.0)
    nf = o.nf
    N_d = DeviceArray(o.sol.shape, o.sol.dtype)
    N_d.upload(np.zeros_like(o.sol))
    N_ref = np.zeros_like(o.sol)
    for a, d in enumerate("xyz"):
        M.UiUpdate(p, N_d.ptr, d)
        (mhd_u_update_literal if physics == "MHD" else hd_u_update_literal)(N_ref, o.sol, 0.0, o, a)
        if physics == "MHD":
            M.BiUpdate(p, N_d.ptr, d)
            mhd_b_update_literal(N_ref, o.sol, 0.0, o, a)
    got = N_d.download()
    for f in range(nf):
        assert rel_l2(got[f], N_ref[f]) < 1e-12, f


def test_taylor_green_fp64_10_steps():
    say("TG fp64")
    u, b = ic.taylor_green(32, dtype=np.float64)
    o, p = pair(32, T=np.float64, nu=1e-2, eta=1e-2, dt=1e-3, fields=list(u) + list(b), mode="literal")
    run_steps(o, 10)
    M.stepforward(p, 10)
    assert rel_l2(p.sol, o.sol) < 1e-12
    for f in range(6):
        assert rel_l2(p.vars.field(f), o.vars[f]) < 1e-12
    assert p.clock.step == 10 and math.isclose(p.clock.t, float(o.t), rel_tol=1e-15)


def test_taylor_green_fp32_100_steps():
    say("TG fp32")
    u, b = ic.taylor_green(32, dtype=np.float64)
    fields = list(u) + list(b)
    o, p = pair(32, T=np.float32, nu=1e-2, eta=1e-2, dt=1e-3, fields=fields, mode="literal")
    run_steps(o, 100)
    M.stepforward(p, 100)
    err = rel_l2(p.sol, o.sol)
    say(f"  FP32 vs FP32 oracle after 100 steps: {err:.3e}")
    assert err < 1e-4
    # error of the FP32 run against the FP64 oracle, per field and for the energies
    o64 = OracleProblem(nx=32, nu=1e-2, eta=1e-2, B_field=True, T=np.float64, dt=1e-3, mode="fused")
    o64.set_ic(*fields)
    run_steps(o64, 100)
    for f in range(6):
        e = rel_l2(p.sol[f], o64.sol[f])
        say(f"  field {f}: FP32 GPU vs FP64 oracle rel L2 {e:.3e}")
        assert e < 1e-4
    d32 = M.Diagnostics(p)
    d64 = full_diagnostics(o64)
    for k in ("E_kin", "E_mag"):
        say(f"  {k}: {d32[k]:.9g} vs {d64[k]:.9g}")
        assert math.isclose(d32[k], d64[k], rel_tol=1e-5)
    assert d32["max_div_b"] < 1e-4


@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_lsrk54_fp64(physics):
    say("LSRK54", physics)
    o, p = pair(32, physics=physics, stepper="LSRK54", kmax=8.0, dt=2e-3)
    run_steps(o, 5)
    M.stepforward(p, 5)
    assert rel_l2(p.sol, o.sol) < 1e-12


def test_hd_rk4_fp64():
    say("HD RK4")
    o, p = pair(32, physics="HD", kmax=10.0)
    run_steps(o, 10)
    M.stepforward(p, 10)
    assert rel_l2(p.sol, o.sol) < 1e-12


def test_cfl_time_integrator():
    say("CFL integrator")
    o, p = pair(32, T=np.float64, nu=1e-3, eta=1e-3, kmax=6.0)
    time_integrator(o, 1e9, 4, CFL_Coef=0.3)
    taken = M.TimeIntegrator(p, math.inf, 4, CFL_Coef=0.3)
    assert taken == 5 and p.clock.step == 5
    assert math.isclose(p.clock.dt, float(o.dt), rel_tol=1e-14)
    assert rel_l2(p.sol, o.sol) < 1e-12
    # finite t0: loop stops on time
    o2, p2 = pair(16, T=np.float64, kmax=4.0)
    time_integrator(o2, 0.0035, 100, usr_dt=1e-3)
    taken = M.TimeIntegrator(p2, 0.0035, 100, usr_dt=1e-3)
    assert taken == o2.step == 4
    assert rel_l2(p2.sol, o2.sol) < 1e-12


def test_host_hooks_run_once_per_step():
    """increment!(diags) and prob.usr_func run after every stepforward! (integrator.jl:115,131-133);
    the per-step host loop they force gives the device loop's state."""
    say("host hooks")
    _, a = pair(32, T=np.float64, nu=1e-3, eta=1e-3, kmax=6.0)
    _, b = pair(32, T=np.float64, nu=1e-3, eta=1e-3, kmax=6.0)
    seen, energies = [], []
    a.usr_func.append(lambda prob: seen.append(prob.clock.step))
    taken = M.TimeIntegrator(a, math.inf, 4, CFL_Coef=0.3, diags=[lambda prob: energies.append(
        M.Diagnostics(prob)["E_kin"])])
    M.TimeIntegrator(b, math.inf, 4, CFL_Coef=0.3)
    assert taken == 5 and seen == [1, 2, 3, 4, 5] and len(energies) == 5
    assert energies[-1] == M.Diagnostics(a)["E_kin"]
    assert math.isclose(a.clock.t, b.clock.t, rel_tol=1e-14)
    assert rel_l2(a.sol, b.sol) < 1e-13


def test_user_cfl_function():
    """A user CFL_function(prob, t_diff; Coef) sets dt before every step (integrator.jl:53-54,107)."""
    say("user CFL")
    _, a = pair(16, T=np.float64, nu=2e-3, eta=1e-3, kmax=4.0)
    _, b = pair(16, T=np.float64, nu=2e-3, eta=1e-3, kmax=4.0)
    calls = []

    def my_cfl(prob, t_diff, Coef=0.3):
        calls.append((t_diff, Coef))
        prob.clock.dt = 5e-4

    M.TimeIntegrator(a, math.inf, 3, CFL_Coef=0.2, CFL_function=my_cfl)
    M.TimeIntegrator(b, math.inf, 3, usr_dt=5e-4)
    dl = 2 * math.pi / 16
    assert len(calls) == 4 and all(math.isclose(t, 0.2 * dl * dl / 2e-3) and c == 0.2 for t, c in calls)
    assert rel_l2(a.sol, b.sol) < 1e-14 and a.clock.step == 4


def test_cfl_fp32_matches_oracle():
    o, p = pair(32, T=np.float32, nu=1e-3, eta=1e-3, kmax=6.0)
    time_integrator(o, 1e9, 2, CFL_Coef=0.3)
    M.TimeIntegrator(p, math.inf, 2, CFL_Coef=0.3)
    # later dt values come from FP32 fields that differ from the oracle's at rounding level
    assert math.isclose(p.clock.dt, float(o.dt), rel_tol=1e-6)
    assert rel_l2(p.sol, o.sol) < 1e-5


def test_forcing_fp64():
    say("forcing")
    o, p = pair(32, kmax=6.0)
    o.forcing = Forcing(seed=11, kf=2.0, amp=0.3)
    M.set_forcing(p, 11, 2.0, 0.3)
    run_steps(o, 5)
    M.stepforward(p, 5)
    assert rel_l2(p.sol, o.sol) < 1e-12


def test_diagnostics_fp64():
    say("diagnostics")
    o, p = pair(32, kmax=8.0)
    run_steps(o, 3)
    M.stepforward(p, 3)
    d = M.Diagnostics(p)
    r = full_diagnostics(o)
    for k in ("E_kin", "E_mag", "H_kin", "H_mag", "H_cross"):
        assert math.isclose(d[k], r[k], rel_tol=1e-11, abs_tol=1e-11), (k, d[k], r[k])
    assert d["max_div_b"] < 1e-10 and d["max_div_u"] < 1e-10
    ke, me = M.ProbDiagnostic(p)
    assert ke == float(f"{r['E_kin']:.3g}") and me == float(f"{r['E_mag']:.3g}")


def test_nan_is_reported():
    o, p = pair(16, kmax=3.0)
    s = p.sol
    s[0, 1, 1, 1] = np.nan
    p.sol = s
    M.stepforward(p, 1)
    with pytest.raises(M.MHDError):
        M.ProbDiagnostic(p)


def test_graph_replay_matches_eager():
    say("graph vs eager")
    _, a = pair(32, T=np.float32, kmax=8.0)
    _, b = pair(32, T=np.float32, kmax=8.0)
    M.lib().mhd_use_graph(b._h, 0)
    for i in range(3):
        M.stepforward(a, 2)
        M.stepforward(b, 2)
        _ = a.sol  # eager host work between replays
        assert np.array_equal(a.sol, b.sol), i


@pytest.mark.parametrize("T,physics,stepper,n", [(np.float64, "MHD", "RK4", 32), (np.float32, "MHD", "RK4", 48),
                                                 (np.float64, "HD", "LSRK54", 32), (np.float32, "MHD", "LSRK54", 32)])
def test_step_pruning_gives_the_same_state(T, physics, stepper, n):
    """Steps before the last of a call update kept modes only; the state after every call
    (aliased modes included: the fields reach |k| = n/2, so the products fill them) is the
    same as with every step computed in full, bit for bit."""
    say("pruning", T.__name__, physics, stepper, n)
    o, a = pair(n, T=T, physics=physics, stepper=stepper, kmax=n / 2.0)
    _, b = pair(n, T=T, physics=physics, stepper=stepper, kmax=n / 2.0)
    M.set_step_pruning(b, False)
    for k in (5, 1, 3):
        M.stepforward(a, k)
        M.stepforward(b, k)
        assert np.array_equal(a.sol, b.sol), k
        # prob.vars (stored by each call's last, full step only) match as well
        for f in range(a.sol.shape[0]):
            assert np.array_equal(a.vars.field(f), b.vars.field(f)), (k, f)
    if T == np.float64 and stepper == "RK4":
        run_steps(o, 9)
        assert rel_l2(a.sol, o.sol) < 1e-12
    M.TimeIntegrator(a, math.inf, 4, CFL_Coef=0.3)
    M.TimeIntegrator(b, math.inf, 4, CFL_Coef=0.3)
    assert np.array_equal(a.sol, b.sol)
    for f in range(a.sol.shape[0]):
        assert np.array_equal(a.vars.field(f), b.vars.field(f))


@pytest.mark.parametrize("n", [24, 48, 96])
def test_rhs_parity_fp32_three_factor(n):
    say("rhs fp32", n)
    o, p = pair(n, T=np.float32, kmax=n / 2.0)
    N_gpu, _ = _calcN_gpu(p, o.sol.copy())
    N_ref = np.zeros_like(o.sol)
    o_calcN(N_ref, o.sol.copy(), 0.0, o)
    assert rel_l2(N_gpu, N_ref) < 2e-5


@pytest.mark.parametrize("var", [[1, 1, 1, 1, 1], [2, 2, 2, 2, 0], [0, 2, 1, 0, 1], [0, 0, 0, 0, 2],
                                 [1, 2, 0, 1, 3], [3, 3, 3, 3, 3]])
@pytest.mark.parametrize("T,n,tol", [(np.float32, 64, 2e-5), (np.float64, 32, 1e-12)])
def test_variants_parity(var, T, n, tol):
    say("variants", var, T.__name__, n)
    o, p = pair(n, T=T, kmax=n / 2.0)
    assert M.variants(p, var) == [v for v in var]
    run_steps(o, 2)
    M.stepforward(p, 2)
    assert rel_l2(p.sol, o.sol) < tol


def test_autotune_keeps_results():
    say("autotune")
    o, p = pair(64, T=np.float64, kmax=10.0)
    rep = M.autotune(p, 2)
    assert "xrow_fused" in rep
    run_steps(o, 2)
    M.stepforward(p, 2)
    assert rel_l2(p.sol, o.sol) < 1e-12


def test_fp32_rhs_128_against_oracle():
    say("rhs fp32 128")
    o, p = pair(128, T=np.float32, kmax=40.0, mode="fused")
    N_gpu, _ = _calcN_gpu(p, o.sol.copy())
    N_ref = np.zeros_like(o.sol)
    o_calcN(N_ref, o.sol.copy(), 0.0, o)
    assert rel_l2(N_gpu, N_ref) < 2e-5


def test_errors_raise():
    with pytest.raises(M.MHDError):
        M.Problem("MI355X", nx=33)
    with pytest.raises(M.MHDError):
        M.Problem("MI355X", nx=32, hbm_budget=1000)
    p = M.Problem("MI355X", nx=16, B_field=True)
    with pytest.raises(M.MHDError):
        M.TimeIntegrator(p, 1.0, 2, save=True)
    with pytest.raises(M.MHDError):
        M.TimeIntegrator(p, 1.0, 2, usr_dt=1e-3, CFL_function=lambda *a: None)
    with pytest.raises(M.MHDError):
        M.UiUpdate(p, 0, "w")
