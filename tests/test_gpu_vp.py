"""Volume penalisation (Problem(...; VP_method=true)) on the MI355X against the oracle, through the
C ABI: VP_UᵢUpdate!/VP_BᵢUpdate! inside every right-hand side (src/Solver/VPSolver.jl:21-56) and
the divergence corrections TimeIntegrator! applies around every step (src/integrator.jl:84-87,
118-121).  Float64 <= 1e-12 relative L2, Float32 <= 1e-5 against the Float32 oracle."""

import math
import sys

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import ic
from oracle import OracleProblem
from oracle.integrator import rel_l2, run_steps, time_integrator
from oracle.rhs import calcN as o_calcN
from oracle.vp import set_vp

from test_gpu_parity import _calcN_gpu

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def vp_pair(n=32, T=np.float64, physics="MHD", dt=2e-3, nu=0.01, eta=0.02, kmax=6.0):
    B = physics == "MHD"
    fields = ic.solenoidal_modes(4, k0=2.0, kmax=kmax, mhd=B).real_fields(n, nf=6 if B else 3, dtype=np.float64)
    o = OracleProblem(nx=n, nu=nu, eta=eta, B_field=B, T=T, dt=dt, mode="fused")
    o.set_ic(*fields)
    p = M.Problem("MI355X", nx=n, ν=nu, η=eta, B_field=B, T=T, dt=dt, VP_method=True)
    M.SetUpProblemIC(p, *fields)
    chi = M.Cylindrical_Mask_Function(p.grid, R2=0.6 * math.pi)
    shape = (n, n, n)
    X, Y, Z = ic.grid_coords(n)
    U0 = [0.5 * np.cos(Z), np.zeros(shape), 0.25 * np.ones(shape)]
    B0 = [np.zeros(shape), 0.3 * np.ones(shape), 0.2 * np.sin(X)] if B else None
    if B:
        M.SetUpProblemIC(p, U0x=U0[0], U0y=U0[1], U0z=U0[2], B0x=B0[0], B0y=B0[1], B0z=B0[2])
    else:
        M.SetUpProblemIC(p, U0x=U0[0], U0y=U0[1], U0z=U0[2])
    M.SetUpVP(p, chi=chi)
    set_vp(o, chi=chi, U0=U0, B0=B0)
    return o, p


@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_vp_rhs_matches_oracle(physics):
    say("vp rhs", physics)
    o, p = vp_pair(physics=physics, kmax=16.0)
    N_gpu, sol_after = _calcN_gpu(p, o.sol.copy())
    N_ref = np.zeros_like(o.sol)
    s_ref = o.sol.copy()
    o_calcN(N_ref, s_ref, 0.0, o)
    assert rel_l2(N_gpu, N_ref) < 1e-12
    assert np.array_equal(sol_after, s_ref)


@pytest.mark.parametrize("physics", ["MHD", "HD"])
@pytest.mark.parametrize("T,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_vp_steps_match_oracle(physics, T, tol):
    say("vp steps", physics, T.__name__)
    o, p = vp_pair(T=T, physics=physics)
    run_steps(o, 4)
    M.stepforward(p, 4)
    err = rel_l2(p.sol, o.sol)
    say(f"  4 steps rel L2 {err:.3e}")
    assert err < tol


def test_vp_time_integrator_applies_the_corrections():
    say("vp integrator")
    o, p = vp_pair(T=np.float64, physics="MHD")
    time_integrator(o, math.inf, 3, CFL_Coef=0.3)
    taken = M.TimeIntegrator(p, math.inf, 3, CFL_Coef=0.3)
    assert taken == 4
    assert math.isclose(p.clock.dt, float(o.dt), rel_tol=1e-13)
    assert rel_l2(p.sol, o.sol) < 1e-12
    # the corrections leave k . u = k . b = 0 on every mode (the Nyquist planes aside, whose
    # wavenumber is its own negative)
    g = o.grid
    sol = p.sol
    n = g.nkr - 1
    inner = np.ones(g.shape_c, bool)
    inner[:, :, n] = False
    inner[:, g.l.shape[1] // 2, :] = False
    inner[g.m.shape[0] // 2, :, :] = False
    for f0 in (0, 3):
        div = g.kr * sol[f0] + g.l * sol[f0 + 1] + g.m * sol[f0 + 2]
        assert np.max(np.abs(div[inner])) < 1e-10 * np.max(np.abs(sol[f0:f0 + 3]))


def test_vp_dealiased_steps_equal_full_steps():
    _, a = vp_pair(n=24, T=np.float32, physics="MHD", kmax=12.0)
    _, b = vp_pair(n=24, T=np.float32, physics="MHD", kmax=12.0)
    M.set_step_pruning(b, False)
    for k in (3, 2):
        M.stepforward(a, k)
        M.stepforward(b, k)
        assert np.array_equal(a.sol, b.sol), k


def test_vp_refusals():
    from mhdflows_amd.devmem import DeviceArray
    _, p = vp_pair(n=16, physics="MHD", kmax=4.0)
    N = DeviceArray(p.sol.shape, p.sol.dtype)
    with pytest.raises(M.MHDError):
        M.UiUpdate(p, N.ptr, "x")                      # the per-direction updates refuse VP
    with pytest.raises(M.MHDError):
        M.SetUpND(p, 0.1, *[np.ones((16, 16, 16))] * 3)
    q = M.Problem("MI355X", nx=16, B_field=True)
    with pytest.raises(M.MHDError):
        M.SetUpProblemIC(q, U0x=np.ones((16, 16, 16)))   # U0 needs VP_method
