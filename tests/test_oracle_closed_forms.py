"""Closed-form and invariant pins of the oracle's right-hand side and steppers (CPU only)."""

import math

import numpy as np
import pytest

from mhdflows_amd import ic
from oracle import OracleProblem
from oracle.diagnostics import energies, full_diagnostics, cfl_dt, t_diff
from oracle.integrator import rel_l2, run_steps, time_integrator
from oracle.rhs import calcN, r2c
from oracle.steppers import LSRK54_A, LSRK54_B


def rk4_factor(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def lsrk54_poly_coeffs():
    """Stability polynomial of the 2N recursion S=A S+h λ y, y+=B S, as Taylor coefficients."""
    # represent y and S as polynomials in z = hλ (coefficient arrays, degree <= 5)
    y = np.zeros(6); y[0] = 1
    S = np.zeros(6)
    for A, B in zip(LSRK54_A, LSRK54_B):
        zy = np.roll(y, 1); zy[0] = 0
        S = A * S + zy
        y = y + B * S
    return y


def test_lsrk54_order_conditions():
    c = lsrk54_poly_coeffs()
    assert np.allclose(c[:5], [1, 1, 1 / 2, 1 / 6, 1 / 24], atol=1e-12)
    assert abs(c[5] - 1 / 120) > 1e-6  # fifth-order term differs (4th-order scheme)


@pytest.mark.parametrize("mode", ["literal", "fused"])
@pytest.mark.parametrize("stepper", ["RK4", "LSRK54"])
def test_single_mode_decay(mode, stepper):
    """u = A(0, sin x, 0), b = B(0, 0, sin x): nonlinear terms vanish, each mode decays by R(-ν k^2 dt)."""
    n, nu, eta, dt, ns = 8, 0.3, 0.7, 0.01, 10
    p = OracleProblem(nx=n, nu=nu, eta=eta, B_field=True, T=np.float64, dt=dt, mode=mode, stepper=stepper)
    X, Y, Z = ic.grid_coords(n)
    zero = np.zeros_like(X)
    p.set_ic(zero, 1.3 * np.sin(X), zero, zero, zero, 0.4 * np.sin(X))
    s0 = p.sol.copy()
    run_steps(p, ns)
    if stepper == "RK4":
        fu, fb = rk4_factor(-nu * dt) ** ns, rk4_factor(-eta * dt) ** ns
    else:
        c = lsrk54_poly_coeffs()
        fu = sum(c[i] * (-nu * dt) ** i for i in range(6)) ** ns
        fb = sum(c[i] * (-eta * dt) ** i for i in range(6)) ** ns
    assert rel_l2(p.sol[1], fu * s0[1]) < 1e-13
    assert rel_l2(p.sol[5], fb * s0[5]) < 1e-13
    assert np.max(np.abs(p.sol[[0, 2, 3, 4]])) < 1e-12
    assert math.isclose(float(p.t), ns * dt, rel_tol=1e-14)


AXES = [(d, q) for d in range(3) for q in range(3) if q != d]


@pytest.mark.parametrize("mode", ["literal", "fused"])
@pytest.mark.parametrize("d,q", AXES)
def test_linear_alfven_wave(mode, d, q):
    """Uniform B0 along axis d, transverse perturbation along q varying with x_d: exact linear
    dynamics d/dt[û,b̂] = [[-νk², i k B0],[i k B0, -ηk²]][û,b̂]; RK4 = 4th-order matrix Taylor polynomial."""
    n, nu, eta, B0, dt, ns = 16, 0.05, 0.02, 1.7, 0.02, 10
    p = OracleProblem(nx=n, nu=nu, eta=eta, B_field=True, T=np.float64, dt=dt, mode=mode)
    C = ic.grid_coords(n)
    xd = C[d]
    zero = np.zeros_like(xd)
    u = [zero.copy() for _ in range(3)]
    b = [zero.copy() for _ in range(3)]
    u[q] = 0.1 * np.sin(xd) + 0.05 * np.cos(2 * xd)
    b[q] = 0.07 * np.cos(xd)
    b[d] = B0 + zero
    p.set_ic(*u, *b)
    s0 = p.sol.copy()
    run_steps(p, ns)
    kgrid = [p.grid.kr, p.grid.l, p.grid.m][d]
    kd = np.broadcast_to(kgrid, p.grid.shape_c)
    uh0, bh0 = s0[q], s0[3 + q]
    expect_u = np.zeros_like(uh0)
    expect_b = np.zeros_like(bh0)
    for idx in zip(*np.nonzero(np.abs(uh0) + np.abs(bh0) > 1e-9)):
        k = float(kd[idx])
        Amat = np.array([[-nu * k * k, 1j * k * B0], [1j * k * B0, -eta * k * k]])
        hA = dt * Amat
        P = np.eye(2) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
        v = np.linalg.matrix_power(P, ns) @ np.array([uh0[idx], bh0[idx]])
        expect_u[idx], expect_b[idx] = v
    assert rel_l2(p.sol[q], expect_u) < 1e-12
    assert rel_l2(p.sol[3 + q], expect_b) < 1e-12
    assert rel_l2(p.sol[3 + d], s0[3 + d]) < 1e-14  # uniform field untouched


def _random_problem(n, T, mode, nu=0.01, eta=0.02, seed=3, kmax=3.0, physics="MHD", stepper="RK4", dt=1e-3):
    p = OracleProblem(nx=n, nu=nu, eta=eta, B_field=(physics == "MHD"), T=T, dt=dt, mode=mode, stepper=stepper)
    ms = ic.solenoidal_modes(seed, k0=2.0, kmax=kmax, mhd=(physics == "MHD"))
    f = ms.real_fields(n, nf=p.nf, dtype=np.float64)
    p.set_ic(*f)
    return p


@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_rhs_is_solenoidal(physics):
    p = _random_problem(16, np.float64, "fused", physics=physics, kmax=6.0)
    N = np.zeros_like(p.sol)
    calcN(N, p.sol, 0.0, p)
    g = p.grid
    k = (g.kr, g.l, g.m)
    for base in ([0, 3] if physics == "MHD" else [0]):
        kN = k[0] * N[base] + k[1] * N[base + 1] + k[2] * N[base + 2]
        assert np.max(np.abs(kN)) < 1e-9 * np.max(np.abs(N[base]))


@pytest.mark.parametrize("physics,T,tol", [("MHD", np.float64, 1e-13), ("MHD", np.float32, 2e-5),
                                           ("HD", np.float64, 1e-13), ("HD", np.float32, 2e-5)])
def test_literal_equals_fused(physics, T, tol):
    a = _random_problem(16, T, "literal", physics=physics, kmax=7.0)
    b = a.copy()
    b.mode = "fused"
    run_steps(a, 3)
    run_steps(b, 3)
    assert rel_l2(a.sol, b.sol) < tol
    assert rel_l2(a.vars, b.vars) < tol


@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_ideal_invariants(physics):
    """ν=η=0: total energy (and cross helicity for MHD) conserved to O(dt^4) per step.

    At 32^3 the modes next to the asymmetric mask boundary stay empty over this run;
    at 16^3 they fill and the truncated system drifts by ~2e-7 independently of dt.
    """
    p = _random_problem(32, np.float64, "fused", nu=0.0, eta=0.0, physics=physics, kmax=2.5, dt=2e-3)
    d0 = full_diagnostics(p)
    run_steps(p, 20)
    # diagnostics read vars = stage-4 input; refresh from the final state for the comparison
    N = np.zeros_like(p.sol)
    calcN(N, p.sol.copy(), 0.0, p)
    d1 = full_diagnostics(p)
    e0 = d0["E_kin"] + d0["E_mag"]
    e1 = d1["E_kin"] + d1["E_mag"]
    assert abs(e1 - e0) / e0 < 1e-11
    if physics == "MHD":
        assert abs(d1["H_cross"] - d0["H_cross"]) < 1e-11 * e0
        assert d1["E_kin"] != pytest.approx(d0["E_kin"], rel=1e-6)  # energy is exchanged


def test_taylor_green_energies():
    n = 32
    p = OracleProblem(nx=n, B_field=True, T=np.float64)
    u, b = ic.taylor_green(n, dtype=np.float64)
    p.set_ic(*u, *b)
    ek, em = energies(p)
    assert math.isclose(ek, (2 * math.pi) ** 3 / 4, rel_tol=1e-12)
    assert math.isclose(em, (2 * math.pi) ** 3 * 3 / 16, rel_tol=1e-12)
    d = full_diagnostics(p)
    assert d["max_div_b"] < 1e-12 and d["max_div_u"] < 1e-12


def test_seeded_ic_properties():
    ms = ic.solenoidal_modes(7, k0=4.0, kmax=8.0)
    f = ms.real_fields(24, dtype=np.float64)
    assert np.allclose(np.mean(f[0:3] ** 2) * 3, 3.0, rtol=1e-12)
    assert np.allclose(np.mean(f[3:6] ** 2) * 3, 3.0, rtol=1e-12)
    assert abs(np.mean(np.sum(f[0:3] * f[3:6], axis=0))) < 1e-12
    # identical field on a finer grid (same seed, resolution independent)
    f2 = ic.solenoidal_modes(7, k0=4.0, kmax=8.0).real_fields(48, dtype=np.float64)
    assert np.allclose(f2[:, ::2, ::2, ::2], f, atol=1e-12)
    # spectrum_on_grid is the R2C of the sampled fields
    assert np.allclose(ms.spectrum_on_grid(24), r2c(f), atol=1e-8)


def test_cfl_and_integrator_step_count():
    p = OracleProblem(nx=8, nu=0.1, eta=0.05, B_field=True, T=np.float32)
    X, Y, Z = ic.grid_coords(8)
    zero = np.zeros_like(X)
    p.set_ic(2.0 * np.sin(Y), zero, zero, zero, 0.5 * np.cos(Z), zero)
    td = t_diff(p, 0.3)
    dl32 = np.float32(np.float32(2 * math.pi) / np.float32(8))
    dl = float(dl32)
    assert td == 0.3 * float(dl32 * dl32) / 0.1  # dl^2 in Float32, then promoted (integrator.jl:72)
    vmax = float(np.float32(math.sqrt(float(np.max(p.vars[0] ** 2)))))
    assert cfl_dt(p, td, 0.3) == np.float32(min(0.3 * dl / vmax, td))
    time_integrator(p, 1e9, 4, usr_dt=1e-3)
    assert p.step == 5  # N0+1 steps (integrator.jl:76,104)


@pytest.mark.parametrize("mode", ["literal", "fused"])
def test_hd_equals_mhd_velocity_with_zero_b(mode):
    """With b = 0 the MHD velocity right-hand side is -P(i k.(u u)) - nu k^2 u, the HD one
    (HDSolver.jl:58-74 vs MHDSolver.jl:63-94); the MHD sign of the u u term is fixed by
    cross-helicity conservation and the Alfven wave, so this pins the HD sign."""
    n = 16
    ms = ic.solenoidal_modes(9, k0=2.0, kmax=6.0, mhd=False)
    u = ms.real_fields(n, nf=3, dtype=np.float64)
    zero = np.zeros_like(u[0])
    hd = OracleProblem(nx=n, nu=0.02, B_field=False, T=np.float64, mode=mode)
    hd.set_ic(*u)
    mh = OracleProblem(nx=n, nu=0.02, eta=0.05, B_field=True, T=np.float64, mode=mode)
    mh.set_ic(*u, zero, zero, zero)
    Nh = np.zeros_like(hd.sol)
    Nm = np.zeros_like(mh.sol)
    calcN(Nh, hd.sol, 0.0, hd)
    calcN(Nm, mh.sol, 0.0, mh)
    assert rel_l2(Nh, Nm[0:3]) < 1e-14
    assert np.max(np.abs(Nm[3:6])) == 0.0
    assert np.max(np.abs(Nh)) > 1.0  # a non-trivial right-hand side


@pytest.mark.parametrize("physics", ["HD", "MHD"])
@pytest.mark.parametrize("mode", ["literal", "fused"])
def test_taylor_green_analytic_rhs(physics, mode):
    """u = (sin x cos y cos z, -cos x sin y cos z, 0), b = 0: the incompressible right-hand side
    is -(u.grad)u - grad p - nu k^2 u with p = (cos 2x + cos 2y)(2 + cos 2z)/16, i.e.
    N = (-sin 2x cos 2z / 8, -sin 2y cos 2z / 8, (cos 2x + cos 2y) sin 2z / 8) - 3 nu u."""
    n, nu = 32, 0.01
    p = OracleProblem(nx=n, nu=nu, eta=0.03, B_field=(physics == "MHD"), T=np.float64, mode=mode)
    u, _ = ic.taylor_green(n, dtype=np.float64)
    zero = np.zeros_like(u[0])
    p.set_ic(*u, *([zero] * 3 if physics == "MHD" else []))
    N = np.zeros_like(p.sol)
    calcN(N, p.sol, 0.0, p)
    X, Y, Z = ic.grid_coords(n)
    s, c = np.sin, np.cos
    expect = [-s(2 * X) * c(2 * Z) / 8, -s(2 * Y) * c(2 * Z) / 8, (c(2 * X) + c(2 * Y)) * s(2 * Z) / 8]
    for a in range(3):
        got = np.fft.irfftn(N[a], s=(n, n, n), axes=(0, 1, 2))
        assert np.max(np.abs(got - (expect[a] - 3 * nu * u[a]))) < 1e-13, a
    if physics == "MHD":
        assert np.max(np.abs(N[3:6])) == 0.0


def _abc(k, A, B, C, n):
    """Arnold-Beltrami-Childress field of wavenumber k: curl f = k f."""
    X, Y, Z = ic.grid_coords(n)
    s, c = np.sin, np.cos
    return [A * s(k * Z) + C * c(k * Y), B * s(k * X) + A * c(k * Z), C * s(k * Y) + B * c(k * X)]


def test_helicities_beltrami_closed_forms():
    """For curl u = k_u u: H_kin = sum u.(curl u) dV = k_u E_kin.  For curl b = k_b b the
    Coulomb-gauge potential is A = b / k_b, so H_mag = sum A.b dV = E_mag / k_b.  The signs of
    Curl (VectorCalculus.jl:24-51) and VectorPotential (MHDAnalysis.jl:138-174) and the dV
    of the volume integrals are all fixed by these."""
    n, ku, kb = 24, 2, 3
    u = _abc(ku, 0.7, -0.4, 1.1, n)
    b = _abc(kb, 0.3, 0.9, -0.5, n)
    p = OracleProblem(nx=n, B_field=True, T=np.float64)
    p.set_ic(*u, *b)
    d = full_diagnostics(p)
    assert math.isclose(d["H_kin"], ku * d["E_kin"], rel_tol=1e-12)
    assert math.isclose(d["H_mag"], d["E_mag"] / kb, rel_tol=1e-12)
    dV = (2 * math.pi / n) ** 3
    assert math.isclose(d["E_kin"], dV * float(np.sum(np.square(u))), rel_tol=1e-14)
    # a left-handed b (curl b = -k b) flips the sign of H_mag
    p2 = OracleProblem(nx=n, B_field=True, T=np.float64)
    X, Y, Z = ic.grid_coords(n)
    ul = [np.sin(ku * Z), np.cos(ku * Z), np.zeros_like(Z)]      # curl = +k u
    bl = [np.cos(kb * Z), np.sin(kb * Z), np.zeros_like(Z)]      # curl = -k b
    p2.set_ic(*ul, *bl)
    d2 = full_diagnostics(p2)
    assert math.isclose(d2["H_kin"], ku * d2["E_kin"], rel_tol=1e-12)
    assert math.isclose(d2["H_mag"], -d2["E_mag"] / kb, rel_tol=1e-12)
