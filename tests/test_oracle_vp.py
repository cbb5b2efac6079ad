"""Pins of the oracle's divergence corrections (DivVCorrection!/DivBCorrection!, CPU only):
the Helmholtz split they perform -- a gradient field is removed, a solenoidal one is kept, their
sum returns the solenoidal part, k . F = 0 afterwards, and a second pass changes nothing."""

import numpy as np
import pytest

from mhdflows_amd import ic
from oracle import OracleProblem
from oracle.rhs import c2r
from oracle.vp import div_correction


def _fields(n):
    X, Y, Z = ic.grid_coords(n)
    # gradient of phi = sin(2x) cos(y) + cos(3z)
    G = [2 * np.cos(2 * X) * np.cos(Y), -np.sin(2 * X) * np.sin(Y), -3 * np.sin(3 * Z)]
    # solenoidal: Taylor-Green-like plus a uniform part
    S = [np.sin(X) * np.cos(Y) * np.cos(Z) + 0.4, -np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X) - 0.2]
    return G, S


@pytest.mark.parametrize("which", [0, 1])
def test_helmholtz_split(which):
    n = 16
    G, S = _fields(n)
    zero = [np.zeros((n, n, n))] * 3
    for F, want in ((G, zero), (S, S), ([g + s for g, s in zip(G, S)], S)):
        p = OracleProblem(nx=n, B_field=True, T=np.float64)
        p.set_ic(*(F + zero) if which == 0 else (zero + F))
        div_correction(p, which)
        f0 = 3 * which
        for c in range(3):
            assert np.max(np.abs(p.vars[f0 + c] - want[c])) < 1e-13
            assert np.max(np.abs(c2r(p.sol[f0 + c], (n, n, n)) - p.vars[f0 + c])) < 1e-14
        g = p.grid
        div = g.kr * p.sol[f0] + g.l * p.sol[f0 + 1] + g.m * p.sol[f0 + 2]
        assert np.max(np.abs(div)) < 1e-10
        other = 3 - f0
        assert np.max(np.abs(p.sol[other:other + 3])) == 0.0     # the other triple is untouched


def test_second_pass_changes_nothing():
    n = 12
    rng = np.random.default_rng(0)
    p = OracleProblem(nx=n, B_field=True, T=np.float64)
    p.set_ic(*[rng.standard_normal((n, n, n)) for _ in range(6)])
    div_correction(p, 1)
    once = p.sol.copy()
    div_correction(p, 1)
    assert np.max(np.abs(p.sol - once)) < 1e-12 * np.max(np.abs(once))


# ---- the penalisation terms (VP_UᵢUpdate!, VP_BᵢUpdate!) -------------------------------------
from oracle.rhs import calcN
from oracle.vp import set_vp


def _vp_prob(n=16, physics="MHD", mode="fused", T=np.float64, seed=3):
    p = OracleProblem(nx=n, B_field=physics == "MHD", T=T, nu=0.01, eta=0.02, dt=2e-3, mode=mode)
    p.set_ic(*ic.solenoidal_modes(seed, k0=2.0, kmax=4.0, mhd=physics == "MHD").real_fields(
        n, nf=p.nf, dtype=np.float64))
    return p


def _N(p):
    N = np.zeros_like(p.sol)
    calcN(N, p.sol.copy(), 0.0, p)
    return N


def test_solid_everywhere_damps_a_single_mode_at_rate_one_over_eta():
    """chi = 1, U0 = 0, u = sin(y) x (one solenoidal mode, no advection), nu = 0: N_u = -u/eta with
    eta = dt*13/7 (VPSolver.jl:24), so the penalisation's sign and factor are fixed."""
    n = 16
    X, Y, Z = ic.grid_coords(n)
    zero = np.zeros_like(X)
    p = OracleProblem(nx=n, B_field=True, T=np.float64, nu=0.0, eta=0.0, dt=3e-3, mode="literal")
    p.set_ic(np.sin(Y), zero, zero, zero, zero, zero)
    set_vp(p, chi=np.ones_like(X))
    N = _N(p)
    eta = 3e-3 * 13 / 7
    assert np.max(np.abs(N[0] + p.sol[0] / eta)) < 1e-12 * np.max(np.abs(p.sol[0] / eta))
    assert np.max(np.abs(N[1:])) < 1e-12 * np.max(np.abs(N[0]))


@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_penalisation_vanishes_without_solid_or_slip(physics):
    p = _vp_prob(physics=physics)
    base = _N(p)
    rng = np.random.default_rng(1)
    set_vp(p, chi=np.zeros(p.grid.shape_r))                      # no solid: exactly no change
    assert np.array_equal(_N(p), base)
    from oracle.rhs import c2r
    v = [c2r(p.sol[f], p.grid.shape_r) for f in range(p.nf)]
    mask = (rng.random(p.grid.shape_r) > 0.5).astype(np.float64)
    set_vp(p, chi=mask, U0=v[0:3], B0=v[3:6] if physics == "MHD" else None)   # solid moving with the flow
    assert np.max(np.abs(_N(p) - base)) < 1e-12 * np.max(np.abs(base))


@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_penalisation_literal_equals_fused_and_is_solenoidal(physics):
    rng = np.random.default_rng(5)
    out = {}
    for mode in ("literal", "fused"):
        p = _vp_prob(physics=physics, mode=mode)
        shape = p.grid.shape_r
        set_vp(p, chi=(rng.random(shape) > 0.4).astype(np.float64) if mode == "literal" else out["chi"],
               U0=[0.3 * np.ones(shape), np.zeros(shape), np.zeros(shape)],
               B0=[np.zeros(shape), 0.2 * np.ones(shape), np.zeros(shape)])
        out["chi"] = p.vp["chi"]
        N = _N(p)
        q = p.copy()
        q.vp = None
        out[mode] = N - _N(q)
    assert np.max(np.abs(out["literal"] - out["fused"])) < 1e-12 * np.max(np.abs(out["fused"]))
    g = _vp_prob(physics=physics).grid
    V = out["fused"]
    for f0 in range(0, V.shape[0], 3):
        div = g.kr * V[f0] + g.l * V[f0 + 1] + g.m * V[f0 + 2]
        assert np.max(np.abs(div)) < 1e-10 * np.max(np.abs(V[f0:f0 + 3]))
