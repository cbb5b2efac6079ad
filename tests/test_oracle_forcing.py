"""Pins of the oracle's forcing modules (CPU only): N97 against its closed form, negative damping
against its injected power, A99 against its construction (solenoidal unit directions, zero kx = 0
plane, amplitude scaling) and the uniformity of its counter-based draws."""

import math

import numpy as np
import pytest

from mhdflows_amd import ic
from oracle import OracleProblem
from oracle.forcing import A99, N97, NegativeDamping, counter_uniforms
from oracle.rhs import c2r, calcN


def _prob(n=16, T=np.float64, forcing=None, seed=4, kmax=4.0):
    p = OracleProblem(nx=n, nu=0.01, eta=0.02, B_field=True, T=T, dt=1e-3, mode="fused", forcing=forcing)
    p.set_ic(*ic.solenoidal_modes(seed, k0=2.0, kmax=kmax).real_fields(n, dtype=np.float64))
    return p


def _added(p):
    """N with the problem's forcing minus N without it (same state)."""
    q = p.copy()
    q.forcing = None
    N1, N0 = np.zeros_like(p.sol), np.zeros_like(q.sol)
    calcN(N1, p.sol.copy(), 0.0, p)
    calcN(N0, q.sol.copy(), 0.0, q)
    return N1 - N0


def test_n97_closed_form_and_support():
    n, F0, kf = 16, 0.7, 2
    p = _prob(n, forcing=N97(F0=F0, kf=kf))
    F = _added(p)
    X, Y, Z = ic.grid_coords(n)
    s, c = np.sin, np.cos
    assert np.max(np.abs(c2r(F[0], (n, n, n)) - F0 * s(kf * X) * c(kf * Y) * c(kf * Z))) < 1e-13
    assert np.max(np.abs(c2r(F[1], (n, n, n)) + F0 * c(kf * X) * s(kf * Y) * c(kf * Z))) < 1e-13
    assert np.max(np.abs(F[2:])) == 0.0
    g = p.grid
    assert np.max(np.abs(g.kr * F[0] + g.l * F[1] + g.m * F[2])) < 1e-9      # solenoidal
    big = np.argwhere(np.abs(F[0]) > 1e-9 * np.max(np.abs(F[0])))
    kz = np.where(big[:, 0] < n // 2, big[:, 0], big[:, 0] - n)
    ky = np.where(big[:, 1] < n // 2, big[:, 1], big[:, 1] - n)
    assert set(np.abs(kz)) == {kf} and set(np.abs(ky)) == {kf} and set(big[:, 2]) == {kf}


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_negative_damping_injects_power_P(sign):
    """With f >= 0 every point gains u_i F_i = A f_i u_i^2 >= 0, so the forcing's power
    sum u.F dV equals A sum |u_i^2 f_i| dV = P (a negative P damps with the same magnitude)."""
    n, P = 16, 0.37 * sign
    rng = np.random.default_rng(2)
    f = [rng.uniform(0.1, 2.0, (n, n, n)) for _ in range(3)]
    p = _prob(n, forcing=NegativeDamping(P, *f))
    F = _added(p)
    u = p.vars[0:3]
    Fr = np.stack([c2r(F[i], (n, n, n)) for i in range(3)])
    dV = float(p.grid.dx) ** 3
    assert math.isclose(float(np.sum(u * Fr)) * dV, P, rel_tol=1e-12)
    # mixed-sign f: the same amplitude A, power A sum f u^2 dV < |P|
    f[0] = f[0] - 1.0
    q = _prob(n, forcing=NegativeDamping(abs(P), *f))
    Fq = np.stack([c2r(_added(q)[i], (n, n, n)) for i in range(3)])
    A = q.forcing.amplitude(q)
    assert math.isclose(float(np.sum(q.vars[0:3] * Fq)) * dV,
                        A * sum(float(np.sum(f[i] * q.vars[i] ** 2)) for i in range(3)) * dV, rel_tol=1e-12)
    assert abs(float(np.sum(q.vars[0:3] * Fq)) * dV) < abs(P)


def test_a99_construction():
    n = 16
    a = A99(seed=3, kf=2.0, P=1.0).setup(OracleProblem(nx=n, B_field=True, T=np.float64).grid)
    g = OracleProblem(nx=n, B_field=True, T=np.float64).grid
    F1, F2 = a.force(g, 12)
    live = a.Fk != 0
    assert not np.any(a.Fk[:, :, 0]) and np.all(a.Fk[:, :, 1:] >= 0)
    for F in (F1, F2):   # solenoidal
        assert np.max(np.abs(g.kr * F[0] + g.l * F[1] + g.m * F[2])) < 1e-12 * np.max(np.abs(F))
    # e1, e2 unit and orthogonal to each other and to k wherever k_perp > 0
    e1 = np.stack(a.e1)[:, live]
    e2 = np.stack(a.e2)[:, live]
    assert np.allclose(np.sum(e1 * e1, 0), 1, atol=1e-14) and np.allclose(np.sum(e2 * e2, 0), 1, atol=1e-14)
    assert np.max(np.abs(np.sum(e1 * e2, 0))) < 1e-14
    # |F1| = A Fk |gi| |e1| and |F2| = A Fk |gj| per live mode, with |e^{i theta}| = 1
    w1 = np.sqrt(np.sum(np.abs(F1) ** 2, 0))[live] / a.Fk[live]
    w2 = np.sqrt(np.sum(np.abs(F2) ** 2, 0))[live] / a.Fk[live]
    assert np.all(np.isfinite(w1)) and np.all(w1 >= 0) and np.all(w2 >= 0)
    # amplitude scales with sqrt(P)
    b = A99(seed=3, kf=2.0, P=4.0).setup(g)
    assert np.allclose(b.Fk, 2 * a.Fk, rtol=1e-14, atol=0)
    # different calls draw different forcing; the same call the same
    G1, _ = a.force(g, 13)
    H1, _ = a.force(g, 12)
    assert np.array_equal(H1, F1) and not np.allclose(G1, F1)


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_counter_uniforms_are_uniform_and_independent(T):
    idx = np.arange(200000, dtype=np.uint64)
    u = counter_uniforms(11, 7, idx, T).astype(np.float64)
    v = counter_uniforms(11, 8, idx, T).astype(np.float64)
    assert u.min() >= 0 and u.max() < 1
    # Kolmogorov-Smirnov distance to U(0, 1) and correlation between counters
    us = np.sort(u)
    ks = np.max(np.abs(us - (np.arange(1, us.size + 1) / us.size)))
    assert ks < 2.0 / math.sqrt(us.size)
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.01
    assert abs(np.mean(np.exp(2j * np.pi * u))) < 0.01
