"""CFL and diagnostics (oracle; test infrastructure only).

Reductions are accumulated in Float64 (the reference reduces in the problem type T
at ``src/utils/UserInterface.jl:29``; the build keeps these reductions in FP64).
"""

import math

import numpy as np

from .rhs import c2r, r2c


def cfl_dt(prob, t_diff, coef):
    """``getCFL!`` (``src/integrator.jl:158-198``) for HD/MHD; returns dt in T.

    ``vmax = sqrt(max_a max_x u_a^2)`` (and likewise for b) in T, then
    ``dt = min(Coef*dl/vmax, t_diff)`` in Float64, stored into the T clock.
    """
    T = prob.grid.T.type
    v = prob.vars
    vmax = T(math.sqrt(max(float(np.max(v[a] * v[a])) for a in range(3))))
    if prob.physics == "MHD":
        vamax = T(math.sqrt(max(float(np.max(v[3 + a] * v[3 + a])) for a in range(3))))
        vmax = max(vmax, vamax)
    g = prob.grid
    dl = min(g.dx, g.dy, g.dz)
    with np.errstate(divide="ignore"):
        cfl = coef * float(dl) / float(vmax) if float(vmax) != 0 else math.inf
    return T(min(cfl, t_diff))


def t_diff(prob, coef):
    """Diffusive time-step bound (``src/integrator.jl:60-72``), Float64."""
    if prob.physics == "MHD":
        vi = max(prob.nu, prob.eta)
        nv = max(prob.nnu, prob.neta)
    else:
        vi, nv = prob.nu, prob.nnu
    g = prob.grid
    dl = min(g.dx, g.dy, g.dz)          # type T
    if vi == 0:
        return math.inf
    p = dl ** nv if nv > 1 else dl * dl  # evaluated in T, then promoted (integrator.jl:72)
    return coef * float(p) / vi


def dV(grid):
    return float(grid.dx) * float(grid.dy) * float(grid.dz)


def energies(prob):
    """``ProbDiagnostic`` (``UserInterface.jl:65-86``) without display rounding: E = Σ|f|²dV."""
    v = prob.vars.astype(np.float64)
    ek = float(np.sum(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)) * dV(prob.grid)
    em = float(np.sum(v[3] ** 2 + v[4] ** 2 + v[5] ** 2)) * dV(prob.grid) if prob.physics == "MHD" else 0.0
    if math.isnan(ek) or math.isnan(em):
        raise FloatingPointError("detected NaN! Quit the simulation right now.")
    return ek, em


def _spectral_ops(grid, fields):
    """R2C of real fields in FP64 and the wavenumber triple (FP64)."""
    Fh = r2c(np.asarray(fields, dtype=np.float64))
    k = (grid.kr.astype(np.float64), grid.l.astype(np.float64), grid.m.astype(np.float64))
    return Fh, k


def divergence(grid, f3):
    """``Div`` (``src/utils/VectorCalculus.jl:72-94``): C2R(i k·F̂), evaluated in FP64."""
    Fh, k = _spectral_ops(grid, f3)
    return c2r(1j * (k[0] * Fh[0] + k[1] * Fh[1] + k[2] * Fh[2]), grid.shape_r)


def curl(grid, f3):
    """``Curl`` (``src/utils/VectorCalculus.jl:24-51``) in FP64."""
    Fh, k = _spectral_ops(grid, f3)
    s = grid.shape_r
    return np.stack([c2r(1j * (k[1] * Fh[2] - k[2] * Fh[1]), s),
                     c2r(1j * (k[2] * Fh[0] - k[0] * Fh[2]), s),
                     c2r(1j * (k[0] * Fh[1] - k[1] * Fh[0]), s)])


def vector_potential(grid, b3):
    """``VectorPotential`` (``src/utils/MHDAnalysis.jl:138-174``), Coulomb gauge, FP64."""
    Bh, k = _spectral_ops(grid, b3)
    K = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    inv = np.where(K == 0, 0.0, 1.0 / np.where(K == 0, 1.0, K))   # invKrsq[1,1,1] = 0
    s = grid.shape_r
    return np.stack([c2r(1j * (k[1] * Bh[2] - k[2] * Bh[1]) * inv, s),
                     c2r(1j * (k[2] * Bh[0] - k[0] * Bh[2]) * inv, s),
                     c2r(1j * (k[0] * Bh[1] - k[1] * Bh[0]) * inv, s)])


def full_diagnostics(prob):
    """FP64 E_kin, E_mag, kinetic/magnetic/cross helicity, max|∇·b|, max|∇·u| from ``vars``."""
    g = prob.grid
    v = prob.vars.astype(np.float64)
    d = dV(g)
    out = {}
    out["E_kin"] = float(np.sum(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)) * d
    wu = curl(g, v[0:3])
    out["H_kin"] = float(np.sum(v[0] * wu[0] + v[1] * wu[1] + v[2] * wu[2])) * d   # h_k, MHDAnalysis.jl:94-101
    out["max_div_u"] = float(np.max(np.abs(divergence(g, v[0:3]))))
    if prob.physics == "MHD":
        out["E_mag"] = float(np.sum(v[3] ** 2 + v[4] ** 2 + v[5] ** 2)) * d
        A = vector_potential(g, v[3:6])
        out["H_mag"] = float(np.sum(A[0] * v[3] + A[1] * v[4] + A[2] * v[5])) * d  # h_m, :113-117
        out["H_cross"] = float(np.sum(v[0] * v[3] + v[1] * v[4] + v[2] * v[5])) * d  # Dotproduct
        out["max_div_b"] = float(np.max(np.abs(divergence(g, v[3:6]))))
    else:
        out["E_mag"] = out["H_mag"] = out["H_cross"] = out["max_div_b"] = 0.0
    return out
