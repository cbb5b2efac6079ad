"""Volume-penalisation solver (oracle; test infrastructure only).

``vp_update`` restates ``VP_UᵢUpdate!`` / ``VP_BᵢUpdate!`` (``src/Solver/VPSolver.jl:21-56``) as
``UᵢUpdate!`` / ``BᵢUpdate!`` call them after the advection (``src/Solver/MHDSolver.jl:85-88,
160-163``, ``src/Solver/HDSolver.jl:76-79``): with ``eta = clock.dt*13/7`` evaluated in T
(``dt*13`` then ``/7``) and ``kak2 = k_a .* k^-2`` in T,

    N_a += -(delta(a, j) - k_j*kak2) * R2C(chi/eta * (v_j - V0_j))      for j = 1, 2, 3

with v = u, V0 = U0 (or b, B0).  ``set_vp`` installs chi, U0, B0 (MHDParams_VP's fields).

Restates ``DivBCorrection!`` / ``DivVCorrection!`` (``src/Solver/VPSolver.jl:60-139``, exported
as ``MHDFlows.DivBCorrection!`` / ``DivVCorrection!``, ``src/MHDFlows.jl:62-63``) on an
``OracleProblem``: with ``k`` the grid wavenumbers in T and ``k^-2 = invKrsq`` (0 at k = 0),

    phi = -im * (kx*Fx + ky*Fy + kz*Fz);  phi = phi * k^-2;  F_i -= im*k_i*phi

in T on the three spectral fields of ``sol``, then ``vars`` = C2R of all their modes
(``ldiv!(vars.bx, grid.rfftplan, deepcopy(bxh))``).
"""

import numpy as np

from .rhs import c2r, r2c


def set_vp(prob, chi=None, U0=None, B0=None):
    """Install the penalisation fields (zero where not given) on an OracleProblem."""
    g = prob.grid
    z = np.zeros(g.shape_r, dtype=g.T)
    prob.vp = {"chi": z if chi is None else np.asarray(chi, dtype=g.T),
               "U0": [z if U0 is None or U0[c] is None else np.asarray(U0[c], dtype=g.T) for c in range(3)],
               "B0": [z if B0 is None or B0[c] is None else np.asarray(B0[c], dtype=g.T) for c in range(3)]}


def vp_update(Nf, prob, a, field0):
    """Add the penalisation term of direction a to Nf (field0 = 0: u and U0, 3: b and B0)."""
    if prob.vp is None:
        return
    g = prob.grid
    T = g.T.type
    eta = T(T(prob.dt) * T(13)) / T(7)
    k = (g.kr, g.l, g.m)
    kak2 = (k[a] * g.invKrsq).astype(g.T)
    chi = prob.vp["chi"]
    W = prob.vp["U0"] if field0 == 0 else prob.vp["B0"]
    for j in range(3):
        X = ((chi / eta) * (prob.vars[field0 + j] - W[j])).astype(g.T)
        Xh = r2c(X).astype(g.C)
        coef = ((1 if a == j else 0) - k[j] * kak2).astype(g.T)
        Nf += (-coef * Xh).astype(g.C)


def div_correction(prob, which):
    """which = 0: DivVCorrection! (u, fields 0..2); which = 1: DivBCorrection! (b, fields 3..5)."""
    g = prob.grid
    f0 = 3 * which
    F = [prob.sol[f0 + c] for c in range(3)]
    kx = np.broadcast_to(g.kr, g.shape_c)
    ky = np.broadcast_to(g.l, g.shape_c)
    kz = np.broadcast_to(g.m, g.shape_c)
    s = (kx * F[0] + ky * F[1]) + kz * F[2]
    phi = (-1j * s).astype(prob.sol.dtype)
    phi = (phi * g.invKrsq).astype(prob.sol.dtype)
    for c, k in enumerate((kx, ky, kz)):
        prob.sol[f0 + c] = (F[c] - (1j * k) * phi).astype(prob.sol.dtype)
        prob.vars[f0 + c] = c2r(prob.sol[f0 + c].copy(), g.shape_r).astype(prob.vars.dtype)
