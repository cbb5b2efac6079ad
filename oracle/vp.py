"""Divergence corrections of the volume-penalization solver (oracle; test infrastructure only).

Restates ``DivBCorrection!`` / ``DivVCorrection!`` (``src/Solver/VPSolver.jl:60-139``, exported
as ``MHDFlows.DivBCorrection!`` / ``DivVCorrection!``, ``src/MHDFlows.jl:62-63``) on an
``OracleProblem``: with ``k`` the grid wavenumbers in T and ``k^-2 = invKrsq`` (0 at k = 0),

    phi = -im * (kx*Fx + ky*Fy + kz*Fz);  phi = phi * k^-2;  F_i -= im*k_i*phi

in T on the three spectral fields of ``sol``, then ``vars`` = C2R of all their modes
(``ldiv!(vars.bx, grid.rfftplan, deepcopy(bxh))``).
"""

import numpy as np

from .rhs import c2r


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
