"""Time steppers (oracle; test infrastructure only).

RK4 follows FourierFlows' ``RK4TimeStepper`` as restated in-repo at
``src/DyeModule.jl:62-91`` (``RK4substeps!``, ``substepsol!``, ``RK4update!``);
``addlinearterm!`` with ``L = 0`` (``src/Problems.jl:16``) adds ``0*sol``.
LSRK54 is the five-stage, fourth-order, 2N-storage scheme of Carpenter & Kennedy
(NASA TM-109112, 1994), the scheme behind FourierFlows' ``LSRK54TimeStepper``
(built at ``src/Problems.jl:128``).
"""

import numpy as np

from .rhs import calcN

# Carpenter & Kennedy (1994), RK4(3)5[2N] solution 3 coefficients.
LSRK54_A = (0.0,
            -567301805773.0 / 1357537059087.0,
            -2404267990393.0 / 2016746695238.0,
            -3550918686646.0 / 2091501179385.0,
            -1275806237668.0 / 842570457699.0)
LSRK54_B = (1432997174477.0 / 9575080441755.0,
            5161836677717.0 / 13612068292357.0,
            1720146321549.0 / 2090206949498.0,
            3134564353537.0 / 4481467310338.0,
            2277821191437.0 / 14882151754819.0)
LSRK54_C = (0.0,
            1432997174477.0 / 9575080441755.0,
            2526269341429.0 / 6820363183900.0,
            2006345519317.0 / 3224310063776.0,
            2802321613138.0 / 2924317926251.0)


def _addlinear(RHS, sol):
    RHS += 0 * sol  # addlinearterm!(RHS, L=0, sol)


def rk4_step(prob):
    """One ``stepforward!`` with RK4: four ``calcN!`` evaluations then the update."""
    sol = prob.sol
    T = prob.grid.T.type
    t, dt = prob.t, prob.dt
    half = T(dt / T(2))
    R1 = np.zeros_like(sol)  # stepper buffers are zero-initialised (@devzeros)
    R2 = np.zeros_like(sol)
    R3 = np.zeros_like(sol)
    R4 = np.zeros_like(sol)
    base = 8 * prob.step                          # draw counter of stochastic forcing: 8 step + stage
    prob.force_counter = base
    calcN(R1, sol, t, prob)                       # substep 1 (dealiases prob.sol in place)
    _addlinear(R1, sol)
    sol1 = sol + half * R1                        # substepsol!(sol1, sol, RHS1, dt/2)
    prob.force_counter = base + 1
    calcN(R2, sol1, T(t + half), prob)
    _addlinear(R2, sol1)
    sol1 = sol + half * R2
    prob.force_counter = base + 2
    calcN(R3, sol1, T(t + half), prob)
    _addlinear(R3, sol1)
    sol1 = sol + dt * R3
    prob.force_counter = base + 3
    calcN(R4, sol1, T(t + dt), prob)
    _addlinear(R4, sol1)
    sol += dt * (R1 / 6 + R2 / 3 + R3 / 3 + R4 / 6)   # RK4update!
    prob.t = T(t + dt)
    prob.step += 1


def lsrk54_step(prob):
    """One ``stepforward!`` with LSRK54 (each stage dealiases ``sol`` in place)."""
    sol = prob.sol
    T = prob.grid.T.type
    t, dt = prob.t, prob.dt
    A = [T(a) for a in LSRK54_A]
    B = [T(b) for b in LSRK54_B]
    C = [T(c) for c in LSRK54_C]
    S = np.zeros_like(sol)
    RHS = np.zeros_like(sol)
    for i in range(5):
        prob.force_counter = 8 * prob.step + i
        calcN(RHS, sol, T(t + C[i] * dt), prob)
        _addlinear(RHS, sol)
        S[...] = A[i] * S + dt * RHS
        sol += B[i] * S
    prob.t = T(t + dt)
    prob.step += 1


def step(prob):
    (rk4_step if prob.stepper == "RK4" else lsrk54_step)(prob)
