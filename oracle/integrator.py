"""``TimeIntegrator!`` hot loop (``src/integrator.jl:31-156``), oracle; test infrastructure only.

Hooks that are off for plain HD/MHD (VP, dye, shear, save, dashboard) are omitted.
"""

import numpy as np

from .diagnostics import cfl_dt, t_diff as _t_diff
from .steppers import step


def time_integrator(prob, t0, N0, usr_dt=0.0, CFL_Coef=0.25):
    """Runs ``N0+1`` steps at most (``while N0 >= step && t0 >= t``, ``integrator.jl:104``)."""
    td = _t_diff(prob, CFL_Coef)
    prob.step = 0                                              # :76
    usr = usr_dt != 0.0
    if usr:
        prob.dt = prob.grid.T.type(usr_dt)                     # :80-82
    if prob.vp is not None:                                    # :84-87
        _vp_corrections(prob)
    while N0 >= prob.step and t0 >= prob.t:
        if not usr:
            prob.dt = cfl_dt(prob, td, CFL_Coef)               # :106-109
        step(prob)                                             # :112-113
        if prob.vp is not None:                                # :118-121
            _vp_corrections(prob)


def _vp_corrections(prob):
    """DivVCorrection!, then DivBCorrection! for MHD (the VP branch of TimeIntegrator!)."""
    from .vp import div_correction
    div_correction(prob, 0)
    if prob.physics == "MHD":
        div_correction(prob, 1)


def run_steps(prob, nsteps):
    """Call the stepper exactly ``nsteps`` times with the current ``dt``."""
    for _ in range(nsteps):
        step(prob)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel().astype(np.complex128))
    return float(np.linalg.norm((a.astype(np.complex128) - b.astype(np.complex128)).ravel()) / den)
