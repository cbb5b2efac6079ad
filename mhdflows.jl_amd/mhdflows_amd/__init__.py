"""MI355X-native hot path of MHDFlows.jl: 3D periodic incompressible HD/MHD
right-hand side and RK4/LSRK54 time stepping, driven through the C ABI of
libmhdflows_amd.so (include/mhdflows_amd.h)."""

from ._lib import MHDError, lib  # noqa: F401
from .problem import (  # noqa: F401
    ScaleDecomposition,
    SFr,
    DivVCorrection,
    DivBCorrection,
    spectra,
    spectralline,
    Restart,
    read_checkpoint,
    readMHDFlows,
    savefile,
    A99ForceDriving,
    N97ForceDriving,
    NDForceDriving,
    SetUpFk,
    SetUpForcingField,
    SetUpN97,
    SetUpND,
    BiUpdate,
    Diagnostics,
    HDcalcN_advection,
    MHDcalcN_advection,
    MHDFlowsProblem,
    ProbDiagnostic,
    Problem,
    SetUpProblemIC,
    SetUpProblemIC_modes,
    autotune,
    variants,
    TimeIntegrator,
    UiUpdate,
    calcN,
    getCFL,
    get_sol_modes,
    nothingfunction,
    refresh_vars,
    set_forcing,
    set_step_pruning,
    stepforward,
)

__all__ = [
    "Problem", "SetUpProblemIC", "SetUpProblemIC_modes", "stepforward", "TimeIntegrator", "getCFL", "calcN",
    "MHDcalcN_advection", "HDcalcN_advection", "UiUpdate", "BiUpdate", "ProbDiagnostic",
    "Diagnostics", "set_forcing", "get_sol_modes", "N97ForceDriving", "NDForceDriving",
    "A99ForceDriving", "SetUpN97", "SetUpND", "SetUpFk", "SetUpForcingField", "savefile", "Restart",
    "readMHDFlows", "read_checkpoint", "spectra", "spectralline", "ScaleDecomposition", "SFr", "DivVCorrection", "DivBCorrection", "refresh_vars", "MHDError", "MHDFlowsProblem", "nothingfunction",
]
