"""Shared definition of the precision run (1024^3 Float32 vs a 256^3 Float64 oracle run).

Incompressible MHD, L = 2pi, nu = eta = 2e-4, RK4 with a fixed dt = 5e-4 for 50 steps;
seeded solenoidal fields k^4 exp(-2 (k/4)^2), |k| <= 16 (so 256^3 represents them
exactly), plus the seeded static forcing on the |k| = 2 shell (seed 7, amplitude 1).
The error is measured on the modes a 256^3 grid keeps after dealiasing.
"""

import numpy as np

NU = ETA = 2e-4
DT = 5e-4
STEPS = 50
IC_SEED, K0, KMAX = 2024, 4.0, 16.0
F_SEED, KF, AMP = 7, 2.0, 1.0
N_REF = 256


def kept_signed(n):
    """Signed wavenumbers a grid of n points keeps on a full axis (the 1/3 box mask)."""
    lo = int(np.floor(n / 3.0 + 1e-12))
    return np.arange(-lo, lo)  # [-floor(n/3), floor(n/3) - 1]


def retained(sol, n):
    """sol[f, z, y, kx] of an n^3 grid restricted to the 256^3 retained modes, scaled to the
    256^3 normalisation (the FFT scale is the number of grid points)."""
    ks = kept_signed(N_REF)
    idx = np.where(ks >= 0, ks, n + ks)
    kx = np.arange(0, int(np.floor(N_REF / 3.0 + 1e-12)))
    out = sol[:, idx][:, :, idx][:, :, :, kx]
    return out.astype(np.complex128) * (float(N_REF) / n) ** 3
