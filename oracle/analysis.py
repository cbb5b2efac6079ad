"""Spectra and scale decomposition (oracle; test infrastructure only).

Restates ``spectralline(A::Array{T,3})`` (``src/utils/MHDAnalysis.jl:236-255``) and
``ScaleDecomposition(B1, B2, B3; kf)`` (``:55-82``) on the problem's grid: |k| is
``sqrt((kx^2 + ky^2) + kz^2)`` in T, shells are ``round(|k|)`` with rounding half to even
(Julia's ``round``), and the shell sums are accumulated in Float64 here (the reference sums in T).
"""

import numpy as np

from .rhs import c2r, r2c


def kabs(grid):
    T = grid.T.type
    kx = np.broadcast_to(grid.kr, grid.shape_c).astype(T)
    ky = np.broadcast_to(grid.l, grid.shape_c).astype(T)
    kz = np.broadcast_to(grid.m, grid.shape_c).astype(T)
    return np.sqrt((kx * kx + ky * ky) + kz * kz).astype(T)


def krmax(grid):
    T = grid.T.type
    return int(np.rint(T(np.max(kabs(grid)) + T(1))))


def spectralline(grid, A):
    """(Pk, count) of a real field A ([z][y][x], type T): Pk[r] = sum |A_k|^2 over round(|k|) = r."""
    Ak = r2c(np.asarray(A, dtype=grid.T)).astype(grid.C)
    r = np.rint(kabs(grid).astype(np.float64)).astype(np.int64)
    n = krmax(grid)
    keep = r < n
    p = (Ak.real.astype(np.float64) ** 2 + Ak.imag.astype(np.float64) ** 2)[keep]
    Pk = np.bincount(r[keep], weights=p, minlength=n)[:n]
    count = np.bincount(r[keep], minlength=n)[:n].astype(np.float64)
    return Pk, count


def scale_decomposition(grid, fields, k1, k2):
    """Real fields with only the modes k1 <= |k| <= k2 kept (R2C, mask, C2R in T)."""
    k = kabs(grid).astype(np.float64)
    K = (k >= min(k1, k2)) & (k <= max(k1, k2))
    out = []
    for A in fields:
        Ah = r2c(np.asarray(A, dtype=grid.T)).astype(grid.C)
        out.append(c2r(Ah * K, grid.shape_r).astype(grid.T))
    return out
