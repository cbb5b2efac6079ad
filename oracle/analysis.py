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


# ---- structure functions (src/utils/TurbStatTool.jl) --------------------------------------
def cf(V):
    """``CF(V) = fftshift(real(ifft(abs.(fft(V)).^2)))`` (TurbStatTool.jl:64) of a real [z][y][x]
    array in its own precision (scipy.fft keeps complex64 for Float32)."""
    import scipy.fft as sfft
    V = np.asarray(V)
    X = sfft.fftn(V.astype(np.result_type(V.dtype, np.complex64)))
    P = (np.abs(X) ** 2).astype(V.dtype)
    return np.fft.fftshift(np.real(sfft.ifftn(P)).astype(V.dtype))


def sfc(V):
    """``SFC(V) = 2*(mean(V) .- CF(V))`` (TurbStatTool.jl:69), in the precision of V."""
    T = np.asarray(V).dtype.type
    return (T(2) * (np.mean(V, dtype=T) - cf(V))).astype(T)


def _lag_vectors(shape):
    """(idx, jdx, kdx) of every element of the fftshifted [z][y][x] cube: the reference's
    ``i - div(Nx, 2)`` with i the 1-based index along x (likewise y, z) (:100-103, :140-143)."""
    nz, ny, nx = shape
    k = np.arange(1, nz + 1) - nz // 2
    j = np.arange(1, ny + 1) - ny // 2
    i = np.arange(1, nx + 1) - nx // 2
    return np.meshgrid(k, j, i, indexing="ij")[::-1]


def sf2_1d(vx, vy, vz):
    """``SF₂1D(Vx, Vz, Vy)`` (TurbStatTool.jl:88-117): (SFVr, Mask), SFVr = bin sums / counts over
    kk = round(|(idx, jdx, kdx)|) >= 1, 2R bins with R = ceil(|(nx/2, ny/2, nz/2)|)."""
    SFV = sfc(vx) + sfc(vy) + sfc(vz)
    nz, ny, nx = SFV.shape
    R = int(np.ceil(np.sqrt(float((nx // 2) ** 2 + (ny // 2) ** 2 + (nz // 2) ** 2))))
    idx, jdx, kdx = _lag_vectors(SFV.shape)
    kk = np.rint(np.sqrt((idx * idx + jdx * jdx + kdx * kdx).astype(np.float64))).astype(np.int64)
    sel = kk > 0
    sums = np.bincount(kk[sel] - 1, weights=SFV[sel].astype(np.float64), minlength=2 * R)[:2 * R]
    mask = np.bincount(kk[sel] - 1, minlength=2 * R)[:2 * R].astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        return sums / mask, mask


def sfr_global(vx, vy, vz, bx, by, bz):
    """``SFr_Global(Vx, Vz, Vy, bx, by, bz)`` (TurbStatTool.jl:119-160): (SFVr, Mask) of shape
    (R+1, R+1) indexed [rpar, rperp], theta the angle between the mean field and (idx, jdx, kdx)
    computed in T as ``t(...)`` (:39-48) does."""
    T = np.asarray(vx).dtype.type
    SFV = sfc(vx) + sfc(vy) + sfc(vz)
    nz, ny, nx = SFV.shape
    R = int(np.ceil(np.sqrt(float((nx // 2) ** 2 + (ny // 2) ** 2 + (nz // 2) ** 2))))
    mb = [np.mean(b, dtype=T) for b in (bx, by, bz)]
    idx, jdx, kdx = _lag_vectors(SFV.shape)
    ss = idx * idx + jdx * jdx + kdx * kdx
    sel = ss > 0
    kk = np.sqrt(ss[sel].astype(np.float64))
    dot = mb[0] * idx[sel].astype(T) + mb[1] * jdx[sel].astype(T) + mb[2] * kdx[sel].astype(T)
    x = (dot / np.sqrt((mb[0] * mb[0] + mb[1] * mb[1] + mb[2] * mb[2]) * ss[sel].astype(T))).astype(T)
    th = np.arccos(np.clip(x, T(-1), T(1))).astype(T)
    rpar = np.rint(np.abs(kk * np.cos(th).astype(np.float64))).astype(np.int64)
    rperp = np.rint(np.abs(kk * np.sin(th).astype(np.float64))).astype(np.int64)
    n = R + 1
    flat = rpar * n + rperp
    sums = np.bincount(flat, weights=SFV[sel].astype(np.float64), minlength=n * n)[:n * n].reshape(n, n)
    mask = np.bincount(flat, minlength=n * n)[:n * n].reshape(n, n).astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        return sums / mask, mask
