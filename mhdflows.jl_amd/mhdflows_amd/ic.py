"""Seeded initial conditions for the MHD/HD workloads.

The reference takes real-space initial fields through ``SetUpProblemIC!``
(``src/utils/IC.jl:41-109``); its random generator ``DivFreeSpectraMap``
(``IC.jl:122-179``) is unseeded, so these generators are the build's own:

* ``taylor_green(n)``: the 3D MHD Taylor-Green vortex
  ``u = (sin x cos y cos z, -cos x sin y cos z, 0)``,
  ``b = 0.5 (cos x sin y sin z, sin x cos y sin z, -2 sin x sin y cos z)`` sampled on
  ``x_i = -L/2 + i L/n`` (FourierFlows' default origin).
* ``solenoidal_modes(seed, k0, kmax)``: complex-Gaussian Fourier modes on integer
  wavevectors ``0 < |k| <= kmax`` with shell spectrum ``E(k) ∝ k^4 exp(-2 (k/k0)^2)``,
  Leray-projected, ``b`` orthogonalised against ``u`` (cross-helicity 0), rescaled
  to a per-component rms of 1 for ``u`` and ``b``.  Modes are enumerated in a
  fixed order independent of the grid, so one seed gives the same field on every
  grid that resolves ``kmax``.
"""

import math

import numpy as np


def grid_coords(n, L=2 * math.pi):
    x = -L / 2 + np.arange(n) * (L / n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z


def taylor_green(n, L=2 * math.pi, dtype=np.float32):
    """Returns ``(u, b)`` as arrays of shape ``(3, n, n, n)`` in ``[z][y][x]`` order."""
    X, Y, Z = grid_coords(n, L)
    s, c = np.sin, np.cos
    u = np.stack([s(X) * c(Y) * c(Z), -c(X) * s(Y) * c(Z), np.zeros_like(X)])
    b = 0.5 * np.stack([c(X) * s(Y) * s(Z), s(X) * c(Y) * s(Z), -2 * s(X) * s(Y) * c(Z)])
    return u.astype(dtype), b.astype(dtype)


def _canonical_vectors(kmax):
    K = int(math.floor(kmax))
    out = []
    for r in range(-K, K + 1):
        for q in range(-K, K + 1):
            for p in range(-K, K + 1):
                s = p * p + q * q + r * r
                if s == 0 or s > kmax * kmax:
                    continue
                if p > 0 or (p == 0 and q > 0) or (p == 0 and q == 0 and r > 0):
                    out.append((p, q, r))
    return np.array(out, dtype=np.int64)


class ModeSet:
    """Sparse Fourier representation: ``f(x) = Σ_canonical 2 Re(a(k) e^{i k·x})``."""

    def __init__(self, kvec, au, ab, L=2 * math.pi):
        self.kvec = kvec          # (M, 3) integer wavevectors, canonical half
        self.au = au              # (M, 3) complex velocity amplitudes
        self.ab = ab              # (M, 3) complex magnetic amplitudes (or None)
        self.L = L

    @property
    def kmax(self):
        return float(np.sqrt((self.kvec ** 2).sum(1)).max())

    def half_spectrum(self, n, nf=6):
        """R2C-convention half spectrum ``(nf, n, n, n//2+1)`` complex128 (values ``n^3 a``)."""
        nk = n // 2 + 1
        out = np.zeros((nf, n, n, nk), np.complex128)
        N = float(n) ** 3
        amps = [self.au] + ([self.ab] if nf == 6 else [])
        for ci, A in enumerate(amps):
            for c in range(3):
                f = 3 * ci + c
                for sgn in (1, -1):
                    kx = sgn * self.kvec[:, 0]
                    ky = sgn * self.kvec[:, 1]
                    kz = sgn * self.kvec[:, 2]
                    v = A[:, c] if sgn == 1 else np.conj(A[:, c])
                    sel = kx >= 0
                    out[f, kz[sel] % n, ky[sel] % n, kx[sel]] = N * v[sel]
        return out

    def real_fields(self, n, nf=6, dtype=np.float32):
        """Real fields on an ``n^3`` grid (FP64 synthesis, cast to ``dtype``)."""
        import numpy.fft as npfft
        X = self.half_spectrum(n, nf)
        # the x0 = -L/2 origin multiplies each mode by exp(-i k·L/2) = (-1)^(p+q+r) for L=2π
        return npfft.irfftn(X * self._origin_phase(n), s=(n, n, n), axes=(1, 2, 3)).astype(dtype)

    def _origin_phase(self, n):
        k = np.fft.fftfreq(n, 1.0 / n)
        kr = np.arange(n // 2 + 1)
        sh = -self.L / 2 * (2 * math.pi / self.L)
        ph = (np.exp(1j * sh * k)[:, None, None] * np.exp(1j * sh * k)[None, :, None]
              * np.exp(1j * sh * kr)[None, None, :])
        return ph[None]

    def half_entries(self, n, nf=6):
        """Sparse half-spectrum entries of the fields sampled on an ``n^3`` grid:
        ``(idx, vals)`` with ``idx`` (M, 3) int32 array indices (kx, l, m) and ``vals``
        (M, nf) complex128 (R2C convention, origin ``-L/2``)."""
        N = float(n) ** 3
        amps = [self.au] + ([self.ab] if nf == 6 else [])
        idx, vals = [], []
        for sgn in (1, -1):
            k = sgn * self.kvec
            sel = k[:, 0] >= 0
            if sgn == -1:
                sel &= self.kvec[:, 0] == 0  # kx = 0 plane needs both members of the pair
            ks = k[sel]
            ph = np.exp(-1j * math.pi * ks.sum(1)) if abs(self.L - 2 * math.pi) < 1e-15 else \
                np.exp(-1j * (self.L / 2) * (2 * math.pi / self.L) * ks.sum(1))
            v = np.stack([(A[sel, c] if sgn == 1 else np.conj(A[sel, c])) for A in amps for c in range(3)], 1)
            idx.append(np.stack([ks[:, 0], ks[:, 1] % n, ks[:, 2] % n], 1))
            vals.append(N * v * ph[:, None])
        return np.concatenate(idx).astype(np.int32), np.concatenate(vals)

    def spectrum_on_grid(self, n, nf=6):
        """Half spectrum of the real fields sampled on ``x_i = -L/2 + i L/n`` (R2C convention)."""
        return self.half_spectrum(n, nf) * self._origin_phase(n)


def solenoidal_modes(seed, k0=4.0, kmax=16.0, mhd=True, L=2 * math.pi):
    kv = _canonical_vectors(kmax)
    M = len(kv)
    rng = np.random.Generator(np.random.PCG64(seed))
    g = rng.standard_normal((M, 2, 3, 2))  # (mode, u/b, component, re/im)
    kk = np.sqrt((kv.astype(np.float64) ** 2).sum(1))
    E = kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
    A = np.sqrt(E / (4 * math.pi * kk ** 2))

    def make(which):
        a = (g[:, which, :, 0] + 1j * g[:, which, :, 1]) / math.sqrt(2.0) * A[:, None]
        kf = kv.astype(np.float64)
        a = a - kf * ((kf * a).sum(1) / (kk ** 2))[:, None]
        return a

    au = make(0)
    ab = make(1) if mhd else None

    def inner(a, b):
        return float(2.0 * np.real((a * np.conj(b)).sum()))

    if mhd:
        ab = ab - (inner(ab, au) / inner(au, au)) * au
    au = au * math.sqrt(3.0 / inner(au, au))
    if mhd:
        ab = ab * math.sqrt(3.0 / inner(ab, ab))
    return ModeSet(kv, au, ab, L)
