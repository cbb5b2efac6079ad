"""Seeded large-scale solenoidal forcing (oracle restatement; test infrastructure only).

The reference's forcing hook is ``addforcing!`` -> ``params.calcF!(N, sol, t, ...)``
(``src/pgen.jl:231-234``).  Its closest module, A99 (``src/pgen/A99ForceDriving.jl:93-127``),
draws unseeded random numbers on every call, so the build defines a seeded,
time-independent forcing on the shell ``round(|k|) == kf``:

* enumerate integer vectors ``(p, q, r)`` with ``r`` outermost and ``p`` innermost over
  ``[-K, K]^3``, ``K = ceil(kf + 0.5)``; keep ``(kf-0.5)^2 <= p^2+q^2+r^2 < (kf+0.5)^2``
  and the canonical half ``p>0 or (p==0 and q>0) or (p==q==0 and r>0)``;
* for canonical mode ``i`` and component ``c``: ``phi = 2*pi*U(seed, 3*i+c)`` with
  ``U`` the top 53 bits of ``splitmix64(seed*0x9E3779B97F4A7C15 + 3*i+c+1)``;
  ``f_c = amp * exp(i*phi)``;
* Leray-project ``f`` against ``k = (p*2pi/Lx, q*2pi/Ly, r*2pi/Lz)``;
* ``F(k) = f``, ``F(-k) = conj(f)``; the right-hand side of the velocity gets
  ``nx*ny*nz * F`` on the half-spectrum entries with ``kx >= 0``.
"""

import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def uniform(seed, counter):
    x = (seed * 0x9E3779B97F4A7C15 + counter + 1) & MASK64
    return (splitmix64(x) >> 11) * (1.0 / 9007199254740992.0)


def forced_modes(kf):
    K = int(math.ceil(kf + 0.5))
    lo2, hi2 = (kf - 0.5) ** 2, (kf + 0.5) ** 2
    out = []
    for r in range(-K, K + 1):
        for q in range(-K, K + 1):
            for p in range(-K, K + 1):
                s = p * p + q * q + r * r
                if not (lo2 <= s < hi2):
                    continue
                if p > 0 or (p == 0 and q > 0) or (p == 0 and q == 0 and r > 0):
                    out.append((p, q, r))
    return out


class Forcing:
    def __init__(self, seed, kf=2.0, amp=1.0):
        self.seed, self.kf, self.amp = int(seed), float(kf), float(amp)

    def mode_table(self, grid):
        """List of ((p,q,r), complex 3-vector f) for the canonical half."""
        tab = []
        for i, (p, q, r) in enumerate(forced_modes(self.kf)):
            f = np.array([self.amp * np.exp(1j * 2 * math.pi * uniform(self.seed, 3 * i + c))
                          for c in range(3)], dtype=np.complex128)
            k = np.array([p * 2 * math.pi / grid.Lx, q * 2 * math.pi / grid.Ly,
                          r * 2 * math.pi / grid.Lz])
            f = f - k * (k @ f) / (k @ k)
            tab.append(((p, q, r), f))
        return tab

    def spectrum(self, grid):
        Fh = np.zeros((3,) + grid.shape_c, dtype=np.complex128)
        Ntot = grid.nx * grid.ny * grid.nz
        for (p, q, r), f in self.mode_table(grid):
            for sgn, val in ((1, f), (-1, np.conj(f))):
                kx, ky, kz = sgn * p, sgn * q, sgn * r
                if kx < 0:
                    continue
                Fh[:, kz % grid.nz, ky % grid.ny, kx] = Ntot * val
        return Fh
