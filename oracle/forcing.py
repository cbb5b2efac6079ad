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


# ---------------------------------------------------------------------------------------
# The reference's forcing modules (calcF! implementations), restated
# ---------------------------------------------------------------------------------------

def grid_axes_T(grid):
    """Cell coordinates x_i = -L/2 + i*dx (dx = L/n in T), rounded to T, as 1D arrays
    broadcast to (z, y, x) (FourierFlows ThreeDGrid; mirrored at src/utils/utils.jl:43,58-60)."""
    T = grid.T.type
    ax = []
    for n, L, d in ((grid.nx, grid.Lx, grid.dx), (grid.ny, grid.Ly, grid.dy), (grid.nz, grid.Lz, grid.dz)):
        ax.append((-L / 2 + np.arange(n) * float(d)).astype(T))
    x, y, z = ax
    return x[None, None, :], y[None, :, None], z[:, None, None]


class N97:
    """Taylor-Green dynamo forcing of Nore et al. 1997 (src/pgen/TaylorGreenDynamo.jl).

    ``SetUpN97!`` (:18-33): ``fx = F0 sin(kf x) cos(kf y) cos(kf z)``,
    ``fy = -F0 cos(kf x) sin(kf y) cos(kf z)`` evaluated in T on the grid, R2C in T;
    ``N97ForceDriving!`` (:12-16): ``N[ux] += fxh``, ``N[uy] += fyh`` (Complex{T} adds)."""

    def __init__(self, F0=1.0, kf=2):
        self.F0, self.kf = F0, kf

    def fields(self, grid):
        T = grid.T.type
        x, y, z = grid_axes_T(grid)
        kf, F0 = T(self.kf), T(self.F0)
        s, c = np.sin, np.cos
        fx = (F0 * s(kf * x) * c(kf * y) * c(kf * z)).astype(T)
        fy = (F0 * -c(kf * x) * s(kf * y) * c(kf * z)).astype(T)
        return fx, fy

    def apply(self, N, sol, t, prob):
        from .rhs import r2c
        fx, fy = self.fields(prob.grid)
        N[0] += r2c(fx).astype(prob.grid.C)
        N[1] += r2c(fy).astype(prob.grid.C)


class NegativeDamping:
    """Negative-damping forcing (src/pgen/NegativeDamping.jl).

    ``NDForceDriving!`` (:23-44): with u = vars (the real fields of the current stage input),
    ``A = P / (sum|ux^2 fx| + sum|uy^2 fy| + sum|uz^2 fz|) dV`` and ``N[u_i] += A R2C(f_i u_i)``.
    A is Float64 (P is a Float64 field, :10), so the add is promoted and rounded back to T.
    The sums are accumulated in Float64 here (the reference sums in T)."""

    def __init__(self, P, fx, fy, fz):
        self.P = float(P)
        self.f = (fx, fy, fz)

    def amplitude(self, prob):
        T = prob.grid.T.type
        u = prob.vars
        f = [np.asarray(a, dtype=T) for a in self.f]
        S = sum(float(np.sum(np.abs((u[i].astype(np.float64) ** 2) * f[i].astype(np.float64)))) for i in range(3))
        g = prob.grid
        dV = float(g.dx) * float(g.dy) * float(g.dz)
        return self.P / (S * dV)

    def apply(self, N, sol, t, prob):
        from .rhs import r2c
        T = prob.grid.T.type
        A = self.amplitude(prob)
        for i in range(3):
            Fh = r2c((np.asarray(self.f[i], dtype=T) * prob.vars[i]).astype(T)).astype(prob.grid.C)
            N[i] = (N[i].astype(np.complex128) + A * Fh.astype(np.complex128)).astype(N.dtype)


MASK53 = (1 << 53) - 1


def counter_uniforms(seed, counter, mode_index, T):
    """Counter-based uniforms in [0, 1) for A99 draws: splitmix64 of
    ``seed * 0x9E3779B97F4A7C15 + (counter << 32) + mode_index + 1`` (64-bit wrap), top 53 bits
    (Float64) or top 24 bits (Float32).  Vectorised over ``mode_index`` (uint64 array)."""
    with np.errstate(over="ignore"):
        base = np.uint64((int(seed) * 0x9E3779B97F4A7C15 + (int(counter) << 32) + 1) & MASK64)
        x = base + mode_index.astype(np.uint64)
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    if np.dtype(T) == np.float64:
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return ((z >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)).astype(np.float32)


class A99:
    """Alvelius (1999) stochastic forcing (src/pgen/A99ForceDriving.jl).

    ``SetUpFk`` (:93-127), in T on the grid: ``k = sqrt(Krsq)``, ``k^-1 = sqrt(invKrsq)``,
    ``k_perp = sqrt(kx^2 + ky^2)``, ``Int = sum(exp(-(k-kf)^2/s2) / (k+1)^2)`` over the half
    spectrum, ``Amp = sqrt(P*3*(Lx/dx)*(Ly/dy)*(Lz/dz)/Int*(1/dx/dy/dz))``,
    ``Fk = Amp*sqrt(exp(-(k-kf)^2/s2)/2/pi)*k^-1`` with ``Fk[kx = 0] = 0``;
    ``e1 = (ky, -kx, 0)/k_perp``, ``e2 = (kx kz, ky kz, -k_perp^2)/(k_perp k)`` (NaN -> 0).
    ``A99ForceDriving!`` (:33-60): per mode and call, ``e^{i theta1}`` with theta1 = 2 pi U,
    ``Phi = pi (U_re + i U_im)`` (``rand!`` of the Complex{T} buffer ``nonlinh1`` fills both
    parts), ``gi = -tanh(b (Phi - pi/2)) / tanh(b pi / 2)``, ``N[u] += A Fk e^{i theta1} gi e1``,
    then a new ``e^{i theta2}``, ``gj = sqrt(1 - gi^2)``, ``N[u] += A Fk e^{i theta2} gj e2``.

    The reference draws unseeded ``rand``; here every draw is ``counter_uniforms`` keyed by the
    call's counter (a step's stages: ``8 step + stage``), the mode's half-spectrum index
    ``(m ny + l) nkr + kx`` and the draw (theta1, Phi_re, Phi_im, theta2) in the counter's low
    bits, so the device and this restatement draw the same numbers."""

    def __init__(self, seed, kf=2.0, P=1.0, sigma2=1.0, b=1.0, A=1.0):
        self.seed, self.kf, self.P, self.sigma2, self.b, self.A = int(seed), kf, P, sigma2, b, A

    def setup(self, grid):
        T = grid.T.type
        kx = np.broadcast_to(grid.kr, grid.shape_c).astype(T)
        ky = np.broadcast_to(grid.l, grid.shape_c).astype(T)
        kz = np.broadcast_to(grid.m, grid.shape_c).astype(T)
        k = np.sqrt(grid.Krsq).astype(T)
        kinv = np.sqrt(grid.invKrsq).astype(T)
        kperp = np.sqrt(kx * kx + ky * ky).astype(T)
        kf, s2 = T(self.kf), T(self.sigma2)
        g = np.exp(-((k - kf) ** 2) / s2).astype(T)
        integral = float(np.sum((g * (T(1) / ((k + T(1)) ** 2))).astype(T), dtype=np.float64))
        dx, dy, dz = float(grid.dx), float(grid.dy), float(grid.dz)
        amp = math.sqrt(self.P * 3 * (grid.Lx / dx) * (grid.Ly / dy) * (grid.Lz / dz) / integral * (1 / dx / dy / dz))
        Fk = (T(amp) * np.sqrt(g / T(2) / T(math.pi)) * kinv).astype(T)
        Fk[:, :, 0] = 0
        with np.errstate(divide="ignore", invalid="ignore"):
            e1 = [ky / kperp, -kx / kperp, np.zeros_like(k)]
            e2 = [kx * kz / kperp * kinv, ky * kz / kperp * kinv, -kperp * kinv]
        for e in e1[:2] + e2[:2]:
            e[np.isnan(e)] = 0
        self.amp, self.Fk = amp, Fk
        self.e1 = [a.astype(T) for a in e1]
        self.e2 = [a.astype(T) for a in e2]
        return self

    def force(self, grid, counter):
        """The two forcing terms of one call, each complex of shape (3, nz, ny, nkr).

        Mixed precision as in the reference for T = Float32: ``2pi`` and ``pi/2`` are Float64
        literals, so the phases and the tanh argument are evaluated in Float64 and rounded to
        Complex{T} when stored; ``Phi *= pi`` and ``tanh(b*pi/2)`` stay in T."""
        T = grid.T.type
        C = grid.C.type
        if not hasattr(self, "Fk"):
            self.setup(grid)
        idx = np.arange(grid.nz * grid.ny * grid.nkr, dtype=np.uint64).reshape(grid.shape_c)
        u = [counter_uniforms(self.seed, (counter << 2) + d, idx, T) for d in range(4)]
        two_pi = 2 * math.pi
        e_i1 = np.exp(1j * (u[0].astype(np.float64) * two_pi)).astype(C)
        Phi = ((u[1].astype(C) + C(1j) * u[2].astype(C)) * T(math.pi)).astype(C)
        b = T(self.b)
        den = T(np.tanh(T(b * T(math.pi)) / T(2)))
        gi = (-np.tanh(np.float64(b) * (Phi.astype(np.complex128) - math.pi / 2)) / np.float64(den)).astype(C)
        e_i2 = np.exp(1j * (u[3].astype(np.float64) * two_pi)).astype(C)
        gj = np.sqrt(C(1) - gi * gi).astype(C)
        A = T(self.A)
        live = self.Fk != 0
        w1 = (A * self.Fk * e_i1 * gi).astype(C)
        w2 = (A * self.Fk * e_i2 * gj).astype(C)
        F1 = np.zeros((3,) + grid.shape_c, dtype=C)
        F2 = np.zeros((3,) + grid.shape_c, dtype=C)
        for c in range(3):
            if c < 2:
                F1[c] = np.where(live, w1 * self.e1[c], 0)
            F2[c] = np.where(live, w2 * self.e2[c], 0)
        return F1, F2

    def apply(self, N, sol, t, prob):
        F1, F2 = self.force(prob.grid, prob.force_counter)
        N[0:2] += F1[0:2].astype(N.dtype)      # :48-49
        N[0:3] += F2.astype(N.dtype)           # :54-56
