"""Grid conventions of FourierFlows' ``ThreeDGrid`` (oracle; test infrastructure only).

Array layout: C order ``[z][y][x]`` for real fields and ``[z][y][kx]`` for the
half spectrum, i.e. byte-identical to Julia's column-major ``T[x,y,z]`` /
``Complex{T}[kr,l,m]`` (``src/Structure/datastructure.jl:8-9``).
"""

import math

import numpy as np

ALIASED_FRACTION = 1.0 / 3.0


def aliased_ranges(n, aliased_fraction=ALIASED_FRACTION):
    """0-based aliased index ranges ``(kalias, kralias)`` for an axis of length ``n``.

    Follows FourierFlows ``getaliasedwavenumbers`` as called at
    ``src/utils/utils.jl:82-84``: ``L=(1-a)/2``, ``R=(1+a)/2``,
    ``iL=floor(L*n)+1``, ``iR=ceil(R*n)`` (1-based), ``kalias=iL:iR``,
    ``kralias=iL:nkr``.  Float64 arithmetic, as in Julia.
    Returns python ``range`` objects of 0-based indices.
    """
    a = float(aliased_fraction)
    L = (1.0 - a) / 2.0
    R = (1.0 + a) / 2.0
    iL = math.floor(L * n) + 1
    iR = math.ceil(R * n)
    nkr = n // 2 + 1
    return range(iL - 1, iR), range(iL - 1, nkr)


def kept_range(n):
    """Kept signed wavenumber indices ``[lo, hi]`` along a full (C2C) axis."""
    kal, _ = aliased_ranges(n)
    lo = -(n - kal.stop)  # first kept negative index after the aliased block
    hi = kal.start - 1
    return lo, hi


def keep_masks(nx, ny, nz):
    """Boolean keep-masks (x over nkr, y over ny, z over nz)."""
    _, kral = aliased_ranges(nx)
    lal, _ = aliased_ranges(ny)
    mal, _ = aliased_ranges(nz)
    mx = np.ones(nx // 2 + 1, bool)
    mx[kral.start:kral.stop] = False
    my = np.ones(ny, bool)
    my[lal.start:lal.stop] = False
    mz = np.ones(nz, bool)
    mz[mal.start:mal.stop] = False
    return mx, my, mz


def fftfreq_T(n, L, dtype):
    """``fftfreq(n, 2π/L*n)`` evaluated in Float64 then converted to T (``utils.jl:63-66``)."""
    mult = ((2.0 * math.pi / L) * n) / n
    idx = np.fft.fftfreq(n, 1.0 / n)  # integers 0..n/2-1, -n/2..-1
    return (idx * mult).astype(dtype)


def rfftfreq_T(n, L, dtype):
    mult = ((2.0 * math.pi / L) * n) / n
    idx = np.arange(n // 2 + 1, dtype=np.float64)
    return (idx * mult).astype(dtype)


class Grid:
    """Wavenumbers and k^2 arrays in the problem's real type ``T``.

    ``Krsq = kr^2 + l^2 + m^2`` and ``invKrsq = 1/Krsq`` with ``invKrsq[0,0,0] = 0``
    (``src/utils/utils.jl:73-75``), all evaluated in T.
    """

    def __init__(self, nx, ny=None, nz=None, Lx=2 * math.pi, Ly=None, Lz=None, T=np.float32):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        Ly = Lx if Ly is None else Ly
        Lz = Lx if Lz is None else Lz
        for n in (nx, ny, nz):
            if n % 2:
                # Int(nx/2+1) raises InexactError for odd n (utils.jl:20,55)
                raise ValueError(f"grid sizes must be even, got {n}")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.T = np.dtype(T)
        self.C = np.dtype(np.complex64 if self.T == np.float32 else np.complex128)
        self.nkr = nx // 2 + 1
        T = self.T
        self.kr = rfftfreq_T(nx, Lx, T)[None, None, :]
        self.l = fftfreq_T(ny, Ly, T)[None, :, None]
        self.m = fftfreq_T(nz, Lz, T)[:, None, None]
        kr2 = (self.kr * self.kr).astype(T)
        l2 = (self.l * self.l).astype(T)
        m2 = (self.m * self.m).astype(T)
        self.Krsq = ((kr2 + l2).astype(T) + m2).astype(T)
        with np.errstate(divide="ignore"):
            self.invKrsq = (T.type(1) / self.Krsq).astype(T)
        self.invKrsq[0, 0, 0] = 0
        self.mx, self.my, self.mz = keep_masks(nx, ny, nz)
        self.keep = self.mz[:, None, None] & self.my[None, :, None] & self.mx[None, None, :]
        # grid spacing as stored in the grid (type T): dx = Lx/nx
        self.dx = T.type(T.type(Lx) / T.type(nx))
        self.dy = T.type(T.type(Ly) / T.type(ny))
        self.dz = T.type(T.type(Lz) / T.type(nz))

    @property
    def shape_r(self):
        return (self.nz, self.ny, self.nx)

    @property
    def shape_c(self):
        return (self.nz, self.ny, self.nkr)

    def coords(self, dtype=np.float64):
        """Cell coordinates ``x_i = x0 + i*dx`` with ``x0 = -L/2`` (mirrored ``utils.jl:43,58-60``)."""
        x = -self.Lx / 2 + np.arange(self.nx) * (self.Lx / self.nx)
        y = -self.Ly / 2 + np.arange(self.ny) * (self.Ly / self.ny)
        z = -self.Lz / 2 + np.arange(self.nz) * (self.Lz / self.nz)
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        return X.astype(dtype), Y.astype(dtype), Z.astype(dtype)


def dealias_(fh, grid):
    """FourierFlows ``dealias!`` on a half-spectrum array (in place), called at ``pgen.jl:155``.

    Zeros ``fh[kralias,:,:]``, ``fh[:,lalias,:]`` and ``fh[:,:,malias]`` in Julia
    index order, i.e. axes ``-1``, ``-2``, ``-3`` here.  Works on a stack of fields.
    """
    fh[..., ~grid.mx] = 0
    fh[..., ~grid.my, :] = 0
    fh[..., ~grid.mz, :, :] = 0
    return fh


def hermitian_project(X):
    """Project the kx=0 and kx=Nyquist planes of a half spectrum onto their Hermitian part.

    ``P_H X (0,l,m) = (X(0,l,m) + conj X(0,-l,-m)) / 2`` (indices mod n); same on
    the Nyquist plane.  ``rfftn(irfftn(X)) == P_H(X)`` for pocketfft/FFTW-style c2r.
    """
    Y = X.copy()
    for p in (0, X.shape[-1] - 1):
        P = X[..., p]
        mirror = np.conj(np.roll(np.flip(P, axis=(-1, -2)), shift=(1, 1), axis=(-1, -2)))
        Y[..., p] = (P + mirror) / 2
    return Y
