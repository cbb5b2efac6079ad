"""Right-hand sides of the incompressible HD/MHD equations (oracle; test infrastructure only).

State layout: ``sol[f]`` is the half spectrum of field ``f`` (``[z][y][kx]``),
``f = 0..5`` = ux, uy, uz, bx, by, bz (``src/Structure/datastructure.jl:90``:
indices 1..6).  ``vars[f]`` holds the real fields (``[z][y][x]``).
"""

import numpy as np
import scipy.fft as sfft

from .grid import dealias_, hermitian_project

WORKERS = -1  # scipy.fft workers: all host cores, as FFTW threads = Sys.CPU_THREADS


def r2c(x):
    """Unnormalised forward real FFT, ``mul!(xh, grid.rfftplan, x)``."""
    return sfft.rfftn(x, axes=(-3, -2, -1), workers=WORKERS)


def c2r(X, shape):
    """Normalised inverse real FFT, ``ldiv!(x, grid.rfftplan, X)`` (scaled by 1/N)."""
    return sfft.irfftn(X, s=shape, axes=(-3, -2, -1), workers=WORKERS)


def _kvec(grid):
    return (grid.kr, grid.l, grid.m)


def _delta(a, b):
    return 1 if a == b else 0


def _promoted_diffusion(Nf, Krsq, coef, uh, power=1):
    """``N += -Krsq^p * coef * uh`` with ``coef`` a Float64 (``MHDParams.jl:45-47``).

    Julia promotes the product to Complex{Float64} and rounds the sum back to the
    storage type of ``N`` (``MHDSolver.jl:94,98,168,172``).
    """
    K = Krsq if power == 1 else (Krsq ** power).astype(Krsq.dtype)
    term = (-(K.astype(np.float64)) * float(coef)) * uh.astype(np.complex128)
    Nf[...] = (Nf.astype(np.complex128) + term).astype(Nf.dtype)


# ---------------------------------------------------------------------------
# literal mode: the reference's operation structure, one product at a time
# ---------------------------------------------------------------------------

def mhd_u_update_literal(N, sol, t, prob, a):
    """``UᵢUpdate!`` (``src/Solver/MHDSolver.jl:27-103``), direction ``a`` in 0..2."""
    g, v = prob.grid, prob.vars
    k = _kvec(g)
    ka, invK = k[a], g.invKrsq
    Na = N[a]
    Na *= 0                                                     # :62
    u = (v[0], v[1], v[2])
    b = (v[3], v[4], v[5])
    for i in range(3):                                          # :63-64
        for j in range(3):
            if j >= i:
                nonlin1 = b[i] * b[j] - u[i] * u[j]             # :73
                nonlinh1 = r2c(nonlin1)                         # :74
                coef = (1j * k[i]) * (_delta(a, j) - (ka * k[j]) * invK)
                Na += coef * nonlinh1                           # :77
                if i != j:
                    coef = (1j * k[j]) * (_delta(a, i) - (ka * k[i]) * invK)
                    Na += coef * nonlinh1                       # :79
    vp_update(Na, prob, a, 0)                                   # :85-88
    uh = r2c(u[a])                                              # :91-93
    _promoted_diffusion(Na, g.Krsq, prob.nu, uh)                # :94
    if prob.nnu > 1:                                            # :97-99
        _promoted_diffusion(Na, g.Krsq, prob.nu, uh, power=prob.nnu)


def mhd_b_update_literal(N, sol, t, prob, a):
    """``BᵢUpdate!`` (``src/Solver/MHDSolver.jl:106-177``), direction ``a`` in 0..2."""
    g, v = prob.grid, prob.vars
    k = _kvec(g)
    Nb = N[3 + a]
    Nb *= 0                                                     # :143
    u = (v[0], v[1], v[2])
    b = (v[3], v[4], v[5])
    for j in range(3):                                          # :147
        if a != j:
            nonlin1 = u[a] * b[j] - b[a] * u[j]                 # :150
            nonlinh1 = r2c(nonlin1)                             # :152
            Nb += (1j * k[j]) * nonlinh1                        # :155
    vp_update(Nb, prob, a, 3)                                   # :160-163
    bh = r2c(b[a])                                              # :166-167
    _promoted_diffusion(Nb, g.Krsq, prob.eta, bh)               # :168
    if prob.neta > 1:                                           # :171-173
        _promoted_diffusion(Nb, g.Krsq, prob.eta, bh, power=prob.neta)


def mhd_advection_literal(N, sol, t, prob):
    """``MHDcalcN_advection!`` (``src/Solver/MHDSolver.jl:330-351``)."""
    shape = prob.grid.shape_r
    for f in range(6):                                          # :333-338 (deepcopy: c2r input untouched)
        prob.vars[f] = c2r(sol[f].copy(), shape).astype(prob.grid.T)
    for a in range(3):                                          # :341-343
        mhd_u_update_literal(N, sol, t, prob, a)
    for a in range(3):                                          # :346-348
        mhd_b_update_literal(N, sol, t, prob, a)


def hd_u_update_literal(N, sol, t, prob, a):
    """HD ``UᵢUpdate!`` (``src/Solver/HDSolver.jl:25-93``)."""
    g, v = prob.grid, prob.vars
    k = _kvec(g)
    ka, invK = k[a], g.invKrsq
    Na = N[a]
    Na *= 0                                                     # :55
    u = (v[0], v[1], v[2])
    for i in range(3):
        for j in range(3):
            if i <= j:
                uuh = r2c(u[i] * u[j])                          # :62-65
                coef = (-1j * k[i]) * (_delta(a, j) - (ka * k[j]) * invK)
                Na += coef * uuh                                # :68
                if i != j:
                    coef = (-1j * k[j]) * (_delta(a, i) - (ka * k[i]) * invK)
                    Na += coef * uuh                            # :70
    vp_update(Na, prob, a, 0)                                   # HDSolver.jl:76-79
    uh = r2c(u[a])                                              # :82-84
    _promoted_diffusion(Na, g.Krsq, prob.nu, uh)                # :85
    if prob.nnu > 1:                                            # :88-90
        _promoted_diffusion(Na, g.Krsq, prob.nu, uh, power=prob.nnu)


def hd_advection_literal(N, sol, t, prob):
    """``HDcalcN_advection!`` (``src/Solver/HDSolver.jl:95-108``)."""
    shape = prob.grid.shape_r
    for f in range(3):
        prob.vars[f] = c2r(sol[f].copy(), shape).astype(prob.grid.T)
    for a in range(3):
        hd_u_update_literal(N, sol, t, prob, a)


# ---------------------------------------------------------------------------
# fused mode: the same mathematics with 15 (MHD) / 9 (HD) FFTs per evaluation
# ---------------------------------------------------------------------------

def mhd_advection_fused(N, sol, t, prob):
    """Same result as ``mhd_advection_literal`` with 6 C2R + 9 R2C transforms.

    ``UᵢUpdate!`` summed over all ``i,j`` equals the Leray projection of
    ``D_j = Σ_i i k_i T̂_ij`` with ``T = b⊗b − u⊗u`` symmetric; ``BᵢUpdate!`` uses the
    antisymmetric ``M_aj = u_a b_j − b_a u_j`` (3 independent components); the
    diffusion acts on ``R2C(C2R(ŝ)) = P_H(ŝ)``.
    """
    g = prob.grid
    T = g.T
    shape = g.shape_r
    for f in range(6):
        prob.vars[f] = c2r(sol[f], shape).astype(T)
    u = prob.vars[0:3]
    b = prob.vars[3:6]
    k = _kvec(g)
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    Th = {}
    for (i, j) in pairs:
        Th[(i, j)] = r2c(b[i] * b[j] - u[i] * u[j])
        Th[(j, i)] = Th[(i, j)]
    D = [sum((1j * k[i]) * Th[(i, j)] for i in range(3)) for j in range(3)]
    kD = sum(k[j] * D[j] for j in range(3))
    sp = hermitian_project(sol)
    for a in range(3):
        Na = (D[a] - k[a] * (kD * g.invKrsq)).astype(g.C)
        vp_update(Na, prob, a, 0)
        _promoted_diffusion(Na, g.Krsq, prob.nu, sp[a])
        if prob.nnu > 1:
            _promoted_diffusion(Na, g.Krsq, prob.nu, sp[a], power=prob.nnu)
        N[a] = Na
    Mh = {}
    for (a, j) in [(0, 1), (0, 2), (1, 2)]:
        Mh[(a, j)] = r2c(u[a] * b[j] - b[a] * u[j])
        Mh[(j, a)] = -Mh[(a, j)]
    for a in range(3):
        Nb = sum((1j * k[j]) * Mh[(a, j)] for j in range(3) if j != a).astype(g.C)
        vp_update(Nb, prob, a, 3)
        _promoted_diffusion(Nb, g.Krsq, prob.eta, sp[3 + a])
        if prob.neta > 1:
            _promoted_diffusion(Nb, g.Krsq, prob.eta, sp[3 + a], power=prob.neta)
        N[3 + a] = Nb


def hd_advection_fused(N, sol, t, prob):
    g = prob.grid
    T = g.T
    shape = g.shape_r
    for f in range(3):
        prob.vars[f] = c2r(sol[f], shape).astype(T)
    u = prob.vars[0:3]
    k = _kvec(g)
    Uh = {}
    for i in range(3):
        for j in range(i, 3):
            Uh[(i, j)] = r2c(u[i] * u[j])
            Uh[(j, i)] = Uh[(i, j)]
    D = [sum((-1j * k[i]) * Uh[(i, j)] for i in range(3)) for j in range(3)]
    kD = sum(k[j] * D[j] for j in range(3))
    sp = hermitian_project(sol[0:3])
    for a in range(3):
        Na = (D[a] - k[a] * (kD * g.invKrsq)).astype(g.C)
        vp_update(Na, prob, a, 0)
        _promoted_diffusion(Na, g.Krsq, prob.nu, sp[a])
        if prob.nnu > 1:
            _promoted_diffusion(Na, g.Krsq, prob.nu, sp[a], power=prob.nnu)
        N[a] = Na


# ---------------------------------------------------------------------------
# calcN! wrappers (src/pgen.jl:153-181)
# ---------------------------------------------------------------------------

def calcN(N, sol, t, prob):
    """``MHDcalcN!`` / ``HDcalcN!``: dealias the stage input in place, advection, forcing."""
    dealias_(sol, prob.grid)                                    # pgen.jl:155 / :174
    literal = prob.mode == "literal"
    if prob.physics == "MHD":
        (mhd_advection_literal if literal else mhd_advection_fused)(N, sol, t, prob)
        add_forcing(N, sol, t, prob)                            # pgen.jl:159
    else:
        # pgen.jl:176-178: forcing is added before the advection, which zeroes N
        # (HDSolver.jl:55), so HD forcing never reaches the right-hand side.
        add_forcing(N, sol, t, prob)
        (hd_advection_literal if literal else hd_advection_fused)(N, sol, t, prob)


def add_forcing(N, sol, t, prob):
    """``addforcing!`` (``pgen.jl:231-234``): the problem's calcF!, i.e. one of the reference's
    modules (``forcing.N97``, ``forcing.NegativeDamping``, ``forcing.A99``, each with ``apply``)
    or the build's seeded large-scale forcing (``forcing.Forcing``)."""
    if prob.forcing is None:
        return
    if hasattr(prob.forcing, "apply"):
        prob.forcing.apply(N, sol, t, prob)
        return
    Fh = prob.forcing.spectrum(prob.grid)  # (3, nz, ny, nkr), already scaled by N
    N[0:3] += Fh.astype(prob.grid.C)


from .vp import vp_update  # noqa: E402  (oracle.vp imports the transforms above)
