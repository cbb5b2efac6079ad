"""Host-side mirror of the MHDFlows.jl problem API on the MI355X device.

Julia names ending in ``!`` drop the ``!`` here (``SetUpProblemIC!`` ->
``SetUpProblemIC``, ``TimeIntegrator!`` -> ``TimeIntegrator`` ...).  Keyword
arguments keep the reference's names, Greek letters included (``ν``, ``η``,
``nν``, ``nη``); ASCII spellings (``nu``, ``eta``, ``nnu``, ``neta``) are
accepted too.  Errors raise :class:`MHDError` (the reference's ``error(...)``).
"""

import ctypes as C
import math

import numpy as np

from . import _lib as L
from ._lib import MHDError, check, lib


def nothingfunction(*args):
    """Default ``calcF``/``usr_func`` hook (``src/pgen.jl:7``)."""
    return None


class _Clock:
    """``prob.clock``: t, dt (stored in T on the device) and step."""

    def __init__(self, prob):
        self._p = prob

    def _read(self):
        t, dt, st = C.c_double(), C.c_double(), C.c_int64()
        check(lib().mhd_get_clock(self._p._h, C.byref(t), C.byref(dt), C.byref(st)))
        return t.value, dt.value, st.value

    @property
    def t(self):
        return self._read()[0]

    @t.setter
    def t(self, v):
        _, dt, st = self._read()
        check(lib().mhd_set_clock(self._p._h, float(v), dt, st))

    @property
    def dt(self):
        return self._read()[1]

    @dt.setter
    def dt(self, v):
        t, _, st = self._read()
        check(lib().mhd_set_clock(self._p._h, t, float(v), st))

    @property
    def step(self):
        return self._read()[2]

    @step.setter
    def step(self, v):
        t, dt, _ = self._read()
        check(lib().mhd_set_clock(self._p._h, t, dt, int(v)))


class _Vars:
    """``prob.vars``: real fields copied from the device on access."""

    NAMES = ("ux", "uy", "uz", "bx", "by", "bz")

    def __init__(self, prob):
        self._p = prob

    def field(self, f):
        p = self._p
        out = np.empty((p.grid.nz, p.grid.ny, p.grid.nx), dtype=p.T)
        check(lib().mhd_get_vars(p._h, f, out.ctypes.data_as(C.c_void_p)))
        return out

    def __getattr__(self, name):
        if name in self.NAMES:
            f = self.NAMES.index(name)
            if f < self._p.nf:
                return self.field(f)
        raise AttributeError(name)


class _Grid:
    """Grid sizes and wavenumbers (``ThreeDGrid`` conventions, ``src/utils/utils.jl:42-98``)."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, T):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nkr = nx // 2 + 1
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.T = T
        self.dx = T(T(Lx) / T(nx))
        self.dy = T(T(Ly) / T(ny))
        self.dz = T(T(Lz) / T(nz))

    def _axis(self, n, L, d):
        # x_i = -L/2 + i dx rounded to T (FourierFlows' default origin, mirrored at utils.jl:43,58-60)
        return (-L / 2 + np.arange(n) * float(d)).astype(self.T)

    @property
    def x(self):
        return self._axis(self.nx, self.Lx, self.dx)

    @property
    def y(self):
        return self._axis(self.ny, self.Ly, self.dy)

    @property
    def z(self):
        return self._axis(self.nz, self.Lz, self.dz)


class _Params:
    def __init__(self, nu, eta, nnu, neta, calcF):
        self.ν = self.nu = nu
        self.η = self.eta = eta
        self.nν = self.nnu = nnu
        self.nη = self.neta = neta
        self.calcF = calcF
        self.ux_ind, self.uy_ind, self.uz_ind = 1, 2, 3
        self.bx_ind, self.by_ind, self.bz_ind = 4, 5, 6


class _Flag:
    def __init__(self, b):
        self.b, self.e, self.vp, self.c, self.s = b, False, False, False, False


class MHDFlowsProblem:
    """A problem whose state lives in MI355X HBM, owned by the HIP library."""

    def __init__(self, cfg, T, calcF):
        self.cfg = cfg
        self.T = T
        self.C = np.complex64 if T == np.float32 else np.complex128
        self.nf = 6 if cfg.physics == L.MHD_PHYS_MHD else 3
        h = C.c_void_p()
        check(lib().mhd_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.grid = _Grid(cfg.nx, cfg.ny, cfg.nz, cfg.Lx, cfg.Ly, cfg.Lz, T)
        self.params = _Params(cfg.nu, cfg.eta, cfg.n_nu, 0, calcF)
        self.flag = _Flag(cfg.physics == L.MHD_PHYS_MHD)
        self.clock = _Clock(self)
        self.vars = _Vars(self)
        self.stepper = "RK4" if cfg.stepper == L.MHD_RK4 else "LSRK54"
        self.usr_func = [nothingfunction]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mhd_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- state ---------------------------------------------------------------
    @property
    def sol_shape(self):
        g = self.grid
        return (self.nf, g.nz, g.ny, g.nkr)

    @property
    def sol(self):
        out = np.empty(self.sol_shape, dtype=self.C)
        check(lib().mhd_get_sol(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    @sol.setter
    def sol(self, value):
        a = np.ascontiguousarray(value, dtype=self.C)
        if a.shape != self.sol_shape:
            raise MHDError(-1, f"sol must have shape {self.sol_shape}")
        check(lib().mhd_set_sol(self._h, a.ctypes.data_as(C.c_void_p)))

    def device_ptrs(self):
        s, v = C.c_void_p(), C.c_void_p()
        check(lib().mhd_device_ptrs(self._h, C.byref(s), C.byref(v)))
        return s.value, v.value

    def __repr__(self):
        return (f"MHDFlows Problem on MI355X: {self.grid.nx}x{self.grid.ny}x{self.grid.nz}, "
                f"B-field {'ON (Ideal MHD)' if self.flag.b else 'OFF'}, T={np.dtype(self.T).name}, "
                f"timestepper {self.stepper}")


def _pick(kw, names, default):
    for n in names:
        if n in kw:
            return kw.pop(n)
    return default


def Problem(dev="MI355X", **kw):
    """``Problem(dev; nx, ny, nz, Lx, Ly, Lz, ν, nν, η, nη, B_field, stepper, calcF, T, dt, ...)``
    (``src/pgen.jl:64-127``).  Only plain HD/MHD on the MI355X device is supported here;
    the flags of other physics raise as unsupported.

    ``aliased_fraction`` is accepted and ignored, as the reference does not forward
    it to the grid (``src/pgen.jl:91`` vs ``:107``): the 1/3 rule always applies.
    ``nη`` is accepted and ignored, as the reference does not forward it to the
    parameters (``src/pgen.jl:114-116``).
    """
    if dev not in ("MI355X", "GPU"):
        raise MHDError(-1, f"unknown device {dev!r}: this package drives the MI355X device")
    nx = int(_pick(kw, ["nx"], 64))
    ny = int(_pick(kw, ["ny"], nx))
    nz = int(_pick(kw, ["nz"], nx))
    Lx = float(_pick(kw, ["Lx"], 2 * math.pi))
    Ly = float(_pick(kw, ["Ly"], Lx))
    Lz = float(_pick(kw, ["Lz"], Lx))
    nu = float(_pick(kw, ["ν", "nu"], 0.0))
    nnu = int(_pick(kw, ["nν", "nnu"], 0))
    eta = float(_pick(kw, ["η", "eta"], 0.0))
    _pick(kw, ["nη", "neta"], 0)
    B = bool(_pick(kw, ["B_field"], False))
    vp = bool(_pick(kw, ["VP_method"], False))
    for flag in ("EMHD", "Compressibility", "Shear", "Dye_Module"):
        if _pick(kw, [flag], False):
            raise MHDError(-2, f"{flag} is not part of the MI355X hot path")
    stepper = _pick(kw, ["stepper"], "RK4")
    calcF = _pick(kw, ["calcF"], nothingfunction)
    if calcF is not nothingfunction and not getattr(calcF, "_device_forcing", False):
        # addforcing! calls calcF! inside every right-hand side (src/pgen.jl:231-234); a host
        # callable cannot run inside the device step, so it is refused rather than dropped
        raise MHDError(-2, "a user calcF! cannot run inside the device step; use the library's "
                           "forcing modules instead")
    T = _pick(kw, ["T"], np.float32)
    dt = float(_pick(kw, ["dt"], 0.0))
    _pick(kw, ["aliased_fraction"], 1 / 3)
    device = int(_pick(kw, ["device"], 0))
    budget = int(_pick(kw, ["hbm_budget"], 0))
    for k in ("usr_vars", "usr_params", "usr_func", "cₛ"):
        kw.pop(k, None)
    if kw:
        raise MHDError(-1, f"unknown keyword(s) {sorted(kw)}")
    T = np.dtype(T).type
    if T not in (np.float32, np.float64):
        raise MHDError(-1, "T must be Float32 or Float64")
    if stepper not in ("RK4", "LSRK54", "LSRK4"):
        raise MHDError(-1, f"stepper {stepper!r} is not supported on the MI355X device")
    cfg = L.Config(nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Lz=Lz, nu=nu, eta=eta if B else 0.0,
                   n_nu=nnu, n_eta=0, precision=L.MHD_F32 if T == np.float32 else L.MHD_F64,
                   physics=L.MHD_PHYS_MHD if B else L.MHD_PHYS_HD,
                   stepper=L.MHD_RK4 if stepper == "RK4" else L.MHD_LSRK54,
                   device=device, stream=None, hbm_budget=budget, dt=dt)
    prob = MHDFlowsProblem(cfg, T, calcF)
    if vp:   # MHDParams_VP / HDParams_VP: chi, U0 (and B0) start as zero fields (datastructure.jl:79-95)
        check(lib().mhd_set_vp(prob._h, 1, None, None, None))
        prob.flag.vp = True
    return prob


def SetUpProblemIC(prob, ux=None, uy=None, uz=None, bx=None, by=None, bz=None, **kw):
    """``SetUpProblemIC!`` (``src/utils/IC.jl:41-109``): real fields in ``[z][y][x]`` order.
    ``U₀x``..``U₀z`` and ``B₀x``..``B₀z`` (ASCII ``U0x``.. ``B0z``) set the volume-penalisation
    fields of a ``VP_method`` problem (:93-105)."""
    vpk = {}
    for name, alias in (("U₀x", "U0x"), ("U₀y", "U0y"), ("U₀z", "U0z"), ("B₀x", "B0x"), ("B₀y", "B0y"), ("B₀z", "B0z")):
        v = kw.pop(name, kw.pop(alias, None))
        if v is not None and not (isinstance(v, (list, tuple)) and len(v) == 0):
            vpk[alias] = v
    if kw:
        raise MHDError(-1, f"unknown keyword(s) {sorted(kw)}")
    if vpk:
        if not prob.flag.vp:
            raise MHDError(-1, "U₀/B₀ fields need a VP_method problem")
        SetUpVP(prob, **vpk)
    g = prob.grid
    arrs = [ux, uy, uz] + ([bx, by, bz] if prob.nf == 6 else [])
    keep = []
    ptrs = (C.c_void_p * prob.nf)()
    for f, a in enumerate(arrs):
        if a is None or (isinstance(a, (list, tuple)) and len(a) == 0):
            ptrs[f] = None
            continue
        a = np.ascontiguousarray(a, dtype=prob.T)
        if a.shape != (g.nz, g.ny, g.nx):
            raise MHDError(-1, f"field {f} must have shape {(g.nz, g.ny, g.nx)}")
        keep.append(a)
        ptrs[f] = a.ctypes.data
    check(lib().mhd_set_ic_real(prob._h, ptrs))


def SetUpVP(prob, chi=None, U0x=None, U0y=None, U0z=None, B0x=None, B0y=None, B0z=None, **kw):
    """Volume-penalisation fields of a ``VP_method`` problem: ``params.χ`` (what the reference's
    user copies in, e.g. ``Cylindrical_Mask_Function``) and ``params.U₀x..z``, ``B₀x..z``
    (``src/Structure/MHDParams.jl:1-40``), real ``[z][y][x]`` arrays; a field not given keeps its
    value.  Every right-hand side then adds VP_UᵢUpdate!/VP_BᵢUpdate! (src/Solver/VPSolver.jl:21-56)."""
    chi = kw.pop("χ", chi)
    if kw:
        raise MHDError(-1, f"unknown keyword(s) {sorted(kw)}")
    if not prob.flag.vp:
        raise MHDError(-1, "SetUpVP needs a problem created with VP_method=True")
    g = prob.grid
    keep = []

    def ptr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=prob.T), (g.nz, g.ny, g.nx)))
        keep.append(a)
        return a.ctypes.data

    U = (C.c_void_p * 3)(*[ptr(a) for a in (U0x, U0y, U0z)])
    B = (C.c_void_p * 3)(*[ptr(a) for a in (B0x, B0y, B0z)])
    check(lib().mhd_set_vp(prob._h, 1, ptr(chi), U, B))


def Cylindrical_Mask_Function(grid, R2=0.82 * math.pi, R1=0.0, **kw):
    """``Cylindrical_Mask_Function(grid; R₂, R₁)`` (``src/utils/IC.jl:15-27``): 0 where
    ``R₂ >= sqrt(x^2 + y^2) >= R₁`` (the fluid), 1 elsewhere (the solid), as a ``[z][y][x]`` array
    of the grid's type (the reference's BitArray copied into ``params.χ``)."""
    R2 = kw.pop("R₂", R2)
    R1 = kw.pop("R₁", R1)
    if kw:
        raise MHDError(-1, f"unknown keyword(s) {sorted(kw)}")
    x = grid.x[None, None, :]
    y = grid.y[None, :, None]
    R = np.sqrt(x * x + y * y)          # in T, compared with the Float64 radii
    fluid = (R2 >= R.astype(np.float64)) & (R.astype(np.float64) >= R1)
    return np.broadcast_to(np.where(fluid, 0, 1), (grid.nz, grid.ny, grid.nx)).astype(grid.T)


def SetUpProblemIC_modes(prob, idx, vals):
    """Spectral initial condition from sparse half-spectrum entries (see ``ic.ModeSet.half_entries``)."""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=prob.C)
    if idx.ndim != 2 or idx.shape[1] != 3 or vals.shape != (idx.shape[0], prob.nf):
        raise MHDError(-1, "idx must be (M,3) and vals (M,nf)")
    check(lib().mhd_set_sol_modes(prob._h, idx.shape[0], idx.ctypes.data_as(C.c_void_p),
                                  vals.ctypes.data_as(C.c_void_p)))


def get_sol_modes(prob, idx):
    """Entries ``(kx, l, m)`` (rows of ``idx``) of ``prob.sol``, shape ``(M, nf)``."""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    if idx.ndim != 2 or idx.shape[1] != 3:
        raise MHDError(-1, "idx must be (M,3)")
    out = np.empty((idx.shape[0], prob.nf), dtype=prob.C)
    check(lib().mhd_get_sol_modes(prob._h, idx.shape[0], idx.ctypes.data_as(C.c_void_p),
                                  out.ctypes.data_as(C.c_void_p)))
    return out


def stepforward(prob, nsteps=1):
    """``stepforward!`` (``src/timestepper/timestepper.jl:4-6``) with the problem's stepper."""
    check(lib().mhd_step(prob._h, int(nsteps)))


def getCFL(prob, t_diff=None, Coef=0.3):
    """``getCFL!`` (``src/integrator.jl:158-198``); sets and returns ``clock.dt``.  The library
    bounds dt by the diffusive limit of ``Coef``; a ``t_diff`` passed here bounds it as well."""
    dt = C.c_double()
    check(lib().mhd_cfl_dt(prob._h, float(Coef), C.byref(dt)))
    if t_diff is not None and t_diff < dt.value:
        prob.clock.dt = t_diff
        return float(t_diff)
    return dt.value


def diffusion_dt(prob, CFL_Coef):
    """The diffusive time-step bound ``t_diff`` of ``TimeIntegrator!`` (``src/integrator.jl:60-72``)."""
    pr, g = prob.params, prob.grid
    if prob.flag.b:
        vi, nv = max(pr.nu, pr.eta), max(pr.nnu, pr.neta)
    else:
        vi, nv = pr.nu, pr.nnu
    dl = min(g.Lx / g.nx, g.Ly / g.ny, g.Lz / g.nz)
    return CFL_Coef * dl ** nv / vi if nv > 1 else CFL_Coef * dl ** 2 / vi


def TimeIntegrator(prob, t0, N0, usr_dt=0.0, CFL_Coef=0.25, CFL_function=nothingfunction, diags=(),
                   dynamic_dashboard=False, loop_number=100, save=False, save_loc="", filename="",
                   file_number=0, dump_dt=0):
    """``TimeIntegrator!`` (``src/integrator.jl:31-156``): at most ``N0+1`` steps while ``t <= t0``.

    The hot loop runs on the device (dt from ``getCFL!`` computed there).  Diagnostics
    callables in ``diags``, ``prob.usr_func`` hooks, a user ``CFL_function(prob, t_diff,
    Coef=...)`` (which sets ``prob.clock.dt``) and saving force a per-step loop on the host;
    without them the loop has no host synchronisation.
    """
    file_path_and_name = ""
    if save:
        if len(save_loc) == 0 or len(filename) == 0 or dump_dt == 0:
            raise MHDError(-1, "Save Function Turned ON but save_loc/filename/dump_dt is not declared!\n")
        file_path_and_name = save_loc + filename                   # integrator.jl:44-51
        savefile(prob, file_number, file_path_and_name=file_path_and_name)
        file_number += 1
    user_cfl = CFL_function is not nothingfunction
    if user_cfl and usr_dt > 0.0:                                 # CFL_Init, integrator.jl:200-206
        raise MHDError(-1, "User define both CFL_function and usr_dt")
    hooks = [f for f in prob.usr_func if f is not nothingfunction]
    vp = prob.flag.vp
    if not diags and not hooks and not save and not user_cfl and not vp:
        steps = C.c_int64()
        check(lib().mhd_integrate(prob._h, float(t0), int(N0), float(usr_dt), float(CFL_Coef), C.byref(steps)))
        return steps.value
    # per-step loop with host hooks (integrator.jl:104-146)
    taken = 0
    t_next_save = prob.clock.t + dump_dt                          # :75
    prob.clock.step = 0
    if usr_dt != 0.0:
        prob.clock.dt = usr_dt
    t_diff = diffusion_dt(prob, CFL_Coef)
    if vp:                                                        # :84-87
        DivVCorrection(prob)
        if prob.flag.b:
            DivBCorrection(prob)
    while N0 >= prob.clock.step and t0 >= prob.clock.t:
        if user_cfl:
            CFL_function(prob, t_diff, Coef=CFL_Coef)             # sets prob.clock.dt
        elif usr_dt == 0.0:
            getCFL(prob, Coef=CFL_Coef)
        stepforward(prob, 1)
        taken += 1
        for d in diags:                                           # increment!(diags)
            d(prob)
        if vp:                                                    # :118-121
            DivVCorrection(prob)
            if prob.flag.b:
                DivBCorrection(prob)
        for f in hooks:                                           # usr_func
            f(prob)
        if save and prob.clock.t >= t_next_save:                  # :136-141
            ProbDiagnostic(prob)
            savefile(prob, file_number, file_path_and_name=file_path_and_name)
            t_next_save += dump_dt
            file_number += 1
    return taken


def _dev_ptr(x):
    if isinstance(x, int):
        return C.c_void_p(x)
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    raise MHDError(-1, "device arrays are passed as integer device pointers")


def calcN(prob, N, sol, t=0.0):
    """``MHDcalcN!``/``HDcalcN!`` (``src/pgen.jl:153-181``) on device pointers ``N``, ``sol``."""
    check(lib().mhd_calcN(prob._h, _dev_ptr(N), _dev_ptr(sol), float(t)))


def MHDcalcN_advection(prob, N, sol, t=0.0):
    """``MHDcalcN_advection!`` (``src/Solver/MHDSolver.jl:330-351``)."""
    check(lib().mhd_advection(prob._h, _dev_ptr(N), _dev_ptr(sol), float(t)))


HDcalcN_advection = MHDcalcN_advection

_DIRS = {"x": 0, "y": 1, "z": 2}


def UiUpdate(prob, N, direction="x"):
    """``UᵢUpdate!(N, sol, t, clock, vars, params, grid; direction)`` (``MHDSolver.jl:27-103``)."""
    if direction not in _DIRS:
        raise MHDError(-1, "Warning : Unknown direction is declerad")
    check(lib().mhd_u_update(prob._h, _dev_ptr(N), _DIRS[direction]))


def BiUpdate(prob, N, direction="x"):
    """``BᵢUpdate!`` (``MHDSolver.jl:106-177``)."""
    if direction not in _DIRS:
        raise MHDError(-1, "Warning : Unknown direction is declerad")
    check(lib().mhd_b_update(prob._h, _dev_ptr(N), _DIRS[direction]))


def Diagnostics(prob):
    """FP64 energies, helicities and divergences of the current ``vars``."""
    d = L.Diag()
    check(lib().mhd_diagnostics(prob._h, C.byref(d)))
    return {k: getattr(d, k) for k in ("E_kin", "E_mag", "H_kin", "H_mag", "H_cross", "max_div_b", "max_div_u")} | {
        "has_nan": bool(d.has_nan)}


def ProbDiagnostic(prob):
    """``ProbDiagnostic`` (``src/utils/UserInterface.jl:65-86``): KE (and ME), 3 significant digits."""
    d = Diagnostics(prob)
    if d["has_nan"] or math.isnan(d["E_kin"]):
        raise MHDError(-5, "detected NaN! Quit the simulation right now.")

    def r3(x):
        return float(f"{x:.3g}")

    return (r3(d["E_kin"]), r3(d["E_mag"])) if prob.flag.b else r3(d["E_kin"])


# ---- calcF! modules of the reference (src/pgen/*.jl), evaluated inside the device step ----
def _device_module(name, doc):
    def f(*args):
        raise MHDError(-2, f"{name} runs inside the device step; pass it as calcF and call its setup")
    f.__name__ = name
    f.__doc__ = doc
    f._device_forcing = True
    return f


N97ForceDriving = _device_module("N97ForceDriving", "``N97ForceDriving!`` (src/pgen/TaylorGreenDynamo.jl:12-16).")
NDForceDriving = _device_module("NDForceDriving", "``NDForceDriving!`` (src/pgen/NegativeDamping.jl:23-44).")
A99ForceDriving = _device_module("A99ForceDriving", "``A99ForceDriving!`` (src/pgen/A99ForceDriving.jl:33-60).")


def _real_ptrs(prob, fields):
    g = prob.grid
    keep, ptrs = [], (C.c_void_p * 3)()
    for c, a in enumerate(fields):
        if a is None:
            ptrs[c] = None
            continue
        a = np.ascontiguousarray(np.broadcast_to(a, (g.nz, g.ny, g.nx)), dtype=prob.T)
        keep.append(a)
        ptrs[c] = a.ctypes.data
    return keep, ptrs


def SetUpN97(prob, F0=1, kf=2):
    """``SetUpN97!(prob; F0, kf)`` (src/pgen/TaylorGreenDynamo.jl:18-33): the Taylor-Green
    forcing ``fx = F0 sin(kf x) cos(kf y) cos(kf z)``, ``fy = -F0 cos(kf x) sin(kf y) cos(kf z)``
    evaluated in T on the grid, transformed on the device and added to N_u every call."""
    g, T = prob.grid, prob.T
    x = g.x[None, None, :]
    y = g.y[None, :, None]
    z = g.z[:, None, None]
    kf, F0 = T(kf), T(F0)
    fx = F0 * np.sin(kf * x) * np.cos(kf * y) * np.cos(kf * z)
    fy = F0 * -np.cos(kf * x) * np.sin(kf * y) * np.cos(kf * z)
    keep, ptrs = _real_ptrs(prob, [fx, fy, None])
    check(lib().mhd_set_forcing_field(prob._h, ptrs))


def SetUpForcingField(prob, fx=None, fy=None, fz=None):
    """A static forcing field added to the velocity right-hand side (any real fields in
    ``[z][y][x]`` order; the device-step counterpart of a calcF! that adds a fixed field)."""
    keep, ptrs = _real_ptrs(prob, [fx, fy, fz])
    check(lib().mhd_set_forcing_field(prob._h, ptrs))


def SetUpND(prob, P, fx, fy, fz):
    """``SetUpND!(prob, P, fx, fy, fz)`` (src/pgen/NegativeDamping.jl:16-22)."""
    keep, ptrs = _real_ptrs(prob, [fx, fy, fz])
    check(lib().mhd_set_forcing_nd(prob._h, float(P), ptrs))


def SetUpFk(prob, kf=2, P=1, sigma2=1, seed=0, b=1.0, A=1.0, **kw):
    """``SetUpFk(prob; kf, P, σ²)`` (src/pgen/A99ForceDriving.jl:93-127) with ``usr_vars.b`` and
    ``usr_vars.A``; ``seed`` keys the counter-based draws that replace the reference's rand()."""
    sigma2 = kw.pop("σ²", sigma2)
    if kw:
        raise MHDError(-1, f"unknown keyword(s) {sorted(kw)}")
    check(lib().mhd_set_forcing_a99(prob._h, int(seed), float(kf), float(P), float(sigma2), float(b), float(A)))


# ---- analysis of the current vars on the device (src/utils/MHDAnalysis.jl) ----
_FIELDS = ("ux", "uy", "uz", "bx", "by", "bz")


def _field_index(prob, field):
    f = _FIELDS.index(field) if isinstance(field, str) else int(field)
    if not 0 <= f < prob.nf:
        raise MHDError(-1, f"field {field!r} is not a field of this problem")
    return f


def spectra(prob):
    """Shell spectra of every vars field: ``{name: Pk}`` plus ``"count"`` (modes per shell),
    each of length krmax (see ``spectralline``)."""
    nb = C.c_int32()
    check(lib().mhd_spectra(prob._h, None, 0, C.byref(nb)))
    out = np.zeros((prob.nf + 1, nb.value), np.float64)
    check(lib().mhd_spectra(prob._h, out.ctypes.data_as(C.c_void_p), nb.value, C.byref(nb)))
    d = {_FIELDS[f]: out[f] for f in range(prob.nf)}
    d["count"] = out[prob.nf]
    return d


def spectralline(prob, field="ux"):
    """``spectralline(A)`` (``src/utils/MHDAnalysis.jl:236-255``) of a vars field: ``Pk[r]`` is the
    sum of ``|A_k|^2`` over the half-spectrum modes with ``round(|k|) = r`` (0-based r), and
    ``kr[r] = r + 1`` (the reference's 1-based shell label) for every shell that holds a mode,
    0 otherwise.  Pk is accumulated in Float64."""
    d = spectra(prob)
    Pk = d[_FIELDS[_field_index(prob, field)]]
    kr = np.where(d["count"] > 0, np.arange(1, Pk.size + 1), 0).astype(prob.T)
    return Pk, kr


def ScaleDecomposition(prob, fields=("ux", "uy", "uz"), kf=(1, 5)):
    """``ScaleDecomposition(B1, B2, B3; kf)`` (``src/utils/MHDAnalysis.jl:17-82``) of vars fields:
    the modes with ``kf[1] <= |k| <= kf[2]`` back in real space, one ``[z][y][x]`` array each."""
    if isinstance(fields, (str, int)):
        fields = (fields,)
    idx = [_field_index(prob, f) for f in fields]
    if idx != list(range(idx[0], idx[0] + len(idx))):
        raise MHDError(-1, "fields must be consecutive components of u or of b")
    g = prob.grid
    out = np.empty((len(idx), g.nz, g.ny, g.nx), prob.T)
    check(lib().mhd_scale_decomposition(prob._h, idx[0], len(idx), float(min(kf)), float(max(kf)),
                                        out.ctypes.data_as(C.c_void_p)))
    return tuple(out)


# ---- vector calculus on the device (src/utils/VectorCalculus.jl, src/utils/MHDAnalysis.jl) ----
_OP = {"x": 3, "y": 4, "z": 5, "laplace": 6, "curl": 7, "div": 8, "vecpot": 9}


def _field_op(prob, op, fields, nout):
    g = prob.grid
    keep = []
    ins = (C.c_void_p * 3)()
    for c, a in enumerate(fields):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=prob.T), (g.nz, g.ny, g.nx)))
        keep.append(a)
        ins[c] = a.ctypes.data
    outs = [np.empty((g.nz, g.ny, g.nx), prob.T) for _ in range(nout)]
    optr = (C.c_void_p * 3)(*[o.ctypes.data for o in outs])
    check(lib().mhd_field_op(prob._h, op, ins, optr))
    return outs


def Curl(prob, B1, B2, B3):
    """``Curl(B1, B2, B3)`` (``src/utils/VectorCalculus.jl:27-50``) on the problem's grid:
    C2R(im k × R2C(B)) in T; real ``[z][y][x]`` fields in and out."""
    return tuple(_field_op(prob, _OP["curl"], (B1, B2, B3), 3))


def Div(prob, B1, B2, B3):
    """``Div(B1, B2, B3)`` (``src/utils/VectorCalculus.jl:70-92``): C2R(im k · R2C(B))."""
    return _field_op(prob, _OP["div"], (B1, B2, B3), 1)[0]


def partial_i(prob, B1, direction):
    """``∂i(B1, direction)`` (``src/utils/VectorCalculus.jl:101-123``), direction "x"/"y"/"z"."""
    if direction not in ("x", "y", "z"):
        raise MHDError(-1, "Wrong driection declared")
    return _field_op(prob, _OP[direction], (B1,), 1)[0]


def LaplaceSolver(prob, B):
    """``LaplaceSolver(B)`` (``src/utils/VectorCalculus.jl:146-168``): C2R(R2C(B) k^-2) with the
    reference's sign (it returns Φ with -ΔΦ = B for zero-mean B)."""
    return _field_op(prob, _OP["laplace"], (B,), 1)[0]


def VectorPotential(prob, B1, B2, B3):
    """``VectorPotential(B1, B2, B3)`` (``src/utils/MHDAnalysis.jl:129-180``), Coulomb gauge:
    C2R(im (k × B̂) k^-2)."""
    return tuple(_field_op(prob, _OP["vecpot"], (B1, B2, B3), 3))


def h_k(prob, iv, jv, kv, L=2 * math.pi):
    """``h_k(iv, jv, kv; L)`` (``src/utils/MHDAnalysis.jl:94-111``): (curl v · v) dV, the product in
    T times the Float64 dV = (L/nx)(L/ny)(L/nz) (a Float64 array, as the reference returns)."""
    g = prob.grid
    c = Curl(prob, iv, jv, kv)
    T = prob.T
    v = [np.asarray(a, dtype=T) for a in (iv, jv, kv)]
    dV = (L / g.nx) * (L / g.ny) * (L / g.nz)
    return (c[0] * v[0] + c[1] * v[1] + c[2] * v[2]).astype(T) * dV


def h_m(prob, ib, jb, kb):
    """``h_m(ib, jb, kb)`` (``src/utils/MHDAnalysis.jl:113-117``): A · B pointwise in T (no dV)."""
    A = VectorPotential(prob, ib, jb, kb)
    T = prob.T
    b = [np.asarray(a, dtype=T) for a in (ib, jb, kb)]
    return A[0] * b[0] + A[1] * b[1] + A[2] * b[2]


def getL(prob, iv, jv, kv):
    """``getL(iv, jv, kv)`` (``src/utils/MHDAnalysis.jl:187-203``): (y kv, z iv, x jv - y iv) with
    the grid points; pointwise, on the host."""
    g = prob.grid
    x, y, z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    T = prob.T
    iv, jv, kv = (np.asarray(a, dtype=T) for a in (iv, jv, kv))
    return (y * kv - 0), -(0 - z * iv), (x * jv - y * iv)


def DivVCorrection(prob):
    """``DivVCorrection!(prob)`` (``src/Solver/VPSolver.jl:100-139``): remove the gradient part
    of u from sol (phi = -i (k . u) / k^2, u -= i k phi) and refresh vars.ux..uz."""
    check(lib().mhd_div_correction(prob._h, 0))


def DivBCorrection(prob):
    """``DivBCorrection!(prob)`` (``src/Solver/VPSolver.jl:60-98``): the same for b (MHD)."""
    check(lib().mhd_div_correction(prob._h, 1))


def SFr(prob, mode="1D"):
    """Velocity structure functions of the current vars (``src/utils/TurbStatTool.jl``):
    ``mode="1D"`` is ``SF₂1D(ux, uy, uz)`` (:88-117), a vector of 2 ceil(|(nx/2, ny/2, nz/2)|)
    radial bins; ``mode="Global"`` is ``SFr(...; mode="Global")`` = ``SFr_Global`` (:119-160), the
    (nrow, nrow) matrix ``SFVr[rpar, rperp]`` against the mean magnetic field.  Empty bins are NaN
    (the reference's 0/0).  Returns (values, counts)."""
    m = {"1D": 0, "Global": 1}.get(mode)
    if m is None:
        raise MHDError(-1, f"mode must be '1D' or 'Global', not {mode!r}")
    nb, nrow = C.c_int32(), C.c_int32()
    check(lib().mhd_structure_function(prob._h, m, None, 0, C.byref(nb), C.byref(nrow)))
    out = np.empty(2 * nb.value, np.float64)
    check(lib().mhd_structure_function(prob._h, m, out.ctypes.data_as(C.c_void_p), out.size, C.byref(nb),
                                       C.byref(nrow)))
    vals, cnt = out[:nb.value], out[nb.value:]
    if m == 1:   # column-major [rpar, rperp] -> numpy [rpar, rperp]
        vals = vals.reshape(nrow.value, nrow.value).T
        cnt = cnt.reshape(nrow.value, nrow.value).T
    return vals, cnt


# ---- checkpoint / restart (src/integrator.jl:208-288, src/utils/IC.jl:245-257) ----
CHECKPOINT_SUFFIX = ".mhdf"
_DTYPES = {1: np.float32, 2: np.float64, 3: np.complex64, 4: np.complex128, 5: np.int64}


def savefile(prob, file_number, file_path_and_name="", exact=False):
    """``savefile(prob, file_number; file_path_and_name)`` (``src/integrator.jl:259-288``): writes
    ``<file_path_and_name>_t_NNNN.mhdf`` with the datasets i/j/k_velocity, i/j/k_mag_field and
    time (the library's container; libhdf5 is not available).  ``exact=True`` adds the spectral
    state and clock for a bit-exact ``Restart``.  Returns the file name."""
    name = f"{file_path_and_name}_t_{int(file_number):04d}{CHECKPOINT_SUFFIX}"
    check(lib().mhd_savefile(prob._h, name.encode(), 1 if exact else 0))
    return name


def Restart(prob, file_path_and_name, exact=False):
    """``Restart!(prob, file_path_and_name)`` (``src/integrator.jl:208-257``): real fields into
    vars and sol, clock.t from the file; ``exact=True`` restores the saved spectral state and
    clock instead (bit-exact continuation)."""
    check(lib().mhd_restart(prob._h, str(file_path_and_name).encode(), 1 if exact else 0))


def read_checkpoint(path):
    """All datasets of a checkpoint file as numpy arrays (real fields in ``[z][y][x]`` order)."""
    import struct
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] != b"MHDFLOWS":
        raise MHDError(-1, f"{path} is not an MHDFlows checkpoint file")
    ver, cnt = struct.unpack_from("<II", buf, 8)
    if ver != 1:
        raise MHDError(-1, f"{path}: unknown version {ver}")
    out = {}
    for i in range(cnt):
        name, dt, nd, d0, d1, d2, d3, off, nb = struct.unpack_from("<48sII4QQQ", buf, 16 + 104 * i)
        dims = (d0, d1, d2, d3)[:nd]
        a = np.frombuffer(buf, dtype=_DTYPES[dt], count=nb // np.dtype(_DTYPES[dt]).itemsize, offset=off)
        out[name.rstrip(b"\0").decode()] = a.reshape(tuple(reversed(dims))) if nd else a[0]
    return out


def readMHDFlows(FileName):
    """``readMHDFlows(FileName)`` (``src/utils/IC.jl:245-257``): velocity and magnetic fields as
    Float32 arrays and the time."""
    d = read_checkpoint(FileName)
    f32 = [np.asarray(d[k], dtype=np.float32) for k in ("i_velocity", "j_velocity", "k_velocity",
                                                          "i_mag_field", "j_mag_field", "k_mag_field")]
    return (*f32, d["time"])


def set_forcing(prob, seed, kf=2.0, amp=1.0):
    """Seeded large-scale solenoidal forcing (the ``calcF!`` hook, ``src/pgen.jl:231-234``)."""
    check(lib().mhd_set_forcing(prob._h, int(seed), float(kf), float(amp)))


def autotune(prob, reps=5):
    """Pick the fastest compiled variant of each hot launch for this grid (see mhd_autotune)."""
    buf = C.create_string_buffer(4096)
    check(lib().mhd_autotune(prob._h, int(reps), buf, len(buf)))
    return buf.value.decode()


def variants(prob, set_to=None):
    """Read (and optionally set) the launch variants [zinv, yinv, yfwd, zfwd, row]."""
    out = (C.c_int32 * 5)()
    src = None
    if set_to is not None:
        src = (C.c_int32 * 5)(*[int(v) for v in set_to])
    check(lib().mhd_set_variants(prob._h, src, out))
    return list(out)


def set_pipeline(prob, slabs):
    """Slab pipeline of each right-hand side (see ``mhd_set_pipeline``; 1 = off): the same
    results, the row kernel of one slab overlapping the y passes of another."""
    check(lib().mhd_set_pipeline(prob._h, int(slabs)))


def set_step_pruning(prob, enable=True):
    """Dealiased steps (see ``mhd_set_step_pruning``): every step of a call but the last
    updates kept modes only; the state after the call is the same either way."""
    check(lib().mhd_set_step_pruning(prob._h, 1 if enable else 0))


def refresh_vars(prob):
    check(lib().mhd_refresh_vars(prob._h))
