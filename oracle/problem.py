"""Oracle problem object mirroring ``Problem(dev; ...)`` (test infrastructure only).

Follows ``src/pgen.jl:64-127`` (defaults and which keywords reach the solver),
``src/Structure/datastructure.jl:5-13,27-35,69-108`` (vars/params) and
``src/Problems.jl:118-140`` (clock in T, ``sol`` zeroed).
"""

import math

import numpy as np

from .grid import Grid
from .rhs import r2c


class OracleProblem:
    def __init__(self, nx=64, ny=None, nz=None, Lx=2 * math.pi, Ly=None, Lz=None,
                 nu=0.0, nnu=0, eta=0.0, neta=0, B_field=False, stepper="RK4",
                 T=np.float32, dt=0.0, mode="literal", forcing=None):
        if stepper not in ("RK4", "LSRK54"):
            raise ValueError(f"unsupported stepper {stepper!r}")
        self.grid = Grid(nx, ny, nz, Lx, Ly, Lz, T)
        self.physics = "MHD" if B_field else "HD"
        self.nf = 6 if B_field else 3
        self.stepper = stepper
        self.mode = mode
        # SetParams receives ν, η, nν but not nη (pgen.jl:114-116), so nη stays 0.
        self.nu, self.eta, self.nnu, self.neta = float(nu), float(eta), int(nnu), 0
        g = self.grid
        self.sol = np.zeros((self.nf,) + g.shape_c, dtype=g.C)
        self.vars = np.zeros((self.nf,) + g.shape_r, dtype=g.T)
        Tt = g.T.type
        self.t = Tt(0)
        self.dt = Tt(dt)
        self.step = 0
        self.forcing = forcing
        self.force_counter = 1 << 40   # draws of a right-hand side outside a step (calcN calls)
        self.vp = None                 # volume penalisation: {"chi", "U0", "B0"} real fields (oracle.vp)

    def set_ic(self, ux=None, uy=None, uz=None, bx=None, by=None, bz=None):
        """``SetUpProblemIC!`` (``src/utils/IC.jl:41-109``): copy real fields to vars, R2C into sol."""
        fields = [ux, uy, uz] + ([bx, by, bz] if self.physics == "MHD" else [])
        for f, arr in enumerate(fields):
            if arr is None:
                continue
            self.vars[f] = np.asarray(arr, dtype=self.grid.T)
            self.sol[f] = r2c(self.vars[f]).astype(self.grid.C)

    def copy(self):
        p = OracleProblem.__new__(OracleProblem)
        p.__dict__.update(self.__dict__)
        p.sol = self.sol.copy()
        p.vars = self.vars.copy()
        return p
