"""Ensembles of independent runs stepped together on one MI355X (``mhd_batch_*``).

Runs with the same grid and precision form a group that shares every launch (per-run
pointers, parameters and device clocks in a descriptor array); each group's step is one
captured HIP graph and the groups run on their own streams, so uneven grids overlap.
"""

import ctypes as C
import math

import numpy as np

from . import _lib as L
from ._lib import MHDError, check, lib


def make_config(nx, ny=None, nz=None, L=2 * math.pi, nu=0.0, eta=0.0, T=np.float32, B_field=True,
                stepper="RK4", dt=0.0, device=0, hbm_budget=0, nnu=0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return L_config(nx, ny, nz, L, nu, eta, T, B_field, stepper, dt, device, hbm_budget, nnu)


def L_config(nx, ny, nz, Lb, nu, eta, T, B_field, stepper, dt, device, hbm_budget, nnu):
    T = np.dtype(T).type
    return L.Config(nx=nx, ny=ny, nz=nz, Lx=Lb, Ly=Lb, Lz=Lb, nu=nu, eta=eta if B_field else 0.0,
                    n_nu=nnu, n_eta=0, precision=L.MHD_F32 if T == np.float32 else L.MHD_F64,
                    physics=L.MHD_PHYS_MHD if B_field else L.MHD_PHYS_HD,
                    stepper=L.MHD_RK4 if stepper == "RK4" else L.MHD_LSRK54,
                    device=device, stream=None, hbm_budget=hbm_budget, dt=dt)


def query_bytes(cfg):
    b = C.c_uint64()
    check(lib().mhd_query_bytes(C.byref(cfg), C.byref(b)))
    return b.value


class Ensemble:
    """A batch of runs (``configs``: list of ``Config`` from :func:`make_config`)."""

    def __init__(self, configs):
        n = len(configs)
        arr = (L.Config * n)(*configs)
        h = C.c_void_p()
        check(lib().mhd_batch_create(arr, n, C.byref(h)))
        self._h = h
        self.configs = list(configs)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().mhd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.configs)

    def _types(self, run):
        c = self.configs[run]
        T = np.float32 if c.precision == L.MHD_F32 else np.float64
        Cx = np.complex64 if T == np.float32 else np.complex128
        nf = 6 if c.physics == L.MHD_PHYS_MHD else 3
        return c, T, Cx, nf

    def set_ic_modes(self, run, idx, vals):
        c, T, Cx, nf = self._types(run)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=Cx)
        if vals.shape != (idx.shape[0], nf):
            raise MHDError(-1, "vals must be (M, nf)")
        check(lib().mhd_batch_set_sol_modes(self._h, run, idx.shape[0], idx.ctypes.data_as(C.c_void_p),
                                            vals.ctypes.data_as(C.c_void_p)))

    def set_ic_real(self, run, fields):
        c, T, Cx, nf = self._types(run)
        arrs = [np.ascontiguousarray(f, dtype=T) for f in fields]
        ptrs = (C.c_void_p * nf)(*[a.ctypes.data for a in arrs])
        check(lib().mhd_batch_set_ic_real(self._h, run, ptrs))

    def get_sol_modes(self, run, idx):
        """Entries ``(kx, l, m)`` (rows of ``idx``) of run ``run``'s state, shape ``(M, nf)``."""
        c, T, Cx, nf = self._types(run)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.empty((idx.shape[0], nf), Cx)
        check(lib().mhd_batch_get_sol_modes(self._h, run, idx.shape[0], idx.ctypes.data_as(C.c_void_p),
                                            out.ctypes.data_as(C.c_void_p)))
        return out

    def get_sol(self, run):
        c, T, Cx, nf = self._types(run)
        out = np.empty((nf, c.nz, c.ny, c.nx // 2 + 1), Cx)
        check(lib().mhd_batch_get_sol(self._h, run, out.ctypes.data_as(C.c_void_p)))
        return out

    def clock(self, run):
        t, dt, st = C.c_double(), C.c_double(), C.c_int64()
        check(lib().mhd_batch_get_clock(self._h, run, C.byref(t), C.byref(dt), C.byref(st)))
        return t.value, dt.value, st.value

    def set_cfl(self, run, coef):
        check(lib().mhd_batch_set_cfl(self._h, run, float(coef)))

    def set_params(self, run, nu, eta, nnu=0):
        check(lib().mhd_batch_set_params(self._h, run, float(nu), float(eta), int(nnu)))

    def start_run(self, run, dt=0.0, cfl=0.0, target=2 ** 62):
        check(lib().mhd_batch_start_run(self._h, run, float(dt), float(cfl), int(target)))

    def step(self, nsteps=1):
        check(lib().mhd_batch_step(self._h, int(nsteps)))

    def launch(self, nsteps=1):
        check(lib().mhd_batch_launch(self._h, int(nsteps)))

    def busy(self):
        b = C.c_int32()
        check(lib().mhd_batch_query(self._h, C.byref(b)))
        return bool(b.value)

    def sync(self):
        check(lib().mhd_batch_sync(self._h))

    def set_step_pruning(self, enable=True):
        check(lib().mhd_batch_set_step_pruning(self._h, 1 if enable else 0))

    def autotune(self, reps=5):
        check(lib().mhd_batch_autotune(self._h, int(reps)))

    def use_graph(self, enable=True):
        check(lib().mhd_batch_use_graph(self._h, 1 if enable else 0))

    def multi(self, enable=None):
        """Multi-size launches: True/1 for every eligible group, 2 with the largest group of each
        set on its own stream, False/0 off, None keeps the setting; returns how many size groups
        are stepped by them (groups of one precision, physics and stepper with lengths in
        {32, 48, 64, 96, 128} and no forcing module step as one launch per kernel class)."""
        n = C.c_int32()
        mode = -1 if enable is None else int(enable)
        check(lib().mhd_batch_multi(self._h, mode, C.byref(n)))
        return n.value

    def multi_classes(self, col=None, row=None):
        """Column passes / row kernel of the multi-size sets as one multi-size launch (True) or
        each group's own launch (False); None keeps the choice.  Returns (col, row) of the first
        set ((-1, -1) if there is none); ``autotune`` sets it by timing both forms."""
        out = (C.c_int32 * 2)()
        v = lambda x: -1 if x is None else (1 if x else 0)
        check(lib().mhd_batch_multi_classes(self._h, v(col), v(row), out))
        return int(out[0]), int(out[1])

    def time_steps(self, nsteps):
        ms = C.c_double()
        check(lib().mhd_batch_time_steps(self._h, int(nsteps), C.byref(ms)))
        return ms.value
