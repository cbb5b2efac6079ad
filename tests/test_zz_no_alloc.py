"""mhd_set_sol_modes_noalloc (staging in a library-owned device array, no allocation) gives
the same state as mhd_set_sol_modes, bit for bit, over two calls in a row with the second
call smaller than the first; a call too large for the array is refused and changes nothing."""

import ctypes as C
import sys

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import _lib, ic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _call(fn, p, idx, vals):
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=p.C)
    return fn(p._h, idx.shape[0], idx.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p))


def _same_state(a, b):
    assert np.array_equal(a.sol, b.sol)
    for f in range(a.nf):
        assert np.array_equal(a.vars.field(f), b.vars.field(f)), f


@pytest.mark.parametrize("T,physics", [(np.float32, True), (np.float64, True), (np.float64, False)])
def test_set_sol_modes_noalloc_matches_allocating_entry(T, physics):
    say("noalloc", np.dtype(T).name, "MHD" if physics else "HD")
    L = M.lib()
    kw = dict(nx=32, ν=1e-2, η=1e-2, B_field=physics, T=T, dt=1e-3)
    a = M.Problem("MI355X", **kw)
    b = M.Problem("MI355X", **kw)
    nf = 6 if physics else 3
    for seed, kmax in ((3, 9.0), (4, 5.0)):            # the second call is smaller
        idx, vals = ic.solenoidal_modes(seed, k0=3.0, kmax=kmax, mhd=physics).half_entries(32, nf=nf)
        say(f"  {idx.shape[0]} modes")
        _lib.check(_call(L.mhd_set_sol_modes, a, idx, vals))
        _lib.check(_call(L.mhd_set_sol_modes_noalloc, b, idx, vals))
        _same_state(a, b)
        assert M.getCFL(a, Coef=0.3) == M.getCFL(b, Coef=0.3)
        M.stepforward(a, 2)
        M.stepforward(b, 2)
        _same_state(a, b)
    # more modes than the 4 MiB staging holds: refused, state unchanged
    before = b.sol
    big = np.zeros((80000, 3), np.int32)
    rc = _call(L.mhd_set_sol_modes_noalloc, b, big, np.zeros((80000, nf), b.C))
    assert rc == -1 and b"scratch" in L.mhd_last_error()
    assert np.array_equal(b.sol, before)
    # zero modes: the state becomes zero
    _lib.check(_call(L.mhd_set_sol_modes_noalloc, b, np.zeros((0, 3), np.int32), np.zeros((0, nf), b.C)))
    assert not np.any(b.sol)
