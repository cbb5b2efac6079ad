"""Slab pipeline of the right-hand side (mhd_set_pipeline): the y-inverse pass, the row kernel
and the y-forward pass over slabs of z-planes on two streams give bitwise the state of the
unsplit launches, for any slab count (row blocks and slabs need not divide evenly)."""

import math
import sys

import numpy as np
import pytest

import mhdflows_amd as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _fields(shape, seed=3):
    """Six smooth real fields: a few random low-wavenumber Fourier modes on the grid."""
    nx, ny, nz = shape
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.arange(nx) * 2 * math.pi / nx
    y = np.arange(ny) * 2 * math.pi / ny
    z = np.arange(nz) * 2 * math.pi / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    out = []
    for _ in range(6):
        f = np.zeros_like(X)
        for _ in range(6):
            kx, ky, kz = rng.integers(-4, 5, size=3)
            f += rng.normal() * np.cos(kx * X + ky * Y + kz * Z + rng.uniform(0, 2 * math.pi))
        out.append(np.ascontiguousarray(f))  # [z][y][x]
    return out


def _run(shape, T, slabs, stepper, graph=True):
    nx, ny, nz = shape
    p = M.Problem("MI355X", nx=nx, ny=ny, nz=nz, ν=2e-3, η=2e-3, B_field=True, T=T, stepper=stepper)
    M.SetUpProblemIC(p, *_fields(shape))
    M.set_pipeline(p, slabs)
    if not graph:
        M.lib().mhd_use_graph(p._h, 0)
    M.TimeIntegrator(p, math.inf, 3, CFL_Coef=0.3)
    M.TimeIntegrator(p, math.inf, 1, CFL_Coef=0.3)
    out = p.sol.copy()
    del p
    return out


@pytest.mark.parametrize("shape,T,stepper", [((64, 64, 64), np.float32, "RK4"), ((32, 48, 96), np.float64, "RK4"),
                                             ((48, 64, 32), np.float32, "LSRK54")])
def test_pipeline_slabs_bitwise(shape, T, stepper):
    ref = _run(shape, T, 1, stepper)
    assert np.isfinite(ref).all()
    for slabs in (2, 3, 5):
        say(f"pipeline {shape} {np.dtype(T).name} {stepper}: {slabs} slabs")
        got = _run(shape, T, slabs, stepper)
        assert np.array_equal(got, ref), slabs
    got = _run(shape, T, 3, stepper, graph=False)
    assert np.array_equal(got, ref)


def test_pipeline_refuses_bad_counts():
    p = M.Problem("MI355X", nx=16, ν=1e-2, η=1e-2, B_field=True, T=np.float32)
    for bad in (0, 17):
        with pytest.raises(M.MHDError):
            M.set_pipeline(p, bad)
