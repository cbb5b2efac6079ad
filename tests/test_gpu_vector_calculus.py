"""Vector calculus of host fields on the MI355X (src/utils/VectorCalculus.jl,
src/utils/MHDAnalysis.jl) against closed forms on Arnold-Beltrami-Childress and trigonometric
fields: Curl(u) = k u for an ABC field, VectorPotential(b) = b / k, Div(grad phi) = lap phi,
∂i, LaplaceSolver (the reference's sign: C2R(R2C(B) k^-2)), h_k = k |u|^2 dV, h_m = |b|^2 / k,
getL's pointwise formula.  Float64 to 1e-12 of the field scale, Float32 to 2e-6."""

import math

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import ic

pytestmark = pytest.mark.gpu


def _abc(n, k, T):
    X, Y, Z = ic.grid_coords(n)
    A, B, C = 0.9, 0.6, 0.4
    u = (A * np.sin(k * Z) + C * np.cos(k * Y), B * np.sin(k * X) + A * np.cos(k * Z),
         C * np.sin(k * Y) + B * np.cos(k * X))
    return [np.asarray(a, dtype=T) for a in u], (X, Y, Z)


def _close(a, b, T, scale):
    tol = 1e-12 if T == np.float64 else 2e-6
    return np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))) <= tol * scale


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_curl_vector_potential_and_helicities_on_abc(T):
    n, k = 32, 3
    p = M.Problem("MI355X", nx=n, B_field=True, T=T)
    u, _ = _abc(n, k, T)
    c = M.Curl(p, *u)
    for a, b in zip(c, u):
        assert _close(a, k * b.astype(np.float64), T, k)
    A = M.VectorPotential(p, *u)
    for a, b in zip(A, u):
        assert _close(a, b.astype(np.float64) / k, T, 1.0)
    dV = (2 * math.pi / n) ** 3
    u2 = sum(a.astype(np.float64) ** 2 for a in u)
    assert _close(M.h_k(p, *u), k * u2 * dV, T, 3 * k * dV)
    assert _close(M.h_m(p, *u), u2 / k, T, 3.0)


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_div_partials_and_laplace_solver(T):
    n = 32
    p = M.Problem("MI355X", nx=n, ny=24, nz=16, Ly=1.5 * math.pi, Lz=math.pi, T=T)
    x, y, z = (p.grid.x.astype(np.float64)[None, None, :], p.grid.y.astype(np.float64)[None, :, None],
               p.grid.z.astype(np.float64)[:, None, None])
    ky, kz = 2 * math.pi / (1.5 * math.pi), 2 * math.pi / math.pi        # fundamental wavenumbers
    phi = np.sin(3 * x) * np.cos(ky * y) + np.cos(2 * kz * z)
    gx = 3 * np.cos(3 * x) * np.cos(ky * y)
    gy = -ky * np.sin(3 * x) * np.sin(ky * y)
    gz = -2 * kz * np.sin(2 * kz * z)
    lap = -(9 + ky ** 2) * np.sin(3 * x) * np.cos(ky * y) - (2 * kz) ** 2 * np.cos(2 * kz * z)
    for d, want in (("x", gx), ("y", gy), ("z", gz)):
        assert _close(M.partial_i(p, phi, d), want, T, 10.0), d
    assert _close(M.Div(p, gx, gy, gz), lap, T, 20.0)
    # LaplaceSolver multiplies by k^-2: cos(2x) -> cos(2x)/4, the mean is dropped
    B = np.cos(2 * x) + np.sin(ky * y + kz * z) + 0.3
    want = np.cos(2 * x) / 4 + np.sin(ky * y + kz * z) / (ky ** 2 + kz ** 2)
    assert _close(M.LaplaceSolver(p, B), np.broadcast_to(want, B.shape), T, 1.0)
    with pytest.raises(M.MHDError):
        M.partial_i(p, phi, "w")


def test_getL_pointwise():
    n = 16
    p = M.Problem("MI355X", nx=n, T=np.float64)
    rng = np.random.default_rng(0)
    iv, jv, kv = (rng.standard_normal((n, n, n)) for _ in range(3))
    Li, Lj, Lk = M.getL(p, iv, jv, kv)
    x, y, z = p.grid.x[None, None, :], p.grid.y[None, :, None], p.grid.z[:, None, None]
    assert np.array_equal(Li, y * kv) and np.array_equal(Lj, z * iv) and np.array_equal(Lk, x * jv - y * iv)
