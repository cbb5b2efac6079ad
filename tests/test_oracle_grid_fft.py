"""Pins of the oracle's grid conventions and FFT semantics (CPU only)."""

import math

import numpy as np
import pytest

from oracle import grid as G
from oracle.rhs import c2r, r2c

# Kept ranges of the FourierFlows 1/3 box rule, y/z axes (signed) and x (half spectrum).
KEPT_TABLE = {32: (-10, 9), 48: (-16, 15), 64: (-21, 20), 96: (-32, 31),
              128: (-42, 41), 256: (-85, 84), 1024: (-341, 340)}


@pytest.mark.parametrize("n", sorted(KEPT_TABLE))
def test_alias_table(n):
    assert G.kept_range(n) == KEPT_TABLE[n]
    mx, my, mz = G.keep_masks(n, n, n)
    # x keeps kr in [0, floor(n/3)-1]
    assert np.flatnonzero(mx).tolist() == list(range(0, KEPT_TABLE[n][1] + 1))
    kept = [i if i < n // 2 else i - n for i in np.flatnonzero(my)]
    assert sorted(kept) == list(range(KEPT_TABLE[n][0], KEPT_TABLE[n][1] + 1))


def test_odd_size_rejected():
    with pytest.raises(ValueError):
        G.Grid(33)


def test_wavenumbers_integer_for_2pi():
    g = G.Grid(16, T=np.float64)
    assert np.array_equal(g.kr.ravel(), np.arange(9))
    assert np.array_equal(g.l.ravel(), np.fft.fftfreq(16, 1 / 16))
    assert g.invKrsq[0, 0, 0] == 0
    assert g.Krsq[0, 2, 3] == 4 + 9


def _dft_matrix(n, sign):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 8), (8, 8, 8)])
def test_r2c_brute_force(shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    nz, ny, nx = shape
    F = np.einsum("zyx,az,by,cx->abc", x, _dft_matrix(nz, -1), _dft_matrix(ny, -1), _dft_matrix(nx, -1))
    assert np.allclose(r2c(x), F[:, :, : nx // 2 + 1], atol=1e-12)


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 8), (8, 8, 8)])
def test_c2r_non_hermitian_brute_force(shape):
    """c2r of an arbitrary half spectrum = inverse DFT of the Hermitian extension of P_H(X)."""
    rng = np.random.default_rng(2)
    nz, ny, nx = shape
    X = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    XH = G.hermitian_project(X)
    full = np.zeros((nz, ny, nx), complex)
    full[:, :, : nx // 2 + 1] = XH
    for kx in range(nx // 2 + 1, nx):
        src = XH[:, :, nx - kx]
        full[:, :, kx] = np.conj(np.roll(np.flip(src, (0, 1)), (1, 1), (0, 1)))
    x = np.einsum("abc,za,yb,xc->zyx", full, _dft_matrix(nz, 1), _dft_matrix(ny, 1), _dft_matrix(nx, 1))
    x = x / (nx * ny * nz)
    assert np.allclose(x.imag, 0, atol=1e-12)
    assert np.allclose(c2r(X, shape), x.real, atol=1e-12)
    # R2C(C2R(X)) is the Hermitian projection
    assert np.allclose(r2c(c2r(X, shape)), XH, atol=1e-12)


def test_dealias_zeroes_expected_planes():
    g = G.Grid(32, T=np.float64)
    X = np.ones((2,) + g.shape_c, complex)
    G.dealias_(X, g)
    assert X[0, 0, 0, 9] == 1 and X[0, 0, 0, 10] == 0
    assert X[0, 0, 9, 0] == 1 and X[0, 0, 10, 0] == 0 and X[0, 0, 32 - 10, 0] == 1 and X[0, 0, 32 - 11, 0] == 0
    assert X[1, 9, 0, 0] == 1 and X[1, 10, 0, 0] == 0 and X[1, 22, 0, 0] == 1 and X[1, 21, 0, 0] == 0


def test_float32_transforms_stay_single():
    x = np.random.default_rng(0).standard_normal((8, 8, 8)).astype(np.float32)
    assert r2c(x).dtype == np.complex64
    assert c2r(r2c(x), x.shape).dtype == np.float32
    assert math.isclose(float(np.max(np.abs(c2r(r2c(x), x.shape) - x))), 0, abs_tol=1e-5)
