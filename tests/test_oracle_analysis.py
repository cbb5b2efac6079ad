"""Pins of the oracle's spectra and scale decomposition against closed forms (CPU only)."""

import numpy as np

from mhdflows_amd import ic
from oracle.analysis import krmax, scale_decomposition, spectralline
from oracle.grid import Grid


def test_single_mode_lands_in_its_shell():
    """cos(k.x) with k = (3, 4, 0), |k| = 5: the half spectrum holds one entry (kx = 3 > 0) of
    magnitude N/2, so Pk = (N/2)^2 at r = 5 and zero elsewhere; every shell up to krmax is hit."""
    n = 16
    g = Grid(n, T=np.float64)
    X, Y, Z = ic.grid_coords(n)
    Pk, count = spectralline(g, np.cos(3 * X + 4 * Y))
    N = n ** 3
    assert np.isclose(Pk[5], (N / 2) ** 2, rtol=1e-12)
    assert np.sum(Pk) - Pk[5] < 1e-12 * Pk[5]
    assert len(Pk) == krmax(g) == int(np.rint(np.sqrt(3) * 8 + 1))
    assert np.all(count > 0)
    assert int(np.sum(count)) == g.nz * g.ny * g.nkr


def test_band_filter_keeps_the_band():
    n = 16
    g = Grid(n, T=np.float64)
    X, Y, Z = ic.grid_coords(n)
    low, high = np.sin(2 * X) * np.cos(Z), np.cos(5 * Y + 3 * Z)     # |k| = sqrt(5), sqrt(34)
    out = scale_decomposition(g, [low + high, 2 * high], 1, 5)
    assert np.max(np.abs(out[0] - low)) < 1e-13
    assert np.max(np.abs(out[1])) < 1e-13
    out = scale_decomposition(g, [low + high], 5.5, 6)
    assert np.max(np.abs(out[0] - high)) < 1e-13
