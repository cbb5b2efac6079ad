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


# ---- structure functions (TurbStatTool.jl) ---------------------------------------------
def test_cf_equals_brute_force_correlation():
    """CF[lag] = sum_x V(x) V(x + lag), fftshifted: a direct O(N^2) sum on a 4 x 6 x 8 cube."""
    from oracle.analysis import cf
    rng = np.random.default_rng(0)
    V = rng.standard_normal((4, 6, 8))
    B = np.zeros_like(V)
    for a in range(4):
        for b in range(6):
            for c in range(8):
                B[a, b, c] = np.sum(V * np.roll(V, (-a, -b, -c), axis=(0, 1, 2)))
    assert np.max(np.abs(np.fft.fftshift(B) - cf(V))) < 1e-12 * np.max(np.abs(B))


def _binned_by_signed_lags(V3, key):
    """SFV of the three components at every signed lag s in [-n/2, n/2) of each axis, binned by
    key(s + 1): the reference's idx = i - n/2 at the fftshifted 1-based index i is s + 1."""
    nz, ny, nx = V3[0].shape
    C = sum(np.real(np.fft.ifftn(np.abs(np.fft.fftn(V)) ** 2)) for V in V3)   # unshifted CF sum
    m2 = 2 * sum(np.mean(V) for V in V3)
    sums, cnt = {}, {}
    for sz in range(-nz // 2, nz // 2):
        for sy in range(-ny // 2, ny // 2):
            for sx in range(-nx // 2, nx // 2):
                b = key(sx + 1, sy + 1, sz + 1)
                if b is None:
                    continue
                v = m2 - 2 * C[sz % nz, sy % ny, sx % nx]
                sums[b] = sums.get(b, 0.0) + v
                cnt[b] = cnt.get(b, 0) + 1
    return sums, cnt


def test_sf2_1d_bins_lags_shifted_by_one():
    from oracle.analysis import sf2_1d
    rng = np.random.default_rng(1)
    v = [rng.standard_normal((6, 8, 10)) + c for c in (0.3, -0.2, 0.0)]
    got, mask = sf2_1d(*v)
    assert len(got) == 2 * int(np.ceil(np.sqrt(5 ** 2 + 4 ** 2 + 3 ** 2)))

    def key(i, j, k):
        kk = int(np.rint(np.sqrt(i * i + j * j + k * k)))
        return kk if kk > 0 else None
    sums, cnt = _binned_by_signed_lags(v, key)
    for kk in range(1, len(got) + 1):
        if kk in cnt:
            assert mask[kk - 1] == cnt[kk] and np.isclose(got[kk - 1], sums[kk] / cnt[kk], rtol=1e-12, atol=1e-9)
        else:
            assert mask[kk - 1] == 0 and np.isnan(got[kk - 1])
    # the zero lag sits at idx = 1, so bin 1 holds it: round(sqrt(3)) = 2 does not, and the
    # lag (-1, -1, -1) (idx = 0) is the one excluded
    assert mask.sum() == 6 * 8 * 10 - 1


def test_sfr_global_axis_aligned_mean_field():
    """Mean field along x: theta is the angle to the x axis, so rpar = |idx| and
    rperp = round(sqrt(jdx^2 + kdx^2))."""
    from oracle.analysis import sfr_global
    rng = np.random.default_rng(2)
    v = [rng.standard_normal((8, 8, 8)) for _ in range(3)]
    b = [np.full((8, 8, 8), 1.3), np.zeros((8, 8, 8)), np.zeros((8, 8, 8))]
    got, mask = sfr_global(*v, *b)

    def key(i, j, k):
        if i == j == k == 0:
            return None
        return abs(i), int(np.rint(np.sqrt(j * j + k * k)))
    sums, cnt = _binned_by_signed_lags(v, key)
    R = int(np.ceil(np.sqrt(3 * 16)))
    assert got.shape == (R + 1, R + 1)
    for (a, c), n in cnt.items():
        assert mask[a, c] == n and np.isclose(got[a, c], sums[(a, c)] / n, rtol=1e-12, atol=1e-9)
    assert mask.sum() == sum(cnt.values()) == 8 ** 3 - 1
