"""Host-side volume-penalisation helpers (CPU only)."""

import math

import numpy as np

import mhdflows_amd as M


def test_cylindrical_mask_matches_the_reference_loop():
    """Cylindrical_Mask_Function (src/utils/IC.jl:15-27): S = 0 where R₂ >= sqrt(x^2 + y^2) >= R₁
    with x, y the grid points (origin -L/2) in T, 1 elsewhere."""
    class G:
        nx, ny, nz, T = 12, 10, 4, np.float32
        x = (-math.pi + np.arange(12) * np.float32(2 * math.pi / 12)).astype(np.float32)
        y = (-math.pi + np.arange(10) * np.float32(2 * math.pi / 10)).astype(np.float32)
    S = M.Cylindrical_Mask_Function(G, R2=0.82 * math.pi, R1=0.3)
    assert S.shape == (4, 10, 12) and S.dtype == np.float32
    for j in range(10):
        for i in range(12):
            R = math.sqrt(float(G.x[i] * G.x[i] + G.y[j] * G.y[j]))
            want = 0 if (0.82 * math.pi >= R >= 0.3) else 1
            assert np.all(S[:, j, i] == want)
    assert 0 < S.sum() < S.size
