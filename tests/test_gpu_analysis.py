"""On-device spectra and scale decomposition against the oracle (Float64 <= 1e-12, Float32
against the Float32 oracle <= 1e-5) and against closed forms."""

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import ic
from oracle.analysis import krmax, scale_decomposition, spectralline

from test_gpu_parity import pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("T,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("physics", ["MHD", "HD"])
def test_spectra_match_oracle(T, tol, physics):
    o, p = pair(32, T=T, physics=physics, kmax=14.0)
    M.stepforward(p, 2)
    d = M.spectra(p)
    assert len(d["count"]) == krmax(o.grid)
    for f, name in enumerate(["ux", "uy", "uz", "bx", "by", "bz"][:p.nf]):
        Pk, count = spectralline(o.grid, p.vars.field(f))
        assert np.array_equal(d["count"], count)
        assert np.max(np.abs(d[name] - Pk)) <= tol * np.max(Pk), name
    Pk, kr = M.spectralline(p, "uy")
    assert np.array_equal(kr, np.arange(1, len(Pk) + 1))   # every shell holds a mode


def test_spectrum_single_mode_closed_form():
    n = 32
    X, Y, Z = ic.grid_coords(n)
    zero = np.zeros_like(X)
    p = M.Problem("MI355X", nx=n, B_field=True, T=np.float64)
    M.SetUpProblemIC(p, np.cos(3 * X + 4 * Y), zero, zero, zero, zero, 0.5 * np.sin(2 * Z))
    d = M.spectra(p)
    N = n ** 3
    assert np.isclose(d["ux"][5], (N / 2) ** 2, rtol=1e-12) and np.sum(d["ux"]) - d["ux"][5] < 1e-9
    # sin(2z) sits on the kx = 0 plane, where the half spectrum holds both k and -k
    assert np.isclose(d["bz"][2], 2 * (N / 4) ** 2, rtol=1e-12)


@pytest.mark.parametrize("T,tol", [(np.float64, 1e-13), (np.float32, 1e-5)])
def test_scale_decomposition_matches_oracle(T, tol):
    o, p = pair(32, T=T, kmax=10.0)
    M.stepforward(p, 1)
    u = [p.vars.field(f) for f in range(3)]
    got = M.ScaleDecomposition(p, ("ux", "uy", "uz"), kf=(2, 6))
    ref = scale_decomposition(o.grid, u, 2, 6)
    for g_, r_ in zip(got, ref):
        assert np.max(np.abs(g_ - r_)) <= tol * np.max(np.abs(r_))
    b = M.ScaleDecomposition(p, "by", kf=(5, 3))
    assert np.max(np.abs(b[0] - scale_decomposition(o.grid, [p.vars.field(4)], 3, 5)[0])) <= tol * 10
    with pytest.raises(M.MHDError):
        M.ScaleDecomposition(p, ("uz", "bx"))
