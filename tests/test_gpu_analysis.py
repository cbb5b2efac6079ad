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


# ---- structure functions (src/utils/TurbStatTool.jl) -------------------------------------
def _with_mean_field(n, ny, nz, T):
    """An MHD problem whose b has a mean field (b0 = (0.7, -0.4, 0.25)) on top of seeded modes."""
    from test_gpu_parity import _resample
    ms = ic.solenoidal_modes(3, k0=2.0, kmax=4.0, mhd=True)
    f = ms.real_fields(n, nf=6, dtype=np.float64) if n == ny == nz else _resample(ms, n, ny, nz, True)
    f = [np.asarray(a) for a in f]
    for c, b0 in zip((3, 4, 5), (0.7, -0.4, 0.25)):
        f[c] = f[c] + b0
    p = M.Problem("MI355X", nx=n, ny=ny, nz=nz, ν=0.01, η=0.01, B_field=True, T=T, dt=1e-3)
    M.SetUpProblemIC(p, *f)
    return p


@pytest.mark.parametrize("shape", [(32, 32, 32), (16, 24, 32)])
@pytest.mark.parametrize("T,tol", [(np.float64, 1e-11), (np.float32, 2e-5)])
def test_structure_function_1d_matches_oracle(shape, T, tol):
    from oracle.analysis import sf2_1d
    p = _with_mean_field(*shape, T)
    M.stepforward(p, 1)
    got, cnt = M.SFr(p, "1D")
    ref, rcnt = sf2_1d(*[p.vars.field(f) for f in range(3)])
    assert np.array_equal(cnt, rcnt)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(got[ok] - ref[ok])) <= tol * np.max(np.abs(ref[ok]))


@pytest.mark.parametrize("shape", [(32, 32, 32), (16, 24, 32)])
def test_structure_function_global_matches_oracle(shape):
    from oracle.analysis import sfr_global
    p = _with_mean_field(*shape, np.float64)
    M.stepforward(p, 1)
    got, cnt = M.SFr(p, "Global")
    ref, rcnt = sfr_global(*[p.vars.field(f) for f in range(6)])
    assert got.shape == ref.shape
    assert np.array_equal(cnt, rcnt)
    ok = ~np.isnan(ref)
    assert np.array_equal(np.isnan(got), ~ok)
    assert np.max(np.abs(got[ok] - ref[ok])) <= 1e-11 * np.max(np.abs(ref[ok]))


def test_structure_function_global_needs_mean_field():
    """With b = 0 the reference's angle is 0/0 and round(Int, NaN) throws; the library refuses."""
    X, Y, Z = ic.grid_coords(16)
    zero = np.zeros_like(X)
    p = M.Problem("MI355X", nx=16, B_field=True, T=np.float64)
    M.SetUpProblemIC(p, np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), zero, zero, zero, zero)
    with pytest.raises(M.MHDError):
        M.SFr(p, "Global")
    _, h = pair(16, T=np.float64, physics="HD", kmax=4.0)
    vals, cnt = M.SFr(h, "1D")
    assert cnt.sum() == 16 ** 3 - 1
    with pytest.raises(M.MHDError):
        M.SFr(h, "Global")


# ---- DivVCorrection! / DivBCorrection! (src/Solver/VPSolver.jl:60-139) --------------------
@pytest.mark.parametrize("T,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
@pytest.mark.parametrize("which", [0, 1])
def test_div_correction_matches_oracle(T, tol, which):
    from oracle import OracleProblem
    from oracle.integrator import rel_l2
    from oracle.vp import div_correction
    n = 32
    X, Y, Z = ic.grid_coords(n)
    # seeded solenoidal modes plus the gradient of sin(2x) cos(y) + cos(3z) and a uniform part
    grad = [2 * np.cos(2 * X) * np.cos(Y), -np.sin(2 * X) * np.sin(Y), -3 * np.sin(3 * Z)]
    sol = ic.solenoidal_modes(7 + which, k0=2.0, kmax=8.0).real_fields(n, dtype=np.float64)
    fields = [np.asarray(sol[c]) + grad[c % 3] + 0.1 * c for c in range(6)]
    o = OracleProblem(nx=n, B_field=True, T=T)
    o.set_ic(*fields)
    p = M.Problem("MI355X", nx=n, B_field=True, T=T)
    M.SetUpProblemIC(p, *fields)
    div_correction(o, which)
    (M.DivVCorrection if which == 0 else M.DivBCorrection)(p)
    assert rel_l2(p.sol, o.sol) < tol
    for f in range(6):
        assert rel_l2(p.vars.field(f), o.vars[f]) < tol, f
    d = M.Diagnostics(p)
    assert (d["max_div_u"] if which == 0 else d["max_div_b"]) < (1e-10 if T == np.float64 else 1e-3)


def test_div_b_correction_needs_b():
    _, h = pair(16, T=np.float64, physics="HD", kmax=4.0)
    with pytest.raises(M.MHDError):
        M.DivBCorrection(h)
    M.DivVCorrection(h)
