"""Multi-size launches: the size groups of an ensemble stepped by one launch per column pass,
row kernel and spectral update.  Every run ends bitwise where the same ensemble stepped
group by group (each group's own launches on its own stream) ends, and where the run alone
ends; a member is checked against the oracle."""

import math
import sys

import numpy as np
import pytest

import mhdflows_amd as M
from mhdflows_amd import ensemble as E
from mhdflows_amd import ic
from oracle import OracleProblem
from oracle.integrator import rel_l2, time_integrator
from oracle.rhs import c2r

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _specs(sizes, per, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [(n, float(rng.uniform(2, 5)), float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-2)))),
             int(rng.integers(1 << 30))) for n in sizes for _ in range(per)]


def _modes(n, k0, seed):
    return ic.solenoidal_modes(seed, k0=k0, kmax=min(12.0, n // 3 - 1))


def _run(specs, T, stepper, physics_b, steps, multi, graph=True, chunks=(None,), classes=(None, None)):
    ens = E.Ensemble([E.make_config(n, nu=nu, eta=nu, T=T, stepper=stepper, B_field=physics_b)
                      for (n, k0, nu, s) in specs])
    fused = ens.multi(multi)
    if multi:
        got = ens.multi_classes(*classes)
        assert got == tuple(1 if c is None else int(c) for c in classes)
    ens.use_graph(graph)
    for r, (n, k0, nu, s) in enumerate(specs):
        idx, vals = _modes(n, k0, s).half_entries(n)
        if not physics_b:
            idx, vals = idx, vals[:, :3]
        ens.set_ic_modes(r, idx, vals)
        ens.set_cfl(r, 0.3)
    for c in chunks:
        ens.step(steps if c is None else c)
    out = [ens.get_sol(r) for r in range(len(specs))], [ens.clock(r) for r in range(len(specs))]
    ens.close()
    return fused, out


@pytest.mark.parametrize("T,stepper,physics_b", [(np.float32, "RK4", True), (np.float32, "LSRK54", True),
                                                 (np.float64, "RK4", True), (np.float32, "RK4", False)])
def test_multi_size_equals_group_launches(T, stepper, physics_b):
    specs = _specs((32, 48, 64, 96, 128) if T == np.float32 else (32, 48, 64), 2, 7)
    say(f"multi-size: {len(specs)} runs, {np.dtype(T).name} {stepper} {'MHD' if physics_b else 'HD'}")
    nfused, (got, clk) = _run(specs, T, stepper, physics_b, 3, True)
    ngroups = len({s[0] for s in specs})
    assert nfused == ngroups
    nsolo, (ref, rclk) = _run(specs, T, stepper, physics_b, 3, False)
    assert nsolo == 0
    for r in range(len(specs)):
        assert np.array_equal(got[r], ref[r]), (r, specs[r][0])
        assert clk[r] == rclk[r]
        assert np.isfinite(got[r]).all()


@pytest.mark.parametrize("classes", [(False, True), (True, False), (False, False)])
def test_multi_size_class_choices(classes):
    """Column passes and row kernels as each group's own launch on the multi-size stream (the
    spectral update stays multi-size): the same state."""
    specs = _specs((32, 48, 64, 96), 2, 9)
    _, (a, _) = _run(specs, np.float32, "RK4", True, 3, True)
    _, (b, _) = _run(specs, np.float32, "RK4", True, 3, True, classes=classes)
    for r in range(len(specs)):
        assert np.array_equal(a[r], b[r]), (r, classes)


def test_multi_size_without_graph_and_in_chunks():
    """Direct launches equal the captured step, and a call split into chunks (each ending
    with a full step) equals the same steps in one call."""
    specs = _specs((32, 64, 96), 2, 11)
    _, (a, _) = _run(specs, np.float32, "RK4", True, 4, True, graph=True, chunks=(2, 2))
    _, (b, _) = _run(specs, np.float32, "RK4", True, 4, True, graph=False, chunks=(2, 2))
    _, (c, _) = _run(specs, np.float32, "RK4", True, 4, False, graph=True, chunks=(2, 2))
    for r in range(len(specs)):
        assert np.array_equal(a[r], b[r]), r
        assert np.array_equal(a[r], c[r]), r


def test_multi_size_member_alone_and_oracle():
    specs = _specs((48, 96), 3, 5)
    _, (got, clk) = _run(specs, np.float32, "LSRK54", True, 3, True)
    for r in (0, len(specs) - 1):
        n, k0, nu, s = specs[r]
        p = M.Problem("MI355X", nx=n, ν=nu, η=nu, B_field=True, T=np.float32, stepper="LSRK54")
        idx, vals = _modes(n, k0, s).half_entries(n)
        M.SetUpProblemIC_modes(p, idx, vals)
        M.TimeIntegrator(p, math.inf, 2, CFL_Coef=0.3)
        assert np.array_equal(got[r], p.sol), r
        del p
        say(f"multi-size member {r} ({n}^3) against the oracle")
        o = OracleProblem(nx=n, nu=nu, eta=nu, B_field=True, T=np.float32, stepper="LSRK54", mode="fused")
        o.sol[...] = _modes(n, k0, s).spectrum_on_grid(n)
        o.vars[...] = c2r(o.sol, (n, n, n)).astype(np.float32)
        time_integrator(o, math.inf, 2, CFL_Coef=0.3)
        err = rel_l2(got[r], o.sol)
        say(f"  rel L2 {err:.3e}")
        assert err < 1e-5


def test_multi_size_largest_group_alone():
    """Mode 2: the group with the most grid points steps on its own stream beside the
    multi-size launches of the others; the same state as every group on its own stream."""
    specs = _specs((32, 48, 64, 96), 2, 13)
    res = []
    for mode in (2, 0):
        ens = E.Ensemble([E.make_config(n, nu=nu, eta=nu, T=np.float32) for (n, k0, nu, s) in specs])
        assert ens.multi(mode) == (3 if mode == 2 else 0)
        for r, (n, k0, nu, s) in enumerate(specs):
            idx, vals = _modes(n, k0, s).half_entries(n)
            ens.set_ic_modes(r, idx, vals)
            ens.set_cfl(r, 0.3)
        ens.step(3)
        res.append([ens.get_sol(r) for r in range(len(specs))])
        ens.close()
    for r in range(len(specs)):
        assert np.array_equal(res[0][r], res[1][r]), r


def test_multi_size_mixed_with_solo_groups():
    """Groups outside the multi-size lengths (16^3, 256^3) and a Float64 group step on their own
    streams beside the multi-size set; all end as when every group steps alone."""
    cfgs, specs = [], []
    for (n, T) in ((16, np.float32), (32, np.float32), (64, np.float32), (32, np.float64), (256, np.float32)):
        specs.append((n, T, 3.0, 2e-3, 100 + n))
        cfgs.append(E.make_config(n, nu=2e-3, eta=2e-3, T=T))
    res = []
    for multi in (True, False):
        ens = E.Ensemble(cfgs)
        nf = ens.multi(multi)
        assert nf == (2 if multi else 0)
        for r, (n, T, k0, nu, s) in enumerate(specs):
            idx, vals = _modes(n, k0, s).half_entries(n)
            ens.set_ic_modes(r, idx, vals)
            ens.set_cfl(r, 0.3)
        ens.step(2)
        res.append([ens.get_sol(r) for r in range(len(specs))])
        ens.close()
    for r in range(len(specs)):
        assert np.array_equal(res[0][r], res[1][r]), r
