"""The reference's forcing modules on the MI355X against the oracle (through the C ABI).

N97 (Taylor-Green dynamo forcing), negative damping and A99 (stochastic, with the counter-based
draws both sides share): Float64 <= 1e-12 relative L2 after steps and for a right-hand side,
Float32 <= 1e-5.  HD discards forcing (src/pgen.jl:173-181), setting a module replaces the
previous one, and dealiased steps give the same state as full steps with every module.
"""

import sys

import numpy as np
import pytest

import mhdflows_amd as M
from oracle.forcing import A99, N97, NegativeDamping
from oracle.integrator import rel_l2, run_steps
from oracle.rhs import calcN as o_calcN

from test_gpu_parity import _calcN_gpu, pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _nd_fields(n, T, seed=2):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.2, 1.5, (n, n, n)).astype(T) for _ in range(3)]


def _setups(kind, n, T):
    """(oracle forcing object, function that sets the same module on a device problem)"""
    if kind == "n97":
        return N97(F0=0.7, kf=2), lambda p: M.SetUpN97(p, F0=0.7, kf=2)
    if kind == "nd":
        f = _nd_fields(n, T)
        return NegativeDamping(0.45, *f), lambda p: M.SetUpND(p, 0.45, *f)
    return A99(seed=9, kf=2.0, P=1.0), lambda p: M.SetUpFk(p, kf=2.0, P=1.0, seed=9)


@pytest.mark.parametrize("kind", ["n97", "nd", "a99"])
@pytest.mark.parametrize("T,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_forcing_module_steps_match_oracle(kind, T, tol):
    say("forcing", kind, T.__name__)
    o, p = pair(32, T=T, kmax=6.0, mode="fused")
    o.forcing, setup = _setups(kind, 32, T)
    setup(p)
    run_steps(o, 4)
    M.stepforward(p, 4)
    err = rel_l2(p.sol, o.sol)
    say(f"  4 steps rel L2 {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("kind", ["n97", "nd", "a99"])
def test_forcing_module_rhs_matches_oracle(kind):
    o, p = pair(32, T=np.float64, kmax=16.0, mode="fused")   # modes in the aliased band too
    o.forcing, setup = _setups(kind, 32, np.float64)
    setup(p)
    N_gpu, sol_after = _calcN_gpu(p, o.sol.copy())
    N_ref = np.zeros_like(o.sol)
    s_ref = o.sol.copy()
    o_calcN(N_ref, s_ref, 0.0, o)      # first calcN call of both: draw counter 2^40
    assert rel_l2(N_gpu, N_ref) < 1e-12
    assert np.array_equal(sol_after, s_ref)


@pytest.mark.parametrize("kind", ["n97", "nd", "a99"])
def test_hd_discards_forcing(kind):
    """HDcalcN! adds the forcing before HDcalcN_advection! zeroes N (src/pgen.jl:176-178)."""
    _, a = pair(16, T=np.float64, physics="HD", kmax=4.0)
    _, b = pair(16, T=np.float64, physics="HD", kmax=4.0)
    _, setup = _setups(kind, 16, np.float64)
    setup(a)
    M.stepforward(a, 3)
    M.stepforward(b, 3)
    assert np.array_equal(a.sol, b.sol)


def test_modules_replace_each_other():
    _, a = pair(16, T=np.float64, kmax=4.0)
    _, b = pair(16, T=np.float64, kmax=4.0)
    _, c = pair(16, T=np.float64, kmax=4.0)
    M.SetUpN97(a, F0=0.7, kf=2)
    M.SetUpND(a, 0.45, *_nd_fields(16, np.float64))
    M.SetUpFk(a, kf=2.0, P=1.0, seed=9)          # a: A99 only
    M.SetUpFk(b, kf=2.0, P=1.0, seed=9)
    M.SetUpN97(c, F0=0.7, kf=2)
    M.set_forcing(c, 0, 0.0, 0.0)                # c: every forcing off
    _, d = pair(16, T=np.float64, kmax=4.0)
    for p in (a, b, c, d):
        M.stepforward(p, 3)
    assert np.array_equal(a.sol, b.sol)
    assert np.array_equal(c.sol, d.sol)
    assert not np.array_equal(a.sol, d.sol)


@pytest.mark.parametrize("kind", ["n97", "nd", "a99"])
def test_dealiased_steps_with_modules_equal_full_steps(kind):
    _, a = pair(24, T=np.float32, kmax=12.0)
    _, b = pair(24, T=np.float32, kmax=12.0)
    _, setup = _setups(kind, 24, np.float32)
    setup(a)
    setup(b)
    M.set_step_pruning(b, False)
    for k in (3, 2):
        M.stepforward(a, k)
        M.stepforward(b, k)
        assert np.array_equal(a.sol, b.sol), k
