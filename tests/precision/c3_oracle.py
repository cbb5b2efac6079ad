"""Generate the oracle side of the precision run: 256^3 Float64 (fused mode), 50 RK4 steps.

Writes <out>.npz with the retained modes of sol and the FP64 diagnostics.  CPU only;
about 100 s per step on 8 cores.  Usage: python tests/precision/c3_oracle.py OUT.npz
"""

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (ROOT, os.path.join(ROOT, "mhdflows.jl_amd"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)

import c3_setup as S  # noqa: E402
from mhdflows_amd import ic  # noqa: E402
from oracle import OracleProblem  # noqa: E402
from oracle import rhs as orhs  # noqa: E402
from oracle.diagnostics import full_diagnostics  # noqa: E402
from oracle.forcing import Forcing  # noqa: E402
from oracle.integrator import run_steps  # noqa: E402


def main(out):
    orhs.WORKERS = os.cpu_count() or 1
    n = S.N_REF
    o = OracleProblem(nx=n, nu=S.NU, eta=S.ETA, B_field=True, T=np.float64, dt=S.DT, mode="fused",
                      forcing=Forcing(S.F_SEED, kf=S.KF, amp=S.AMP))
    o.sol[...] = ic.solenoidal_modes(S.IC_SEED, k0=S.K0, kmax=S.KMAX).spectrum_on_grid(n)
    t0 = time.perf_counter()
    for s in range(S.STEPS):
        run_steps(o, 1)
        print(f"oracle step {s + 1}/{S.STEPS} t={float(o.t):.4f} ({time.perf_counter() - t0:.0f} s)", flush=True)
    d = full_diagnostics(o)
    np.savez(out, retained=S.retained(o.sol, n), E_kin=d["E_kin"], E_mag=d["E_mag"],
             max_div_b=d["max_div_b"], t=float(o.t), steps=S.STEPS)
    print("wrote", out, d, flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
