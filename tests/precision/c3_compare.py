"""GPU side of the precision run: 1024^3 Float32 (and 256^3 Float64 / Float32 for context)
against the 256^3 Float64 oracle run written by c3_oracle.py, on the 256^3 retained modes.

Usage (MI355X): python tests/precision/c3_compare.py ORACLE.npz [OUT.json]
Prints one JSON object: per-field and total relative L2 errors of the retained modes,
E_kin, E_mag (FP64 reductions) and max|div b| of each run next to the oracle's.
"""

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (ROOT, os.path.join(ROOT, "mhdflows.jl_amd"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)

import c3_setup as S  # noqa: E402
import mhdflows_amd as M  # noqa: E402
from mhdflows_amd import ic  # noqa: E402
from mhdflows_amd.devmem import hip  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(n, T):
    p = M.Problem("MI355X", nx=n, ν=S.NU, η=S.ETA, B_field=True, T=T, dt=S.DT, stepper="RK4")
    idx, vals = ic.solenoidal_modes(S.IC_SEED, k0=S.K0, kmax=S.KMAX).half_entries(n)
    M.SetUpProblemIC_modes(p, idx, vals)
    M.set_forcing(p, S.F_SEED, kf=S.KF, amp=S.AMP)
    hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    M.stepforward(p, S.STEPS)
    hip().hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    d = M.Diagnostics(p)
    log(f"c3: {n}^3 {np.dtype(T).name}: {S.STEPS} steps in {wall:.2f} s; reading the state")
    r = S.retained(p.sol, n)
    del p
    return r, d, wall


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def main(oracle, out=None):
    ref = np.load(oracle)
    R = ref["retained"]
    res = {"workload": "incompressible MHD, nu = eta = 2e-4, fixed dt = 5e-4, 50 RK4 steps, seeded "
                       "k^4 exp(-2(k/4)^2) fields |k| <= 16 + static |k| = 2 forcing; errors on the 256^3 "
                       "retained modes against the 256^3 Float64 oracle",
           "oracle": {"E_kin": float(ref["E_kin"]), "E_mag": float(ref["E_mag"]),
                      "max_div_b": float(ref["max_div_b"])}, "runs": []}
    for n, T in ((256, np.float64), (256, np.float32), (1024, np.float32)):
        log(f"c3: running {n}^3 {np.dtype(T).name}")
        r, d, wall = run(n, T)
        names = ["ux", "uy", "uz", "bx", "by", "bz"]
        res["runs"].append({
            "grid": n, "precision": np.dtype(T).name, "seconds": round(wall, 3),
            "rel_l2_all": rel(r, R), "rel_l2": {names[f]: rel(r[f], R[f]) for f in range(6)},
            "E_kin": d["E_kin"], "E_mag": d["E_mag"], "max_div_b": d["max_div_b"],
            "E_kin_rel": abs(d["E_kin"] - float(ref["E_kin"])) / float(ref["E_kin"]),
            "E_mag_rel": abs(d["E_mag"] - float(ref["E_mag"])) / float(ref["E_mag"])})
        log(f"c3: {res['runs'][-1]}")
    txt = json.dumps(res)
    print(txt)
    if out:
        with open(out, "w") as f:
            f.write(txt + "\n")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else None)
