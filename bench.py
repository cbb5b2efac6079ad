"""Benchmark: MHD RK4 steps/s of the 256^3 Float32 workload on one MI355X.

Workload (the single-GPU plain-path configuration): 3D incompressible MHD, 256^3,
Float32, RK4, nu = eta = 1e-3, dt from getCFL! with CFL 0.3 before every step,
seeded solenoidal initial fields with spectrum k^4 exp(-2 (k/4)^2), |k| <= 16,
u_rms = b_rms = 1, cross-helicity 0.  Inputs are resident in HBM before timing.

One JSON line on stdout.  `roofline` is measured with HIP events around every
launch of a separate profiled pass; `cpu_baseline` times the oracle's literal
(reference-structured, 36 FFTs per right-hand side) path on the host cores.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "mhdflows.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(M, ic, n, T, seed=2024):
    import numpy as np
    p = M.Problem("MI355X", nx=n, ν=1e-3, η=1e-3, B_field=True, T=T, stepper="RK4")
    ms = ic.solenoidal_modes(seed, k0=4.0, kmax=16.0)
    fields = ms.real_fields(n, dtype=np.float64)
    M.SetUpProblemIC(p, *fields)
    return p, fields


def profile(M, p, nsteps):
    import ctypes as C
    from mhdflows_amd import _lib as L
    arr = (L.KTime * 32)()
    n = C.c_int32()
    L.check(M.lib().mhd_profile_steps(p._h, nsteps, arr, 32, C.byref(n)))
    out = []
    for i in range(n.value):
        k = arr[i]
        out.append({"name": k.name.decode(), "launches": k.launches, "ms": k.total_ms, "bytes": k.bytes})
    return out


def cpu_baseline(fields, n, steps=1):
    import numpy as np
    from oracle import OracleProblem
    from oracle import rhs as orhs
    from oracle.integrator import time_integrator
    cores = min(16, os.cpu_count() or 1)
    orhs.WORKERS = cores
    o = OracleProblem(nx=n, nu=1e-3, eta=1e-3, B_field=True, T=np.float32, mode="literal")
    o.set_ic(*fields)
    t0 = time.perf_counter()
    time_integrator(o, math.inf, steps - 1, CFL_Coef=0.3)
    dt = time.perf_counter() - t0
    return {"value": steps / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} RK4 step(s) of the {n}^3 FP32 MHD workload through the oracle's literal "
                      f"reference-structured path (36 FFTs per right-hand side, scipy.fft workers={cores}, "
                      f"single-threaded elementwise), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--variants", type=str, default="",
                    help="fixed launch variants zinv,yinv,yfwd,zfwd,row instead of autotuning")
    ap.add_argument("--no-copy", action="store_true", help="skip the reference copy-bandwidth measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (single-GPU build; "
              "see DESIGN.md for how a run would shard)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import mhdflows_amd as M
    from mhdflows_amd import ic
    from mhdflows_amd.devmem import hip

    n = args.n
    log(f"bench: building {n}^3 FP32 MHD problem")
    p, fields = make_problem(M, ic, n, np.float32)
    cfg = p.cfg
    import ctypes as C
    step_bytes = M.lib().mhd_step_bytes_model(C.byref(cfg))

    if args.variants:
        chosen = M.variants(p, [int(v) for v in args.variants.split(",")])
    else:
        log("bench: autotuning launch variants")
        tune_report = M.autotune(p, 5)
        log(tune_report.strip())
        chosen = M.variants(p)

    log(f"bench: warmup {args.warmup} steps")
    M.TimeIntegrator(p, math.inf, max(args.warmup, 1) - 1, CFL_Coef=0.3)
    hip().hipDeviceSynchronize()
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    taken = M.TimeIntegrator(p, math.inf, args.steps - 1, CFL_Coef=0.3)
    hip().hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    assert taken == args.steps
    ms_step = wall * 1e3 / args.steps
    value = args.steps / wall
    d = M.Diagnostics(p)
    if d["has_nan"]:
        raise SystemExit("NaN in the benchmark run")

    log(f"bench: profiling {args.profile_steps} steps with per-launch HIP events")
    prof = profile(M, p, args.profile_steps)
    kern = [k for k in prof if k["bytes"] > 0]
    dom = max(kern, key=lambda k: k["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom["name"],
                "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": bytes_per_launch}
    step_gbs = step_bytes / (ms_step * 1e-3) / 1e9
    kernels = {k["name"]: {"launches": k["launches"], "avg_ms": round(k["ms"] / k["launches"], 4),
                           "GBps": round(k["bytes"] / k["launches"] / (k["ms"] / k["launches"] * 1e-3) / 1e9, 1)
                           if k["bytes"] > 0 else None} for k in prof}

    copy_gbs = None
    if not args.no_copy:
        ms = C.c_double()
        nbytes = 1 << 30
        M._lib.check(M.lib().mhd_bench_copy(nbytes, 20, C.byref(ms)))
        copy_gbs = round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)
        log(f"bench: reference copy kernel {copy_gbs} GB/s (read+write)")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, literal structure)")
        cpu = cpu_baseline(fields, n, args.cpu_steps)

    out = {
        "metric": f"MHD RK4 steps/s at {n}^3 (Float32)",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"incompressible MHD {n}^3 Float32 RK4, nu=eta=1e-3, CFL 0.3 per step, "
                               "seeded solenoidal k^4 exp(-2(k/4)^2) fields |k|<=16",
                   "grid": n, "precision": "Float32", "stepper": "RK4", "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline": {"model_bytes_per_step": step_bytes, "achieved": round(step_gbs, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "copy_kernel_GBps": copy_gbs,
        "variants": {"zpass_inv": chosen[0], "ypass_inv": chosen[1], "ypass_fwd": chosen[2],
                     "zpass_fwd": chosen[3], "xrow_fused": chosen[4]},
        "cpu_baseline": cpu,
        "zone_updates_per_s": round(value * n ** 3, 1),
        "E_kin": d["E_kin"], "E_mag": d["E_mag"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
