"""Benchmark: MHD RK4 steps/s of the 256^3 Float32 workload on one MI355X.

Workload (the single-GPU plain-path configuration): 3D incompressible MHD, 256^3,
Float32, RK4, nu = eta = 1e-3, dt from getCFL! with CFL 0.3 before every step,
seeded solenoidal initial fields with spectrum k^4 exp(-2 (k/4)^2), |k| <= 16,
u_rms = b_rms = 1, cross-helicity 0.  Inputs are resident in HBM before timing.

One JSON line on stdout.  `value` is the rate of one call of --steps steps with the
library's default dealiased steps (every step of a call but the last updates the kept modes
only; the final state is the same bit for bit).  `roofline` is the dominant launch:
algorithmic bytes per launch over its average duration, both measured with HIP events
around every launch of a separate profiled call, and its HBM traffic from the committed
rocprofv3 counter summary.  `step_roofline` gives the whole step's bytes per step time:
the executed launches' algorithmic bytes, the counters' HBM bytes (frac = HBM rate / 8 TB/s),
for the timed call and for a call of full steps; the fixed byte model of the reference
algorithm is reported only as an equivalent rate.  `cpu_baseline` times the CPU
restatement of the reference path (36 FFTs per right-hand side) on the host cores.
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
VARIANT_NAMES = ("zpass_inv", "ypass_inv", "ypass_fwd", "zpass_fwd", "xrow_fused")
COUNTER_PROFILE = os.path.join("profiles", "mhd256_fp32_rk4_launches.md")


def counter_blob(profile):
    """The machine-readable block of a committed rocprofv3 counter summary, or None."""
    try:
        txt = open(os.path.join(ROOT, profile)).read()
        blob = json.loads(txt.split("```json", 1)[1].split("```", 1)[0])
    except (OSError, IndexError, ValueError):
        return None
    if not all("dealiased" in r and "full" in r for r in blob.get("launches", {}).values()):
        return None
    blob["file"] = profile
    return blob


def counter_step_dram(blob, variants, steps_per_call):
    """HBM bytes per RK4 step (4 right-hand sides of the six launches) of a call of
    `steps_per_call` steps, whose last step is full and the others dealiased, from the counter
    summary's per-launch FETCH_SIZE/WRITE_SIZE bytes of both kinds of step; and its source."""
    if blob is None:
        return None, None
    d = sum(rec["dealiased"]["traffic_bytes"] for rec in blob["launches"].values())
    f = sum(rec["full"]["traffic_bytes"] for rec in blob["launches"].values())
    s = max(int(steps_per_call), 1)
    same = "the same launch variants" if blob.get("variants") == variants else \
        f"launch variants {blob.get('variants')} (this run chose others)"
    src = (f"{blob['file']}: 4 x the six launches' calibrated FETCH_SIZE + WRITE_SIZE bytes, "
           f"{s - 1} dealiased + 1 full step per call, taken with {same}")
    return {"dealiased": round(4 * d), "full": round(4 * f), "call": round(4 * ((s - 1) * d + f) / s)}, src


def counter_launch_traffic(blob, launch, variants, steps_per_call):
    """HBM bytes per launch of `launch` in a call of `steps_per_call` steps (the same mix of
    dealiased and full steps), or None when the summary lacks it or used another variant."""
    if blob is None:
        return None, None
    rec = blob.get("launches", {}).get(launch)
    if rec is None or blob.get("variants", {}).get(launch) != variants.get(launch):
        return None, None
    s = max(int(steps_per_call), 1)
    t = ((s - 1) * rec["dealiased"]["traffic_bytes"] + rec["full"]["traffic_bytes"]) / s
    cal = blob["calibration"]
    src = (f"{blob['file']}: FETCH_SIZE x{cal['fetch_factor']} + WRITE_SIZE x{cal['write_factor']} per "
           f"dispatch of {rec['kernel']} (copy-kernel calibrated), {s - 1} dealiased + 1 full step")
    return round(t), src


def step_roofline(model_bytes, steps_per_call, ms_call, exec_call, ms_full, exec_full, dram, dram_src):
    """Whole-step rates: the bytes the executed launches move (the library's per-launch byte
    model of what each launch reads and writes) and the HBM bytes the counters saw, per step
    time; `frac` is the HBM rate of the timed call against the 8 TB/s peak.  The fixed
    three-pass model of the reference algorithm is kept as a labelled equivalent rate."""
    def leg(ms, alg, hbm):
        out = {"ms_per_step": round(ms, 4), "executed_alg_bytes_per_step": round(alg),
               "executed_alg_GBps": round(alg / (ms * 1e-3) / 1e9, 1)}
        if hbm is not None:
            out["dram_bytes_per_step"] = hbm
            out["dram_GBps"] = round(hbm / (ms * 1e-3) / 1e9, 1)
            out["frac"] = round(hbm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        else:
            out["frac"] = round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        return out
    call = leg(ms_call, exec_call, dram["call"] if dram else None)
    call["steps_per_call"] = int(steps_per_call)
    out = {"peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": call["frac"],
           "frac_basis": "HBM bytes per step (counters)" if dram else "executed algorithmic bytes per step",
           "timed_call": call}
    if ms_full is not None:
        out["every_step_full"] = leg(ms_full, exec_full, dram["full"] if dram else None)
    out["dram_source"] = dram_src
    out["model_equivalent"] = {
        "model_bytes_per_step": model_bytes,
        "GBps_timed_call": round(model_bytes / (ms_call * 1e-3) / 1e9, 1),
        "GBps_every_step_full": round(model_bytes / (ms_full * 1e-3) / 1e9, 1) if ms_full else None,
        "note": "fixed three-pass byte model of the reference algorithm (MHD RK4: 438 V_c + 120 V_r) "
                "over the step time: a rate comparable across implementations, not a fraction of peak"}
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(M, ic, n, T, seed=2024):
    """The bench problem; grids above 256^3 get their (|k| <= 16) initial modes set on the
    device instead of through host real fields of n^3 points."""
    import numpy as np
    p = M.Problem("MI355X", nx=n, ν=1e-3, η=1e-3, B_field=True, T=T, stepper="RK4")
    ms = ic.solenoidal_modes(seed, k0=4.0, kmax=16.0)
    if n > 256:
        idx, vals = ms.half_entries(n)
        M.SetUpProblemIC_modes(p, idx, vals)
        return p, None
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


def executed_bytes_per_step(prof, nsteps):
    """Algorithmic bytes of the launches a profiled run executed, per step (dealiased steps
    move fewer bytes than the last, full step of a call)."""
    return sum(k["bytes"] for k in prof) / nsteps


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _ranges(cpus):
    cpus, out, i = sorted(cpus), [], 0
    while i < len(cpus):
        j = i
        while j + 1 < len(cpus) and cpus[j + 1] == cpus[j] + 1:
            j += 1
        out.append(f"{cpus[i]}-{cpus[j]}" if j > i else str(cpus[i]))
        i = j + 1
    return ",".join(out)


def cpu_baseline(sizes=(32, 64, 128, 256), timed=3):
    """The CPU restatement of the reference path (the oracle's literal mode: the reference's
    operation structure, 36 real 3D FFTs per right-hand side, pocketfft with one worker per
    core, single-threaded elementwise work), timed on this host's cores.  32^3 runs the 3D
    Taylor-Green problem (nu = eta = 1e-2, dt = 1e-3, 100 steps); the larger grids run the
    bench workload (nu = eta = 1e-3, CFL 0.3) for one warm-up and `timed` timed steps.
    `value` is the 256^3 rate, the GPU workload's."""
    import numpy as np
    from mhdflows_amd import ic
    from oracle import OracleProblem
    from oracle import rhs as orhs
    from oracle.integrator import run_steps, time_integrator
    try:
        aff = sorted(os.sched_getaffinity(0))
    except AttributeError:
        aff = list(range(os.cpu_count() or 1))
    cores = max(1, min(16, len(aff)))   # the host share of one GPU is 16 cores
    orhs.WORKERS = cores
    readme = {32: 20.4, 64: 5.56, 128: 0.671, 256: 0.0690}   # README.md:77, Ryzen 7 5800X, 8 threads
    rows = {}
    for n in sizes:
        if n == 32:
            u, b = ic.taylor_green(32, dtype=np.float64)
            o = OracleProblem(nx=32, nu=1e-2, eta=1e-2, B_field=True, T=np.float32, dt=1e-3, mode="literal")
            o.set_ic(*u, *b)
            t0 = time.perf_counter()
            run_steps(o, 100)
            dt, nst = time.perf_counter() - t0, 100
            what = "Taylor-Green, dt 1e-3, 100 steps"
        else:
            fields = ic.solenoidal_modes(2024, k0=4.0, kmax=16.0).real_fields(n, dtype=np.float64)
            o = OracleProblem(nx=n, nu=1e-3, eta=1e-3, B_field=True, T=np.float32, mode="literal")
            o.set_ic(*fields)
            del fields
            time_integrator(o, math.inf, 0, CFL_Coef=0.3)
            t0 = time.perf_counter()
            time_integrator(o, math.inf, timed - 1, CFL_Coef=0.3)
            dt, nst = time.perf_counter() - t0, timed
            what = f"bench workload, CFL 0.3, 1 warm-up + {timed} timed steps"
        rows[str(n)] = {"steps_per_s": round(nst / dt, 4), "zone_updates_per_s": round(nst * n ** 3 / dt, 1),
                        "seconds": round(dt, 2), "run": what, "readme_ryzen_5800x_steps_per_s": readme.get(n)}
        log(f"bench: CPU {n}^3 {rows[str(n)]}")
        del o
    top = rows[str(max(sizes))]
    return {"value": top["steps_per_s"], "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{max(sizes)}^3 FP32 MHD RK4, {top['run']}, through the CPU restatement (pocketfft) of the "
                      f"reference path in its literal operation structure (36 FFTs per right-hand side), "
                      f"scipy.fft workers={cores}, elementwise single-threaded",
            "label": "CPU restatement (pocketfft), not the Julia/FFTW reference",
            "cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(), "affinity": _ranges(aff),
            "sizes": rows}


def large_bench(M, n=1024, steps=10, warmup=2, variants=None, full_steps=False, retime_full=True):
    """The 1024^3 Float32 MHD RK4 step (nu = eta = 2e-4, CFL 0.3, seeded |k| <= 16 fields
    set as sparse modes on the device): steps/s of one call of `steps` steps and the step
    roofline of that call and of a call of full steps."""
    import ctypes as C
    import numpy as np
    from mhdflows_amd import ic
    from mhdflows_amd.devmem import hip
    try:
        p = M.Problem("MI355X", nx=n, ν=2e-4, η=2e-4, B_field=True, T=np.float32, stepper="RK4")
    except M.MHDError as e:
        return {"grid": n, "error": str(e)}
    idx, vals = ic.solenoidal_modes(2024, k0=4.0, kmax=16.0).half_entries(n)
    M.SetUpProblemIC_modes(p, idx, vals)
    if variants:
        M.variants(p, variants)
    else:
        M.autotune(p, 2)
    if full_steps:
        M.set_step_pruning(p, False)
    M.TimeIntegrator(p, math.inf, warmup - 1, CFL_Coef=0.3)
    hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    M.TimeIntegrator(p, math.inf, steps - 1, CFL_Coef=0.3)
    hip().hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / steps
    model = M.lib().mhd_step_bytes_model(C.byref(p.cfg))
    prof = profile(M, p, steps)
    exec_call = executed_bytes_per_step(prof, steps)
    d = M.Diagnostics(p)
    ms_full = exec_full = None
    if retime_full and not full_steps:
        M.set_step_pruning(p, False)
        hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        M.TimeIntegrator(p, math.inf, steps - 1, CFL_Coef=0.3)
        hip().hipDeviceSynchronize()
        ms_full = (time.perf_counter() - t0) * 1e3 / steps
        exec_full = executed_bytes_per_step(profile(M, p, steps), steps)
        M.set_step_pruning(p, True)
    vn = M.variants(p)
    named = dict(zip(VARIANT_NAMES, vn))
    blob = counter_blob(os.path.join("profiles", "mhd1024_fp32_rk4_launches.md"))
    dram, src = counter_step_dram(blob, named, 1 if full_steps else steps)
    if full_steps and dram:
        dram["call"] = dram["full"]
    out = {"grid": n, "workload": "incompressible MHD 1024^3 Float32 RK4, nu=eta=2e-4, CFL 0.3 per step, "
                                  "seeded solenoidal k^4 exp(-2(k/4)^2) fields |k|<=16 set as device modes",
           "steps": steps, "value": round(steps / wall, 4), "unit": "steps/s", "ms_per_step": round(ms, 2),
           "every_step_full": full_steps, "variants": named,
           "step_roofline": step_roofline(model, steps, ms, exec_call, ms_full, exec_full, dram, src),
           "kernels": {k["name"]: {"launches": k["launches"], "avg_ms": round(k["ms"] / k["launches"], 4),
                                   "GBps": round(k["bytes"] / k["launches"] / (k["ms"] / k["launches"] * 1e-3) / 1e9, 1)
                                   if k["bytes"] > 0 else None} for k in prof},
           "finite": not d["has_nan"]}
    del p
    return out


def ensemble_bench(M, steps=20, only_mode=None, classes=None):
    """64 runs (16 each of 32^3, 48^3, 64^3, 96^3), Float32, LSRK54, CFL 0.3, seeded
    k^4 exp(-2(k/k0)^2) fields with k0 ~ U[2,6], nu = eta log-uniform in [1e-3, 1e-2]:
    aggregate run-steps/s of the ensemble (each size group by its own launches on its own
    stream; then the multi-size modes: all four groups by multi-size launches, and the
    largest group alone beside the other three) vs the same runs stepped one after another."""
    import numpy as np
    from mhdflows_amd import ensemble as E
    from mhdflows_amd import ic
    rng = np.random.Generator(np.random.PCG64(4))
    specs = []
    for n in (32, 48, 64, 96):
        for _ in range(16):
            specs.append((n, float(rng.uniform(2, 6)), float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-2)))),
                          int(rng.integers(1 << 30))))
    modes = [ic.solenoidal_modes(s, k0=k0, kmax=min(16.0, n // 3 - 1)).half_entries(n) for (n, k0, nu, s) in specs]
    cfgs = [E.make_config(n, nu=nu, eta=nu, T=np.float32, stepper="LSRK54") for (n, k0, nu, s) in specs]
    ens = E.Ensemble(cfgs)
    for r, (idx, vals) in enumerate(modes):
        ens.set_ic_modes(r, idx, vals)
        ens.set_cfl(r, 0.3)
    rates, nmulti, chosen = {}, {}, {}
    for mode in ((0, 1, 2) if only_mode is None else (only_mode,)):
        # group streams (default), multi-size launches, largest group alone
        nmulti[mode] = ens.multi(mode)
        ens.autotune(3)
        if classes is not None:   # fixed per-class choice instead of the tuned one
            ens.multi_classes(*classes)
        chosen[mode] = ens.multi_classes()
        ens.step(2)
        ens.sync()
        rates[mode] = len(specs) * steps / (ens.time_steps(steps) * 1e-3)
    ens.close()
    if only_mode is not None:   # one mode alone (profiling runs): no sequential baseline
        return {"workload": "64 runs: 16 each of 32^3, 48^3, 64^3, 96^3, Float32 LSRK54, CFL 0.3",
                "steps": steps, "multi_mode": only_mode, "groups_in_multi_launches": nmulti[only_mode],
                "classes": chosen[only_mode], "run_steps_per_s": round(rates[only_mode], 1)}
    ens_rate = rates[0]
    # sequential plain path: each run alone (own handle, captured step), one after another
    from mhdflows_amd.devmem import hip
    total = 0.0
    for (n, k0, nu, s), (idx, vals) in zip(specs, modes):
        p = M.Problem("MI355X", nx=n, ν=nu, η=nu, B_field=True, T=np.float32, stepper="LSRK54")
        M.SetUpProblemIC_modes(p, idx, vals)
        M.TimeIntegrator(p, math.inf, 1, CFL_Coef=0.3)
        hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        M.TimeIntegrator(p, math.inf, steps - 1, CFL_Coef=0.3)
        hip().hipDeviceSynchronize()
        total += time.perf_counter() - t0
        del p
    seq_rate = len(specs) * steps / total
    return {"workload": "64 runs: 16 each of 32^3, 48^3, 64^3, 96^3, Float32 LSRK54, CFL 0.3",
            "steps": steps, "ensemble_run_steps_per_s": round(ens_rate, 1),
            "mode": "each size group on its own stream (default)",
            "multi_size": {
                "all_groups_run_steps_per_s": round(rates[1], 1), "groups": nmulti[1],
                "largest_group_alone_run_steps_per_s": round(rates[2], 1), "groups_in_multi_launches": nmulti[2],
                "tuned_classes": {str(m): {"column_passes": "multi" if chosen[m][0] == 1 else "per group",
                                           "row_kernel": "multi" if chosen[m][1] == 1 else "per group",
                                           "spectral": "multi"} for m in (1, 2)}},
            "sequential_run_steps_per_s": round(seq_rate, 1), "speedup": round(ens_rate / seq_rate, 2)}


def ensemble64_bench(M, nruns=64, steps=20, seq_runs=16):
    """64 runs of 64^3 (Float32, LSRK54, CFL 0.3, k0 ~ U[2,6], nu = eta log-uniform in
    [1e-3, 1e-2]) stepped as one ensemble, against the same runs stepped one after another
    (a) each on its own captured step with the device-side CFL and (b) the reference's
    loop, getCFL! read back on the host before every stepforward!.  The sequential rates
    come from the first `seq_runs` runs (same size, so the same cost per run)."""
    import numpy as np
    from mhdflows_amd import ensemble as E
    from mhdflows_amd import ic
    from mhdflows_amd.devmem import hip
    rng = np.random.Generator(np.random.PCG64(11))
    specs = [(float(rng.uniform(2, 6)), float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-2)))),
              int(rng.integers(1 << 30))) for _ in range(nruns)]
    modes = [ic.solenoidal_modes(s, k0=k0, kmax=16.0).half_entries(64) for (k0, nu, s) in specs]
    ens = E.Ensemble([E.make_config(64, nu=nu, eta=nu, T=np.float32, stepper="LSRK54") for (k0, nu, s) in specs])
    for r, (idx, vals) in enumerate(modes):
        ens.set_ic_modes(r, idx, vals)
        ens.set_cfl(r, 0.3)
    ens.autotune(3)
    ens.step(2)
    ens.sync()
    ms = ens.time_steps(steps)
    ens_rate = nruns * steps / (ms * 1e-3)
    ens.close()
    t_graph = t_host = 0.0
    for (k0, nu, s), (idx, vals) in list(zip(specs, modes))[:seq_runs]:
        p = M.Problem("MI355X", nx=64, ν=nu, η=nu, B_field=True, T=np.float32, stepper="LSRK54")
        M.SetUpProblemIC_modes(p, idx, vals)
        M.TimeIntegrator(p, math.inf, 1, CFL_Coef=0.3)
        hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        M.TimeIntegrator(p, math.inf, steps - 1, CFL_Coef=0.3)
        hip().hipDeviceSynchronize()
        t_graph += time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(steps):
            M.getCFL(p, Coef=0.3)
            M.stepforward(p, 1)
        hip().hipDeviceSynchronize()
        t_host += time.perf_counter() - t0
        del p
    seq_graph = seq_runs * steps / t_graph
    seq_host = seq_runs * steps / t_host
    return {"workload": "64 runs of 64^3 MHD, Float32 LSRK54, CFL 0.3", "steps": steps,
            "ensemble_run_steps_per_s": round(ens_rate, 1),
            "sequential_captured_run_steps_per_s": round(seq_graph, 1),
            "sequential_host_cfl_run_steps_per_s": round(seq_host, 1),
            "speedup_vs_captured": round(ens_rate / seq_graph, 2),
            "speedup_vs_host_cfl": round(ens_rate / seq_host, 2)}


def served64_bench(M, nreq=64, steps=(100, 300), seed=5, seq_runs=8):
    """64 requests of 64^3 Float32 MHD (RK4, CFL 0.3, nu = eta log-uniform in [1e-4, 1e-2],
    100-300 steps) arriving together and served by continuous batching in one slot pool,
    against the same requests solved one after another on the single-run captured path
    (rate from the first `seq_runs` requests)."""
    import numpy as np
    from mhdflows_amd import serve
    from mhdflows_amd.devmem import hip
    reqs = serve.poisson_stream(nreq, rate=1e6, seed=seed, steps=steps, sizes=(64,), weights=(1.0,),
                                fp64_frac=0.0)
    srv = serve.Server(hbm_budget=200 << 30, slots={64: nreq}, chunk=25)
    res = srv.run(reqs, realtime=False)
    out = serve.summarize(res, srv.makespan)
    srv.close()
    t = 0.0
    nsteps = 0
    for r in reqs[:seq_runs]:
        p = M.Problem("MI355X", nx=64, ν=r.nu, η=r.eta, B_field=True, T=np.float32, stepper="RK4")
        idx, vals = serve.initial_modes(r)
        M.SetUpProblemIC_modes(p, idx, vals)
        hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        M.TimeIntegrator(p, math.inf, r.steps - 1, CFL_Coef=0.3)
        hip().hipDeviceSynchronize()
        t += time.perf_counter() - t0
        nsteps += r.steps
        del p
    seq = nsteps / t
    return {"workload": f"{nreq} requests of 64^3 MHD Float32 RK4, {steps[0]}-{steps[1]} steps, arriving together",
            "served_steps_per_s": round(out["aggregate_steps_per_s"], 1),
            "p50_latency_s": round(out["p50_latency_s"], 4), "p99_latency_s": round(out["p99_latency_s"], 4),
            "sequential_steps_per_s": round(seq, 1), "speedup": round(out["aggregate_steps_per_s"] / seq, 2)}


def serving_bench(nreq, rate, steps, seed=7):
    """Poisson stream of MHD solve requests (grid 32/64/128/256 with weights .4/.3/.2/.1,
    Float32/Float64 90/10, nu = eta log-uniform in [1e-4, 1e-2], seeded k^4 exp(-2(k/4)^2)
    fields, CFL 0.3), admitted by an HBM budget into slot pools."""
    import numpy as np
    from mhdflows_amd import serve
    reqs = serve.poisson_stream(nreq, rate=rate, seed=seed, steps=steps)
    srv = serve.Server(hbm_budget=200 << 30, chunk=25)
    res = srv.run(reqs, realtime=True)
    out = serve.summarize(res, srv.makespan)
    srv.close()
    out.update({"arrival_rate_per_s": rate, "steps_range": list(steps),
                "fp64_requests": int(sum(1 for r in reqs if r.T == np.float64)),
                "grids": {str(n): int(sum(1 for r in reqs if r.n == n)) for n in (32, 64, 128, 256)}})
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in out.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-full", action="store_true", help="skip the every-step-full timing")
    ap.add_argument("--full-steps", action="store_true",
                    help="compute every step in full for the whole run (counter runs of full steps)")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps of the profiled call (default: --steps, the timed call's mix of steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sizes", type=str, default="32,64,128,256")
    ap.add_argument("--variants", type=str, default="",
                    help="fixed launch variants zinv,yinv,yfwd,zfwd,row instead of autotuning")
    ap.add_argument("--no-copy", action="store_true", help="skip the reference copy-bandwidth measurement")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-run ensemble measurement")
    ap.add_argument("--no-large", action="store_true", help="skip the 1024^3 measurement")
    ap.add_argument("--large-only", action="store_true",
                    help="run the 1024^3 leg alone (its own JSON line); --steps/--warmup apply to it")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-stream measurement")
    ap.add_argument("--ensemble-only", type=int, default=None, metavar="MODE",
                    help="run the 64-run uneven ensemble alone in one multi-size mode (0, 1, 2)")
    ap.add_argument("--ensemble-classes", type=str, default="",
                    help="with --ensemble-only: fixed multi-size classes col,row (1 multi, 0 per group)")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-steps", type=str, default="100,2000")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (single-GPU build; "
              "a single large run would shard as slabs in z with one all-to-all per 3D transform, "
              "ensembles by placing runs on GPUs)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    import numpy as np
    import mhdflows_amd as M
    from mhdflows_amd import ic
    from mhdflows_amd.devmem import hip

    if args.ensemble_only is not None:
        cls = tuple(bool(int(v)) for v in args.ensemble_classes.split(",")) if args.ensemble_classes else None
        log(f"bench: 64-run uneven ensemble, multi-size mode {args.ensemble_only}")
        print(json.dumps(ensemble_bench(M, only_mode=args.ensemble_only, classes=cls)), flush=True)
        return

    fixed = [int(v) for v in args.variants.split(",")] if args.variants else None
    if args.large_only:
        log("bench: 1024^3 Float32 MHD RK4 alone")
        out = large_bench(M, steps=args.steps, warmup=args.warmup, variants=fixed, full_steps=args.full_steps,
                          retime_full=not args.no_full)
        print(json.dumps(out), flush=True)
        return

    n = args.n
    log(f"bench: building {n}^3 FP32 MHD problem")
    p, fields = make_problem(M, ic, n, np.float32)
    del fields
    step_bytes = M.lib().mhd_step_bytes_model(C.byref(p.cfg))

    if fixed:
        chosen = M.variants(p, fixed)
    else:
        log("bench: autotuning launch variants")
        tune_report = M.autotune(p, 5)
        log(tune_report.strip())
        chosen = M.variants(p)
    chosen_named = dict(zip(VARIANT_NAMES, chosen))
    if args.full_steps:
        M.set_step_pruning(p, False)

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

    nprof = args.profile_steps or args.steps
    log(f"bench: profiling {nprof} steps with per-launch HIP events")
    prof = profile(M, p, nprof)
    exec_call = executed_bytes_per_step(prof, nprof)
    ms_full = exec_full = None
    if not args.no_full and not args.full_steps:
        M.set_step_pruning(p, False)
        hip().hipDeviceSynchronize()
        t0 = time.perf_counter()
        M.TimeIntegrator(p, math.inf, args.steps - 1, CFL_Coef=0.3)
        hip().hipDeviceSynchronize()
        ms_full = (time.perf_counter() - t0) * 1e3 / args.steps
        exec_full = executed_bytes_per_step(profile(M, p, nprof), nprof)
        M.set_step_pruning(p, True)
        log(f"bench: every step full: {1e3 / ms_full:.2f} steps/s")
    kern = [k for k in prof if k["bytes"] > 0]
    dom = max(kern, key=lambda k: k["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    blob = counter_blob(COUNTER_PROFILE)
    mix = 1 if args.full_steps else nprof   # a call of full steps has the counters' full-step bytes
    traffic, traffic_src = counter_launch_traffic(blob, dom["name"], chosen_named, mix)
    if args.full_steps and blob and dom["name"] in blob.get("launches", {}):
        traffic = blob["launches"][dom["name"]]["full"]["traffic_bytes"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom["name"],
                "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": round(bytes_per_launch),
                "launches_timed": dom["launches"], "traffic_source": traffic_src}
    dram, dram_src = counter_step_dram(blob, chosen_named, 1 if args.full_steps else args.steps)
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
    del p

    large = None
    if not args.no_large:
        log("bench: 1024^3 Float32 MHD RK4")
        large = large_bench(M)
        log(f"bench: 1024^3 {large}")

    ensemble = None
    if not args.no_ensemble:
        log("bench: ensemble of 64 uneven runs vs sequential plain runs")
        ensemble = ensemble_bench(M)
        log(f"bench: ensemble {ensemble}")
        log("bench: 64 runs of 64^3 as one ensemble vs sequential runs")
        ensemble["n64"] = ensemble64_bench(M)
        log(f"bench: ensemble 64^3 {ensemble['n64']}")
    served = None
    if not args.no_serve:
        log(f"bench: serving stream of {args.serve_requests} requests")
        lo, hi = (int(v) for v in args.serve_steps.split(","))
        served = serving_bench(args.serve_requests, args.serve_rate, (lo, hi))
        log(f"bench: serving {served}")
        log("bench: 64 requests of 64^3 served together vs sequential solves")
        served["n64"] = served64_bench(M)
        log(f"bench: served 64^3 {served['n64']}")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (CPU restatement of the reference path, literal structure)")
        cpu = cpu_baseline(tuple(int(v) for v in args.cpu_sizes.split(",")))

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
                               "seeded solenoidal k^4 exp(-2(k/4)^2) fields |k|<=16"
                               + (", every step full" if args.full_steps else ""),
                   "grid": n, "precision": "Float32", "stepper": "RK4", "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline": step_roofline(step_bytes, args.steps, ms_step, exec_call, ms_full, exec_full, dram,
                                       dram_src),
        "kernels": kernels,
        "copy_kernel_GBps": copy_gbs,
        "variants": chosen_named,
        "cpu_baseline": cpu,
        "n1024": large,
        "ensemble": ensemble,
        "serving": served,
        "zone_updates_per_s": round(value * n ** 3, 1),
        "E_kin": d["E_kin"], "E_mag": d["E_mag"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
