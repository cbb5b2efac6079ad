"""Continuous-batching server for MHD solve requests on one MI355X.

Requests of different grid sizes and precisions arrive over time.  Each (grid,
precision) pair has a slot pool: one :class:`~mhdflows_amd.ensemble.Ensemble` whose runs
are slots.  Pools are created when a request needs one and the HBM budget allows it, and
released when idle and their memory is needed.  A request is admitted into a free slot
at a step boundary (parameters, seeded initial state and a step target written into the
slot's device clock); every pool then advances all its slots by a chunk of steps with
its captured step graph, slots that reached their target are skipped on the device, and
finished requests retire at the next boundary while the other slots keep running.  All
state stays in HBM; the host only reads the clocks (and the final state when asked).
"""

import collections
import dataclasses
import math
import time

import numpy as np

from . import ensemble as ens
from . import ic
from ._lib import MHDError


@dataclasses.dataclass
class Request:
    rid: int
    n: int
    T: type
    steps: int
    nu: float
    eta: float
    seed: int
    arrival: float          # seconds after the stream starts
    k0: float = 4.0
    keep_state: bool = False


@dataclasses.dataclass
class Result:
    rid: int
    n: int
    T: type
    steps: int
    arrival: float
    admitted: float
    finished: float
    sol: np.ndarray = None
    error: str = None       # set when the request was refused (never admitted)

    @property
    def latency(self):
        return self.finished - self.arrival


def initial_modes(req):
    """Seeded solenoidal k^4 exp(-2(k/k0)^2) fields, resolved by the request's grid."""
    kmax = min(16.0, math.floor(req.n / 3) - 1)
    return ic.solenoidal_modes(req.seed, k0=req.k0, kmax=kmax).half_entries(req.n)


class Pool:
    def __init__(self, n, T, slots, stepper, device):
        cfgs = [ens.make_config(n, T=T, stepper=stepper, device=device) for _ in range(slots)]
        self.n, self.T, self.slots = n, T, slots
        self.ensemble = ens.Ensemble(cfgs)
        self.bytes = slots * ens.query_bytes(cfgs[0])
        self.free = list(range(slots))
        self.active = {}  # slot -> (request, admitted time)
        for s in range(slots):  # empty slots never step
            self.ensemble.start_run(s, dt=0.0, cfl=0.0, target=0)

    def close(self):
        self.ensemble.close()


class Server:
    """Admit by HBM budget, step resident slots in chunks, retire finished requests."""

    def __init__(self, hbm_budget, slots=None, chunk=25, stepper="RK4", cfl=0.3, device=0):
        self.budget = int(hbm_budget)
        self.slots = slots or {32: 16, 64: 8, 128: 4, 256: 2}
        self.chunk = chunk
        self.stepper = stepper
        self.cfl = cfl
        self.device = device
        self.pools = {}
        self.used = 0
        self.steps_done = 0

    def _pool_bytes(self, n, T):
        """Device bytes of the slot pool for (n, T); raises MHDError for an unsupported grid."""
        return self.slots.get(n, 2) * ens.query_bytes(ens.make_config(n, T=T, stepper=self.stepper))

    def rejection(self, req):
        """Why ``req`` can never be served (unsupported configuration, or a slot pool larger
        than the whole HBM budget), or None.  Such requests are refused on arrival instead
        of waiting for memory that can never become free."""
        try:
            need = self._pool_bytes(req.n, req.T)
        except MHDError as e:
            return str(e)
        if need > self.budget:
            return (f"the {req.n}^3 {np.dtype(req.T).name} slot pool needs {need} bytes, "
                    f"more than the HBM budget of {self.budget}")
        return None

    def _pool(self, n, T):
        key = (n, np.dtype(T).name)
        p = self.pools.get(key)
        if p is not None:
            return p
        slots = self.slots.get(n, 2)
        need = self._pool_bytes(n, T)
        if self.used + need > self.budget:
            # release idle pools until the new one fits
            for k in list(self.pools):
                q = self.pools[k]
                if not q.active and self.used + need > self.budget:
                    q.close()
                    self.used -= q.bytes
                    del self.pools[k]
            if self.used + need > self.budget:
                return None
        p = Pool(n, T, slots, self.stepper, self.device)
        self.pools[key] = p
        self.used += p.bytes
        return p

    def _admit(self, req, now):
        pool = self._pool(req.n, req.T)
        if pool is None or not pool.free:
            return False
        s = pool.free.pop(0)
        e = pool.ensemble
        e.set_params(s, req.nu, req.eta)
        idx, vals = initial_modes(req)
        e.set_ic_modes(s, idx, vals)
        e.start_run(s, dt=0.0, cfl=self.cfl, target=req.steps)
        pool.active[s] = (req, now)
        return True

    def run(self, requests, realtime=True, log=None):
        """Serve ``requests`` (arrival times in seconds from now); returns {rid: Result}."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        results = {}
        t0 = time.perf_counter()
        now = 0.0
        while pending or waiting or any(p.active for p in self.pools.values()):
            now = time.perf_counter() - t0
            while pending and (pending[0].arrival <= now or not realtime):
                r = pending.popleft()
                why = self.rejection(r)
                if why is None:
                    waiting.append(r)
                else:
                    results[r.rid] = Result(r.rid, r.n, r.T, r.steps, r.arrival, now, now, None, why)
                    if log:
                        log(f"refused {r.rid}: {why}")
            # admit in arrival order; a request that cannot be placed does not block
            # requests of other sizes behind it
            blocked = set()
            for _ in range(len(waiting)):
                r = waiting.popleft()
                key = (r.n, np.dtype(r.T).name)
                if key in blocked or not self._admit(r, now):
                    blocked.add(key)
                    waiting.append(r)
            busy = [p for p in self.pools.values() if p.active]
            if not busy:
                if pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.05)))
                continue
            for p in busy:
                p.ensemble.launch(self.chunk)
            for p in busy:
                p.ensemble.sync()
            now = time.perf_counter() - t0
            for p in busy:
                for s, (req, adm) in list(p.active.items()):
                    step = p.ensemble.clock(s)[2]
                    if step >= req.steps:
                        sol = p.ensemble.get_sol(s) if req.keep_state else None
                        results[req.rid] = Result(req.rid, req.n, req.T, req.steps, req.arrival, adm, now, sol)
                        self.steps_done += req.steps
                        del p.active[s]
                        p.free.append(s)
                        if log:
                            log(f"done {req.rid} n={req.n} {np.dtype(req.T).name} steps={req.steps} "
                                f"latency={results[req.rid].latency:.3f}s")
        self.makespan = time.perf_counter() - t0
        return results

    def close(self):
        for p in self.pools.values():
            p.close()
        self.pools.clear()
        self.used = 0


def poisson_stream(nreq, rate, seed=0, steps=(100, 2000), sizes=(32, 64, 128, 256),
                   weights=(0.4, 0.3, 0.2, 0.1), fp64_frac=0.1, nu_range=(1e-4, 1e-2)):
    """Seeded request stream: Poisson arrivals at `rate` per second, grid from `sizes` with
    `weights`, steps uniform in `steps`, Float64 with probability `fp64_frac`, nu = eta
    log-uniform in `nu_range`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate)
        n = int(rng.choice(sizes, p=np.asarray(weights) / np.sum(weights)))
        T = np.float64 if rng.random() < fp64_frac else np.float32
        st = int(rng.integers(steps[0], steps[1] + 1))
        nu = float(np.exp(rng.uniform(np.log(nu_range[0]), np.log(nu_range[1]))))
        out.append(Request(i, n, T, st, nu, nu, seed * 100003 + i, t))
    return out


def summarize(results, makespan):
    done = [r for r in results.values() if r.error is None]
    lat = np.array([r.latency for r in done]) if done else np.zeros(1)
    steps = sum(r.steps for r in done)
    return {"requests": len(done), "refused": len(results) - len(done),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "aggregate_steps_per_s": steps / makespan,
            "makespan_s": makespan}
