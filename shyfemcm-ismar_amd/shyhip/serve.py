"""Forecast serving: a request loop that keeps many runs resident on one GPU.

Each request is one storm-surge forecast: a basin, a number of steps, and
its forcing (a wind that ramps up, a surge added to the tide on the open
boundary).  The server admits requests while the estimated HBM footprint of
the resident runs stays under a budget and a slot is free; every admitted
run gets its own libshyhip model (its own HIP stream), and worker threads
step the runs concurrently, so the kernels of different requests overlap on
the device.  A run's results do not depend on what else is resident: the
models share no device memory and every kernel is deterministic.

Latency of a request = completion time - arrival time (queueing, set-up and
all its steps); throughput = sum over requests of elements x layers x steps
divided by the wall time of the whole stream.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    basin: object
    nsteps: int
    arrival: float                     # seconds after the stream starts
    wind_max: float = 30.0             # m/s, reached after ramp_steps
    wind_dir: float = 0.0
    ramp_steps: int = 60
    surge: float = 0.5                 # m added to the tide on the open boundary
    tide_phase: float = 0.0
    tracers: dict | None = None
    # filled by the server
    t_admit: float = math.nan
    t_done: float = math.nan
    result: dict = field(default_factory=dict)
    error: str | None = None

    @property
    def elem_layers(self) -> int:
        return int(np.asarray(self.basin.ilhv).sum())


def footprint_bytes(basin, precision: str, ntracer: int = 0) -> int:
    """Estimated device memory of one resident run: about 30 element-layer
    and 20 node-layer state arrays, the zeta matrix with its block factors
    (CSR and fixed-stride copies), its ELL slices and 12 Krylov vectors, plus
    the connectivity and geometry tables."""
    p = 8 if precision == "fp64" else 4
    EL = basin.nel * basin.nlv
    NL = basin.nkn * (basin.nlv + 1)
    nnz = 7 * basin.nkn
    return int(30 * EL * p + (20 + 2 * ntracer) * NL * p + nnz * (8 + 4 + 3 * p + 24) + 12 * basin.nkn * 8
               + 10 * basin.nkn * (8 + p) + basin.nel * (12 + 56 + 6 * 12 + 18 * 12))


def forcing(req: Request, n: int, dt: float):
    """Wind stress and boundary level of step n (1-based) of a request."""
    from . import mesh
    ramp = min(1.0, n / max(1, req.ramp_steps))
    taux, tauy = mesh.wind_stress(req.wind_max * ramp, req.wind_dir)
    level = mesh.tide_level(n * dt, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD),
                            phases=(req.tide_phase, 0.7)) + req.surge * ramp
    return taux, tauy, level


def run_request(req: Request, precision: str, params: dict, model_kw: dict, keep=("znv",),
                deferred: bool = False) -> dict:
    """Create the model of `req`, advance it nsteps, return the kept fields
    (deferred: the steps do not wait for the device; errors surface at the
    final pull)."""
    import shyhip
    m = shyhip.Model(req.basin, precision, params=params, tracers=req.tracers, **model_kw)
    try:
        if deferred:
            m.set_deferred_checks(True)
        dt = m.params_dict["dt"]
        for n in range(1, req.nsteps + 1):
            taux, tauy, level = forcing(req, n, dt)
            m.set_wind(taux, tauy)
            m.step(level)
        m.sync()
        return {k: m.pull(k) for k in keep}
    finally:
        m.close()


class Server:
    """Admission by HBM footprint and slots; worker threads step the runs."""

    def __init__(self, precision="fp32", params=None, model_kw=None, hbm_budget=64 << 30, slots=8, deferred=False):
        self.precision = precision
        self.deferred = deferred
        self.params = dict(params or {})
        self.model_kw = dict(model_kw or {})
        self.budget = hbm_budget
        self.slots = slots
        self._lock = threading.Condition()
        self._resident = 0
        self._active = 0

    def _admit(self, need):
        with self._lock:
            while self._active >= self.slots or (self._active > 0 and self._resident + need > self.budget):
                self._lock.wait()
            self._active += 1
            self._resident += need

    def _release(self, need):
        with self._lock:
            self._active -= 1
            self._resident -= need
            self._lock.notify_all()

    def serve(self, requests, keep=("znv",)):
        """Replay `requests` at their arrival times; returns them completed,
        with the stream's wall time."""
        pending: queue.Queue = queue.Queue()
        t0 = time.perf_counter()

        def worker():
            while True:
                req = pending.get()
                if req is None:
                    return
                need = footprint_bytes(req.basin, self.precision, len(req.tracers or {}))
                self._admit(need)
                req.t_admit = time.perf_counter() - t0
                try:
                    req.result = run_request(req, self.precision, self.params, self.model_kw, keep, self.deferred)
                except Exception as e:   # the error is reported with the request
                    req.error = str(e)
                finally:
                    req.t_done = time.perf_counter() - t0
                    self._release(need)

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.slots)]
        for th in threads:
            th.start()
        for req in sorted(requests, key=lambda r: r.arrival):
            delay = req.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            pending.put(req)
        for _ in threads:
            pending.put(None)
        for th in threads:
            th.join()
        return requests, time.perf_counter() - t0


def summary(requests, wall: float) -> dict:
    done = [r for r in requests if r.error is None]
    lat = np.array([r.t_done - r.arrival for r in done])
    work = sum(r.elem_layers * r.nsteps for r in done)
    return dict(requests=len(requests), failed=len(requests) - len(done),
                p50_s=float(np.percentile(lat, 50)) if len(lat) else None,
                p99_s=float(np.percentile(lat, 99)) if len(lat) else None,
                throughput=work / wall if wall > 0 else None, wall_s=wall)


def poisson_stream(n, rate, seed, sizes=(5000, 50000), steps=(120, 360), nlv=8, median_depth=10.0, dz=3.0):
    """Synthetic request stream: Poisson arrivals (rate per second), mesh
    sizes log-uniform in `sizes`, step counts uniform in `steps`, wind
    ramping to 30 m/s at a random direction, surge 0.3-1.2 m.  The basins are
    coastal bays (log-normal depth of median `median_depth`, z-layers of dz
    metres): a 30 m/s wind over the 1.5 m lagoon of the throughput workload
    sets the surface down by more than a 1 m top layer, which stops the
    model as it stops the reference (check_volume)."""
    from . import mesh
    rng = np.random.default_rng(np.uint64(seed))
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nkn = int(math.exp(rng.uniform(math.log(sizes[0]), math.log(sizes[1]))))
        lx = 40000.0 * math.sqrt(nkn / 50000.0)
        b = mesh.lagoon(nkn, seed=int(seed) + 17 * (i + 1), lx=lx, ly=lx / 2, nlv=nlv, tile=32,
                        median=median_depth, dz=dz)
        out.append(Request(rid=i, basin=b, nsteps=int(rng.integers(steps[0], steps[1] + 1)), arrival=t,
                           wind_max=30.0, wind_dir=float(rng.uniform(0, 2 * math.pi)),
                           surge=float(rng.uniform(0.3, 1.2)), tide_phase=float(rng.uniform(0, 2 * math.pi))))
    return out
