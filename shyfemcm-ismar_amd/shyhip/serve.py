"""Forecast serving: a request loop that keeps many runs resident on one GPU.

Each request is one storm-surge forecast: a basin, a number of steps, and
its forcing (a wind that ramps up, a surge added to the tide on the open
boundary).  The server admits requests while the estimated HBM footprint of
the resident runs stays under a budget and a slot is free; every admitted
run gets its own libshyhip model (its own HIP stream), and worker threads
step the runs concurrently, so the kernels of different requests overlap on
the device.  A run's results do not depend on what else is resident: the
models share no device memory and every kernel is deterministic.

Two servers: Server gives every run its own model and stream (runs overlap
on the device, results bit-identical alone and in the mix); BatchServer
steps all resident runs as the members of one union model (continuous
batching: one launch sequence per step for the whole batch, runs joining
and leaving between steps).

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


# the level of a boundary with no value (mkonst's flag, const.f90:119)
NO_LEVEL = -9988765.0


@dataclass
class _Run:
    req: Request
    need: int
    n: int = 0                 # steps done
    wind: tuple | None = None  # the wind stress last set on its member


class BatchServer:
    """Continuous batching: the runs on the GPU are the members of one union
    model (ensemble.union_basin), and a step advances all of them with one
    launch sequence.  Between steps the batch changes: finished runs leave
    (their fields are pulled), waiting requests join while a slot is free
    and the estimated HBM footprint stays under the budget.  A change builds
    the next union and moves every continuing run's state into it, device to
    device (shyhip_copy_member_state), so its steps go on where they
    stopped; a joining run starts from its own initial state.  Each run has
    its own wind ramp and boundary level, set per member before every step.

    A finished run's fields are pulled at once, but it stays in the union
    (its further steps are discarded) until requests join or the finished
    runs hold more than `idle_frac` of the union's element-layers: a
    rebuild costs the set-up of the union, a finished member only its share
    of each launch.

    A run's fields agree with its run alone to the solver tolerance: the
    union's zeta system is block diagonal but its convergence test is on the
    union residual."""

    def __init__(self, precision="fp32", params=None, model_kw=None, hbm_budget=64 << 30, slots=8, idle_frac=0.5,
                 solo_nodes=None, solo_slots=4):
        # runs of more than solo_nodes nodes fill the device alone: they get
        # a model and stream of their own (Server) beside the batch
        self.solo_nodes = solo_nodes
        self.solo_slots = solo_slots
        self.precision = precision
        self.params = dict(params or {})
        self.model_kw = dict(model_kw or {})
        self.budget = hbm_budget
        self.slots = slots
        self.idle_frac = idle_frac
        self.rebuilds = 0
        self.batch_steps = 0
        self.run_steps = 0
        self.build_s = 0.0          # wall time spent building unions and moving state

    def _build(self, runs):
        import shyhip
        from .ensemble import union_basin
        u = union_basin([r.req.basin for r in runs])
        tracers = None
        names = list(runs[0].req.tracers or {})
        if names:
            tracers = {n: np.concatenate([np.asarray(r.req.tracers[n], np.float64) for r in runs]) for n in names}
        m = shyhip.Model(u.basin, self.precision, params=self.params, tracers=tracers,
                         members=(u.node_start, u.elem_start), **self.model_kw)
        m.set_deferred_checks(True)     # steps queue behind each other
        for r in runs:
            r.wind = None               # a new union starts without wind
        return m, u

    def _step(self, model, union, runs, dt):
        # a run alone sets the level of its first boundary (Model.step);
        # the others keep the no-value flag they start with
        nbnd = int(union.bnd_start[-1])
        levels = np.full(nbnd, NO_LEVEL, np.float32)
        for i, r in enumerate(runs):
            # (a finished run waiting for the next rebuild keeps its last forcing)
            taux, tauy, level = forcing(r.req, min(r.n + 1, r.req.nsteps), dt)
            if r.wind != (taux, tauy):      # the wind changes while it ramps up
                model.set_member_wind(i, taux, tauy)
                r.wind = (taux, tauy)
            if union.bnd_start[i + 1] > union.bnd_start[i]:
                levels[union.bnd_start[i]] = level
        if nbnd:
            model.set_boundary_levels(levels)
        model.step(None)
        for r in runs:
            r.n += 1
        self.batch_steps += 1
        self.run_steps += sum(1 for r in runs if r.n <= r.req.nsteps)

    def serve(self, requests, keep=("znv",)):
        """Replay `requests` at their arrival times; returns them completed,
        with the stream's wall time."""
        import collections
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.t_admit, r.t_done, r.result, r.error = math.nan, math.nan, {}, None
        solo_thread = None
        if self.solo_nodes is not None:
            solo = [r for r in reqs if r.basin.nkn > self.solo_nodes]
            reqs = [r for r in reqs if r.basin.nkn <= self.solo_nodes]
            if solo:
                srv = Server(self.precision, params=self.params, model_kw=self.model_kw, hbm_budget=self.budget,
                             slots=self.solo_slots, deferred=True)
                solo_thread = threading.Thread(target=srv.serve, args=(solo, keep), daemon=True)
                solo_thread.start()
        t0 = time.perf_counter()
        waiting: collections.deque = collections.deque()
        nxt = 0
        model = union = None
        active: list[_Run] = []
        dt = None
        while True:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                waiting.append(reqs[nxt])
                nxt += 1
            # finished runs give their fields (and stay until the next rebuild)
            done = [r for r in active if r.n == r.req.nsteps and math.isnan(r.req.t_done)]
            if done:
                model.sync()
                for r in done:
                    i = active.index(r)
                    r.req.result = {k: _member_field(model, union, i, k) for k in keep}
                    r.req.t_done = time.perf_counter() - t0
            stay = [r for r in active if r.n < r.req.nsteps]
            idle = sum(r.req.elem_layers for r in active if r.n >= r.req.nsteps)
            total = sum(r.req.elem_layers for r in active)
            # waiting requests join while a slot and the budget allow
            join: list[_Run] = []
            resident = sum(r.need for r in stay)
            while waiting and len(stay) + len(join) < self.slots:
                need = footprint_bytes(waiting[0].basin, self.precision, len(waiting[0].tracers or {}))
                if (stay or join) and resident + need > self.budget:
                    break
                req = waiting.popleft()
                req.t_admit = time.perf_counter() - t0
                join.append(_Run(req, need))
                resident += need
            if join or (active and (not stay or idle > self.idle_frac * total)):
                tb = time.perf_counter()
                nxt_runs = stay + join
                new = None
                if nxt_runs:
                    try:
                        new, new_u = self._build(nxt_runs)
                        for i, r in enumerate(stay):
                            new.copy_member_state(i, model, active.index(r))
                    except Exception as e:       # the error is reported with each request
                        for r in nxt_runs:
                            r.req.error = str(e)
                            r.req.t_done = time.perf_counter() - t0
                        nxt_runs, new = [], (new.close() if new is not None else None)
                if model is not None:
                    model.close()
                model, active = new, nxt_runs
                union = new_u if new is not None else None
                self.rebuilds += 1 if new is not None else 0
                if model is not None:
                    dt = model.params_dict["dt"]
                self.build_s += time.perf_counter() - tb
            if not active:
                if nxt >= len(reqs) and not waiting:
                    break
                if not waiting:
                    time.sleep(max(0.0, reqs[nxt].arrival - (time.perf_counter() - t0)))
                continue
            try:
                self._step(model, union, active, dt)
            except Exception as e:
                # a run that stops the model stops its batch (the reference
                # stops the program); the batch's unfinished runs report it
                for r in active:
                    if math.isnan(r.req.t_done):
                        r.req.error = str(e)
                        r.req.t_done = time.perf_counter() - t0
                model.close()
                model, union, active = None, None, []
        if solo_thread is not None:
            solo_thread.join()
        return requests, time.perf_counter() - t0


def _member_field(model, union, i, name):
    a = model.pull(name)
    if a.shape[0] == union.basin.nkn:
        return a[union.node_start[i]:union.node_start[i + 1]]
    return a[union.elem_start[i]:union.elem_start[i + 1]]


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
    model as it stops the reference (check_volume).  A bay's side grows
    with the fourth root of its node count (40 x 20 km at 50k nodes): the
    larger requests are finer meshes of bays of similar size, since the
    wind set-down grows with the fetch, and over a 100 km fetch it dries
    multi-layer elements (setdepth stops)."""
    from . import mesh
    rng = np.random.default_rng(np.uint64(seed))
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nkn = int(math.exp(rng.uniform(math.log(sizes[0]), math.log(sizes[1]))))
        lx = 40000.0 * (nkn / 50000.0) ** 0.25
        b = mesh.lagoon(nkn, seed=int(seed) + 17 * (i + 1), lx=lx, ly=lx / 2, nlv=nlv, tile=32,
                        median=median_depth, dz=dz)
        out.append(Request(rid=i, basin=b, nsteps=int(rng.integers(steps[0], steps[1] + 1)), arrival=t,
                           wind_max=30.0, wind_dir=float(rng.uniform(0, 2 * math.pi)),
                           surge=float(rng.uniform(0.3, 1.2)), tide_phase=float(rng.uniform(0, 2 * math.pi))))
    return out
