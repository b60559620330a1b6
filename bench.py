#!/usr/bin/env python3
"""Throughput of the SHYFEM semi-implicit hydro step + T/S transport on one MI355X.

Headline workload: the 3D lagoon -- 50k nodes of a seeded jittered-point
Delaunay triangulation (~98k elements) in a 40 x 20 km basin, log-normal
depth (median 1.5 m) with intertidal flats, 12 z-layers of 1 m, M2+K1 tide
on the western boundary, uniform 8 m/s wind, T and S tracers (TVD
horizontal and vertical), dt = 60 s, 1440 steps.  Synthetic data; inputs
resident on the device before the timed region.

A "step" is one model time step: set_dry, copy_uvz/copy_depth, hydro (column
solves, zeta assembly, sparse solve, wet/dry loop, final transports, depths,
vertical velocity) and scal3sh for T and S.  Two device paths are timed:
plain FP64 (the headline value: the workload is the plain FP64 GPU run), and
FP32 state/transport with the FP32 Krylov solve inside FP64 iterative
refinement, reported beside it.  Beside them, unless switched
off: a batched ensemble of uneven lagoons (all members in one model) against
the same members run one after the other on the plain FP64 path, a serving
stream with p50/p99 latency, and the CPU oracle on a bounded sample.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shyfemcm-ismar_amd"))

HBM_PEAK_GBS = 8000.0


def _lagoon(args, hmin=0.05):
    from shyhip import mesh
    b = mesh.lagoon(args.nodes, seed=args.seed, lx=40000.0, ly=20000.0, nlv=args.layers, tile=32, hmin=hmin)
    tr = dict(temp=mesh.smooth_tracer(b, args.seed + 1, 15.0, 25.0),
              salt=mesh.smooth_tracer(b, args.seed + 2, 25.0, 38.0))
    return b, tr


def _level(t, phase=0.0):
    from shyhip import mesh
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD), phases=(phase, 0.7))


PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, itpre=2, block_rows=256, tri_sweeps=3)
# scal3sh substeps: 0 = chosen every step by scalar_stability, as the
# reference chooses them (stability.f90:220-260); on this lagoon 3-4
ISTOT = 0
WIND = (8.0, 0.9)


def _block_nnz(b, brows):
    """Entries of the zeta pattern inside the diagonal blocks of brows rows,
    and all entries."""
    e = np.concatenate([b.nen3v[:, [0, 1]], b.nen3v[:, [1, 2]], b.nen3v[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    inside = (e[:, 0] // brows) == (e[:, 1] // brows)
    return int(2 * inside.sum() + b.nkn), int(2 * len(e) + b.nkn)


def _kernel_bytes(b, p, brows, ntr=2):
    """Algorithmic HBM bytes per launch of each timed kernel: every listed
    array read or written once at its stored width p (gathers that hit
    cache are not counted).  The transport kernels carry ntr tracers per
    launch (T and S together)."""
    EL = int(b.ilhv.sum())
    NL = int(b.ilhkv.sum())
    E, N = b.nel, b.nkn
    nnzb, nnz = _block_nnz(b, brows)
    lu = nnzb * (p + 4) + 2 * N * 4             # staged block factors + row starts/diagonals
    spmv = nnz * (p + 4) + N * 4                # values, columns, row pointers
    return {
        # utlov vtlov hdeov read, utlnv vtlnv written, ddxv ddyv (2 planes) written,
        # visv at the interfaces, per element geometry/levels/summaries
        "hydro_intern": 9 * EL * p + (NL + N) * p + E * (12 + 56 + 6 * p + 4 + 48),
        # once per step: utlnv vtlnv utlov vtlov ulnv vlnv read, per element
        # nodes, list positions, geometry (7 doubles), distances (3), levels,
        # external flags and the TVD upwind tables (6 points: element, level,
        # 3 nodes, 3 weights) read; the element-layer record written (15
        # ints, 11 doubles)
        "conz3d_prep": 6 * EL * p + E * (12 + 12 + 56 + 24 + 4 + 4 + 264) + EL * 148,
        # per substep: the record read, 3 horizontal terms per tracer
        # written, each tracer gathered
        "conz3d_elem": 3 * ntr * EL * p + ntr * NL * p + EL * 148,
        # horizontal terms read; hdknv hdkov wlnv difv read, each tracer read
        # and written; per list position its bottom layer and 4 * area, per
        # node its list start and bottom layer
        "conz3d_node": 3 * ntr * EL * p + (4 + 2 * ntr) * NL * p + 2 * N * p + 3 * E * 12 + 2 * N * 4,
        "transports_final": 8 * EL * p + E * (12 + 56 + 3 * p + 2 * p),
        "hydro_vertical": 4 * EL * p + 4 * NL * p + N * p,
        "setdepth": EL * p + NL * p + E * 3 * p,
        # BiCGStab, first kernel: s t p v r0 read, dx read+written, r p v written
        "bicg_a": 10 * N * p + spmv + lu,
        # second kernel: r v r0 read, s t written
        "bicg_b": 5 * N * p + spmv + lu,
        # block ILU(0): values gathered from A, factors written, update lists read
        "bilu_factor": nnzb * (8 + p + 4 + 4 + 4) + 3 * nnzb * 4,
    }


# kernel names of the timed groups as rocprofv3 prints them (value type V)
# kernel-name prefixes (template arguments after the value type vary with
# the iteration-kernel variant) of each timed group's kernel
KERNEL_NAMES = {"bicg_a": "k_bicg_a<{v},", "bicg_b": "k_bicg_b<{v},",
                "bilu_factor": "k_bilu_factor<{v}>", "conz3d_elem": "k_conz_elem<{v}, 2, false, false>",
                "conz3d_prep": "k_conz_prep<{v}>",
                "conz3d_node": "k_conz_node<{v}, 12, 2, false>", "hydro_vertical": "k_hydro_vertical<{v}, 12>",
                "hydro_intern": "k_hydro_intern<{v}, 12, false>"}


def _measured_traffic(path, group, vtype):
    """HBM bytes per launch of a timed group's kernel from the committed
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    from shyhip import profsum
    t = profsum.read_traffic(path)
    if not t or group not in KERNEL_NAMES:
        return None
    pre = KERNEL_NAMES[group].format(v=vtype)
    hits = [v for name, v in t["kernels"].items() if name.startswith(pre)]
    return hits[0]["traffic_bytes"] if len(hits) == 1 else None


def _copy_calibration(mib):
    """One plain device-to-device copy of `mib` MiB (its counters calibrate
    FETCH_SIZE / WRITE_SIZE against a known byte count).  Made through the
    HIP runtime libshyhip itself links (the same loaded instance), never a
    second runtime in the process."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    nb = mib * 1024 * 1024
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nb)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nb)):
        raise RuntimeError("copy calibration: hipMalloc failed")
    try:
        hip.hipMemset(src, 1, C.c_size_t(nb))
        hip.hipDeviceSynchronize()
        if hip.hipMemcpy(dst, src, C.c_size_t(nb), 3):      # hipMemcpyDeviceToDevice
            raise RuntimeError("copy calibration: hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def install_forcing(m, nsteps):
    """The workload's forcing as the model reads it from its files, held on
    the device: a boundary file with the tidal level at every step (cubic
    interpolation, the default for a level boundary) and one wind record
    (a constant wind, converted to stress on the device)."""
    rec_t = PARAMS["dt"] * np.arange(-2, nsteps + 4)
    rec_v = np.array([_level(t) for t in rec_t], np.float32)
    m.set_boundary_series(1, rec_t, rec_v, intpol=4, dtime0=0.0)
    sp, d = WIND
    m.set_wind_series([0.0], [np.float32(sp * math.cos(d))], [np.float32(sp * math.sin(d))])


def run_gpu(b, tr, prec, args, steps=None, baroclinic=False, params=None):
    import shyhip
    steps = args.steps if steps is None else steps
    m = shyhip.Model(b, prec, params=params or PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    m.set_deferred_checks(True)     # device errors surface at the next wait (m.sync below)
    if baroclinic:                  # T and S drive the momentum equations (ibarcl = 1)
        m.set_baroclinic("temp", "salt")
    nt = max(5, min(20, steps))
    install_forcing(m, args.warmup + steps + nt)
    for _ in range(args.warmup):
        m.step(None, istot=ISTOT)
    m.sync()
    t0 = time.perf_counter()
    substeps = []
    for _ in range(steps):
        m.step(None, istot=ISTOT)
        substeps.append(m.last_istot())
    m.sync()
    elapsed = time.perf_counter() - t0
    zn = m.pull("znv")
    iweg = m.pull("iwegv")
    stats = m.stats()
    stats["istot_mean"] = round(float(np.mean(substeps)), 3)
    stats["istot_max"] = int(np.max(substeps))
    # kernel timing pass (events on the library's stream), outside the timed region
    m.set_timing(True)
    for _ in range(nt):
        m.step(None, istot=ISTOT)
    timing = m.timing()
    m.set_timing(False)
    m.close()
    return dict(elapsed=elapsed, timing=timing, nt=nt, znv=zn, iwegv=iweg, stats=stats)


# the settings the autotuner chooses among (all solve and transport the
# same way; they trade iterations and launch shapes)
TUNE_GRID = dict(block_rows=(128, 256, 512), tri_sweeps=(3, 4), node_tile=(64, 256))


def autotune_params(b, tr, prec, steps=20):
    """PARAMS with the block size, triangular sweeps and node tile that run
    the fastest on this basin (shyhip.autotune, timed on the workload
    itself); returns (params, tuner report)."""
    from shyhip import autotune

    def forcing(m, n):
        if n == 0:
            install_forcing(m, 4 * steps + 40)
        m.step(None, istot=ISTOT)

    rep = autotune.tune(b, prec, PARAMS, forcing, steps=steps, warmup=5, grid=TUNE_GRID, cache=False,
                        visv=1e-3, difv=1e-4, tracers=tr)
    return dict(PARAMS, **rep["choice"]), rep


def c3_run(args, params):
    """The mixed-precision configuration at 500k nodes: the FP32 + FP64-refined
    path and the plain FP64 path on the same basin and forcing; per-step
    rates and the FP32 fields' differences from the FP64 run."""
    import shyhip
    from shyhip import mesh
    scale = math.sqrt(args.c3_nodes / 50000.0)      # the same node density as the 50k lagoon
    b = mesh.lagoon(args.c3_nodes, seed=args.seed + 3, lx=40000.0 * scale, ly=20000.0 * scale,
                    nlv=args.c3_layers, tile=32)
    tr = dict(temp=mesh.smooth_tracer(b, args.seed + 4, 15.0, 25.0),
              salt=mesh.smooth_tracer(b, args.seed + 5, 25.0, 38.0))
    EL = int(b.ilhv.sum())
    out = {"workload": f"3D lagoon: {b.nkn} nodes, {b.nel} elements, {args.c3_layers} z-layers, same generator, "
                       f"depth law, tide, wind and tracers as the headline lagoon; {args.c3_steps} steps",
           "elem_layers": EL}
    fields = {}
    for prec in ("fp64", "fp32"):
        m = shyhip.Model(b, prec, params=params[prec], visv=1e-3, difv=1e-4, tracers=tr)
        m.set_deferred_checks(True)
        install_forcing(m, args.c3_steps + 10)
        for _ in range(3):
            m.step(None, istot=ISTOT)
        m.sync()
        t0 = time.perf_counter()
        for _ in range(args.c3_steps):
            m.step(None, istot=ISTOT)
        m.sync()
        el = time.perf_counter() - t0
        out[f"{prec}_ms_per_step"] = round(1e3 * el / args.c3_steps, 4)
        out[f"{prec}_value"] = round(EL * args.c3_steps / el, 1)
        out[f"{prec}_solver"] = m.stats()
        fields[prec] = {k: m.pull(k) for k in ("znv", "iwegv", "temp", "salt")}
        fields[prec]["volume"] = m.volume()
        if prec == "fp64":
            m.set_timing(True)
            for _ in range(5):
                m.step(None, istot=ISTOT)
            t = m.timing()
            m.set_timing(False)
            kb = _kernel_bytes(b, 8, params[prec]["block_rows"])
            per_step = {k: v[0] / 5 for k, v in t.items()}
            avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in t.items()}
            dom = max((k for k in kb if avg.get(k, 0.0) > 0), key=lambda k: per_step[k])
            ach = kb[dom] / (avg[dom] * 1e-3) / 1e9
            out["roofline_fp64"] = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                                    "bytes_per_launch": kb[dom], "avg_launch_ms": round(avg[dom], 5)}
            out["kernel_ms_per_step_fp64"] = {k: round(v, 4) for k, v in per_step.items() if v > 0}
        m.close()
    f32, f64 = fields["fp32"], fields["fp64"]
    valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    out["fp32_vs_fp64"] = {
        "zeta": _zeta_err(b, f32["znv"], f64["znv"], f32["iwegv"], f64["iwegv"]),
        "temp_max_abs": float(np.max(np.abs(f32["temp"] - f64["temp"])[valid])),
        "salt_max_abs": float(np.max(np.abs(f32["salt"] - f64["salt"])[valid])),
        "volume_rel": abs(f32["volume"] - f64["volume"]) / abs(f64["volume"]),
    }
    out["fp32_over_fp64"] = round(out["fp32_value"] / out["fp64_value"], 3)
    return out


def _zeta_err(b, z32, z64, w32, w64):
    """FP32-vs-FP64 zeta differences: over all nodes, and over the nodes
    more than two element rings away from every element whose wet/dry state
    differs between the two runs."""
    dz = np.abs(np.asarray(z32, np.float64) - np.asarray(z64, np.float64))
    near = np.zeros(b.nkn, bool)
    near[b.nen3v[np.nonzero(w32 != w64)[0]].reshape(-1)] = True
    for _ in range(2):
        near[b.nen3v[near[b.nen3v].any(axis=1)].reshape(-1)] = True
    far = dz[~near]
    return {"max": float(dz.max()), "p99": float(np.percentile(dz, 99)), "rms": float(np.sqrt(np.mean(dz ** 2))),
            "elements_wetdry_differ": int((w32 != w64).sum()),
            "max_beyond_2_rings_of_differing_elements": float(far.max()) if far.size else None}


def accuracy_run(b, tr, steps):
    """The FP32 + FP64-refined path and the plain FP64 path side by side
    (same basin, forcing and substeps rule): the zeta difference after
    `steps` steps."""
    import shyhip
    from shyhip import mesh
    ms = {p: shyhip.Model(b, p, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr) for p in ("fp32", "fp64")}
    for m in ms.values():
        m.set_wind(*mesh.wind_stress(*WIND))
    for n in range(steps):
        lev = _level((n + 1) * PARAMS["dt"])
        for m in ms.values():
            m.step(lev, istot=ISTOT)
    z = {p: m.pull("znv") for p, m in ms.items()}
    w = {p: m.pull("iwegv") for p, m in ms.items()}
    for m in ms.values():
        m.close()
    return _zeta_err(b, z["fp32"], z["fp64"], w["fp32"], w["fp64"])


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(b, tr, args, threads):
    """The oracle (reference precision policy, ILUT + BiCGStab, substeps from
    scalar_stability) on the same basin for a bounded time, on `threads`
    host threads: OpenMP over the element columns and over conz3d's colours
    and nodes, where the reference runs in parallel; the solve is serial, as
    SPARSKIT's is.  Returns (steps, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from shyhip import mesh
    pyoracle.load("f32").oc_set_threads(threads)
    om = pyoracle.OracleModel(b, "f32", params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    om.set_wind(*mesh.wind_stress(*WIND))
    t = 0.0
    t0 = time.perf_counter()
    n = 0
    while n < args.cpu_steps or (time.perf_counter() - t0) < args.cpu_seconds and n < 200:
        t += PARAMS["dt"]
        om.step(_level(t), istot=ISTOT)
        n += 1
        if time.perf_counter() - t0 > 4 * args.cpu_seconds:
            break
    el = time.perf_counter() - t0
    pyoracle.load("f32").oc_set_threads(1)
    return n, el


def ensemble_members(args):
    """Uneven lagoons (log-uniform sizes), each with its own wind (0-10 m/s:
    stronger wind sets the surface down by more than the 1 m top z-layer on
    the upwind side of these shallow lagoons, which stops the model as it
    stops the reference, check_volume) and tidal phase, T and S; one
    vertical grid of args.ens_layers layers."""
    from shyhip import mesh
    rng = np.random.default_rng(np.uint64(args.seed + 404))
    out = []
    for i in range(args.ens_members):
        n = int(math.exp(rng.uniform(math.log(args.ens_min), math.log(args.ens_max))))
        lx = 40000.0 * math.sqrt(n / 50000.0)
        b = mesh.lagoon(n, seed=args.seed + 1000 + i, lx=lx, ly=lx / 2, nlv=args.ens_layers, tile=32)
        ph = float(rng.uniform(0, 2 * math.pi))
        wind = mesh.wind_stress(float(rng.uniform(0, 10)), float(rng.uniform(0, 2 * math.pi)))
        out.append(dict(basin=b, wind=wind, level=(lambda t, ph=ph: _level(t, ph)),
                        tracers=dict(temp=mesh.smooth_tracer(b, args.seed + 2000 + i, 15.0, 25.0),
                                     salt=mesh.smooth_tracer(b, args.seed + 3000 + i, 25.0, 38.0))))
    return out


def run_ensemble(mems, prec, steps, warmup, params=None):
    from shyhip import ensemble
    ens = ensemble.Ensemble(mems, prec, params=params or PARAMS, visv=1e-3, difv=1e-4, deferred=True)
    ens.set_level_series(PARAMS["dt"] * (warmup + steps + 2))     # each member's boundary file on the device
    for _ in range(warmup):
        ens.step(istot=ISTOT)
    ens.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step(istot=ISTOT)
    ens.sync()
    el = time.perf_counter() - t0
    z = [(ens.pull("znv", i), ens.pull("iwegv", i)) for i in range(len(mems))]
    ens.close()
    EL = sum(int(m["basin"].ilhv.sum()) for m in mems)
    return EL * steps / el, el, z


def ensemble_zeta_err(mems, r32, r64):
    """|dz| between the FP32 and FP64 ensembles: max and p99 over all nodes,
    and max over the nodes whose incident elements are wet or dry alike in
    both runs (a node next to an element that dried in one run only follows
    a different equation there)."""
    dz_all, dz_same, differ = [], [], 0
    for mem, (za, wa), (zb, wb) in zip(mems, r32, r64):
        b = mem["basin"]
        dz = np.abs(za - zb)
        diff_el = (wa != 0) != (wb != 0)
        touched = np.zeros(b.nkn, bool)
        touched[b.nen3v[diff_el].ravel()] = True
        differ += int(diff_el.sum())
        dz_all.append(dz)
        dz_same.append(dz[~touched])
    a, s = np.concatenate(dz_all), np.concatenate(dz_same)
    return {"max": float(a.max()), "p99": float(np.percentile(a, 99)),
            "max_same_wetdry": float(s.max()) if s.size else None, "elements_wetdry_differ": differ}


def run_members_alone(mems, prec, steps, warmup, params=None):
    """The same members one after the other, each its own model (no batching)."""
    import shyhip
    tot_el, tot_t = 0, 0.0
    for mem in mems:
        m = shyhip.Model(mem["basin"], prec, params=params or PARAMS, visv=1e-3, difv=1e-4, tracers=mem["tracers"])
        m.set_deferred_checks(True)
        m.set_wind(*mem["wind"])
        rec_t = PARAMS["dt"] * np.arange(-2, warmup + steps + 4)
        m.set_boundary_series(1, rec_t, np.array([mem["level"](t) for t in rec_t], np.float32), intpol=4, dtime0=0.0)
        for _ in range(warmup):
            m.step(None, istot=ISTOT)
        m.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.step(None, istot=ISTOT)
        m.sync()
        tot_t += time.perf_counter() - t0
        tot_el += int(mem["basin"].ilhv.sum()) * steps
        m.close()
    return tot_el / tot_t, tot_t


def run_serving(args):
    """The serving workload: a Poisson stream (4 requests/s) of storm-surge
    forecasts on bays of 5k-400k nodes (log-uniform), 360-1440 steps of 60 s
    (6-24 simulated hours), 8 layers, FP32 + refinement, through the
    continuous-batching server (runs joining and leaving one union model
    between steps) and, beside it, through one model and stream per run."""
    from shyhip import serve
    params = dict(dt=60.0, ireib=5, block_rows=256, tri_sweeps=3)

    def stream():
        return serve.poisson_stream(args.srv_requests, rate=4.0, seed=args.seed + 505,
                                    sizes=(5000, args.srv_max_nodes), steps=(args.srv_min_steps, args.srv_max_steps),
                                    nlv=8)
    reqs = stream()
    bsrv = serve.BatchServer("fp32", params=params, slots=args.srv_batch_slots, hbm_budget=64 << 30,
                             solo_nodes=args.srv_solo_nodes, solo_slots=args.srv_slots)
    done, wall = bsrv.serve(reqs)
    s = serve.summary(done, wall)
    s.update(server=f"continuous batching of the runs of at most {args.srv_solo_nodes} nodes "
                    f"({args.srv_batch_slots} slots), larger runs alone ({args.srv_slots} slots)",
             rate_per_s=4.0,
             union_rebuilds=bsrv.rebuilds, batch_steps=bsrv.batch_steps, run_steps=bsrv.run_steps,
             rebuild_s=round(bsrv.build_s, 3),
             workload=f"{len(reqs)} requests, Poisson 4/s, bays of 5k-{args.srv_max_nodes // 1000}k nodes "
                      f"(log-uniform), 8 layers, {args.srv_min_steps}-{args.srv_max_steps} steps of 60 s, "
                      f"wind ramp to 30 m/s, surge 0.3-1.2 m, FP32 + FP64 refinement, hydro only")
    reqs2 = stream()
    srv = serve.Server("fp32", params=params, slots=args.srv_slots, hbm_budget=64 << 30, deferred=True)
    done2, wall2 = srv.serve(reqs2)
    s["one_model_per_run"] = serve.summary(done2, wall2)
    return s


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1440)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nodes", type=int, default=50000)
    ap.add_argument("--layers", type=int, default=12)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-baroclinic", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--acc-steps", type=int, default=120)
    ap.add_argument("--bcl-steps", type=int, default=60)
    ap.add_argument("--ens-members", type=int, default=16)
    ap.add_argument("--ens-min", type=int, default=5000)
    ap.add_argument("--ens-max", type=int, default=40000)
    ap.add_argument("--ens-layers", type=int, default=10)
    ap.add_argument("--ens-steps", type=int, default=120)
    ap.add_argument("--srv-requests", type=int, default=12)
    ap.add_argument("--srv-max-nodes", type=int, default=400000)
    ap.add_argument("--srv-min-steps", type=int, default=360)
    ap.add_argument("--srv-max-steps", type=int, default=1440)
    ap.add_argument("--srv-slots", type=int, default=4)
    ap.add_argument("--srv-batch-slots", type=int, default=8)
    ap.add_argument("--srv-solo-nodes", type=int, default=60000)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--c3-nodes", type=int, default=500000)
    ap.add_argument("--c3-layers", type=int, default=20)
    ap.add_argument("--c3-steps", type=int, default=30)
    ap.add_argument("--only-c3", action="store_true",
                    help="run only the 500k-node configuration (for its profile) and print its object")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (calibrates HBM counters)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "lagoon50k_pmc_traffic.md"),
                    help="committed FETCH_SIZE/WRITE_SIZE summary that fills roofline.traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design note, "
              "not a measured path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.only_c3:
        print(json.dumps({"c3": c3_run(args, {"fp64": dict(PARAMS), "fp32": dict(PARAMS)})}))
        return

    b, tr = _lagoon(args)
    EL = int(b.ilhv.sum())
    # per precision: the settings the autotuner picks on this basin
    params, tuned = {}, {}
    for prec in ("fp64", "fp32"):
        if args.no_autotune:
            params[prec] = dict(PARAMS)
        else:
            params[prec], rep = autotune_params(b, tr, prec)
            tuned[prec] = {"choice": rep["choice"], "ms_per_step": {
                ",".join(f"{k}={t[k]}" for k in TUNE_GRID): round(t["ms"], 4) for t in rep["timings"]}}
    res = {}
    for prec in ("fp64", "fp32"):
        res[prec] = run_gpu(b, tr, prec, args, params=params[prec])
    thr = {p: EL * args.steps / r["elapsed"] for p, r in res.items()}

    # roofline of the headline path's dominant kernel (largest time per step).
    # The headline workload is the lagoon on the plain FP64 GPU path; the
    # FP32 + FP64-refined path runs the same workload beside it.
    head = res["fp64"]
    kb = _kernel_bytes(b, 8, params["fp64"]["block_rows"])
    per_step = {k: v[0] / head["nt"] for k, v in head["timing"].items()}
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in head["timing"].items()}
    dom = max((k for k in kb if avg.get(k, 0.0) > 0), key=lambda k: per_step[k])
    achieved = kb[dom] / (avg[dom] * 1e-3) / 1e9
    kernel_gbs = {k: round(kb[k] / (avg[k] * 1e-3) / 1e9, 1) for k in kb if avg.get(k, 0.0) > 0}
    traffic = _measured_traffic(args.traffic, dom, "double")
    if args.copy_calibration > 0:
        _copy_calibration(args.copy_calibration)

    out = {
        "metric": "elem*layer*steps/s (3D lagoon, plain FP64 GPU path; the FP32 + FP64-refined path, the batched "
                  "FP32+refine ensemble against plain FP64, and serving p50/p99 reported beside it)",
        "value": round(thr["fp64"], 1),
        "unit": "elem*layer*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * res["fp64"]["elapsed"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D lagoon: {b.nkn} nodes, {b.nel} elements, {args.layers} z-layers, T+S (TVD, "
                               f"substeps from scalar_stability as the reference chooses them), "
                               f"dt=60 s, M2+K1 tide (boundary file at every step, cubic interpolation on the "
                               f"device), 8 m/s wind, seed {args.seed}",
                   "elem_layers": EL, "parallelism": "single GPU"},
        "fp32_refined_value": round(thr["fp32"], 1),
        "fp32_refined_ms_per_step": round(1e3 * res["fp32"]["elapsed"] / args.steps, 4),
        "fp32_over_fp64": round(thr["fp32"] / thr["fp64"], 3),
        "zeta_err_fp32_vs_fp64_m": _zeta_err(b, res["fp32"]["znv"], res["fp64"]["znv"], res["fp32"]["iwegv"],
                                             res["fp64"]["iwegv"]),
        "solver": {p: res[p]["stats"] for p in res},
        "settings": {p: {k: params[p].get(k, 0) for k in TUNE_GRID} for p in params},
        "autotune": tuned or None,
        "kernel_ms_per_step_fp64": {k: round(v, 4) for k, v in per_step.items() if v > 0},
        "kernel_ms_per_step_fp32": {k: round(v[0] / res["fp32"]["nt"], 4) for k, v in res["fp32"]["timing"].items()
                                    if v[0] > 0},
        "kernel_algorithmic_GBps_fp64": kernel_gbs,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": kb[dom], "avg_launch_ms": round(avg[dom], 5)},
    }
    if not args.no_accuracy:
        # The workload clips depths at 0.05 m, the drying threshold hzoff:
        # at the rest level thousands of nodes sit exactly on it, and their
        # elements' wet/dry state follows round-off (the reference's own
        # FP32 and FP64 runs disagree on some of them).  The same lagoon with
        # depths clipped at 0.06 m has no such ties.
        bt, trt = _lagoon(args, hmin=0.06)
        out["zeta_err_fp32_vs_fp64_no_threshold_ties_m"] = dict(
            accuracy_run(bt, trt, args.acc_steps), steps=args.acc_steps,
            workload="the same lagoon with depths clipped at 0.06 m instead of 0.05 m (hzoff)")
    if not args.no_baroclinic:
        # T and S drive the momentum equations.  With the workload's fields
        # (15-25 C, 25-38 psu, salinity falling with depth) the added forcing
        # drains the 1 m top layer of multi-layer elements at this shallow
        # lagoon's inlet within tens of steps, which stops the model as it
        # stops the reference (setdepth).  The coupled run uses a stable
        # stratification (T -0.2 C and S +0.3 psu per layer) under 1/10 of
        # the horizontal contrasts, and is short: the oracle run of it stops
        # after about 160 steps the same way.
        lev = np.arange(b.nlv)[None, :]
        trb = {k: (v[:, :1].mean() + 0.1 * (v[:, :1] - v[:, :1].mean())
                   + (-0.2 if k == "temp" else 0.3) * lev).astype(np.float32) for k, v in tr.items()}
        try:
            bcl = {p: run_gpu(b, trb, p, args, steps=args.bcl_steps, baroclinic=True) for p in ("fp64", "fp32")}
            out["baroclinic"] = {
                "workload": "the same lagoon with T and S driving the momentum equations (ibarcl = 1: UNESCO "
                            "density after each transport, baroclinic pressure gradient in the element columns), "
                            "stable stratification, 1/10 of the T/S horizontal contrasts",
                "steps": args.bcl_steps,
                **{f"{p}_ms_per_step": round(1e3 * r["elapsed"] / args.bcl_steps, 4) for p, r in bcl.items()},
                **{f"{p}_value": round(EL * args.bcl_steps / r["elapsed"], 1) for p, r in bcl.items()},
                "solver": {p: r["stats"] for p, r in bcl.items()},
            }
        except Exception as e:      # a model stop is a result of the run, reported, not a crash of the bench
            out["baroclinic"] = {"error": str(e)}
    if not args.no_ensemble:
        mems = ensemble_members(args)
        # every leg with the settings the autotuner picked for its precision
        e32, t32, z32 = run_ensemble(mems, "fp32", args.ens_steps, 5, params["fp32"])
        e64, t64, z64 = run_ensemble(mems, "fp64", args.ens_steps, 5, params["fp64"])
        a64, ta64 = run_members_alone(mems, "fp64", max(10, args.ens_steps // 4), 3, params["fp64"])
        out["ensemble"] = {
            "workload": f"{len(mems)} lagoons of {args.ens_min // 1000}k-{args.ens_max // 1000}k nodes (log-uniform), "
                        f"{args.ens_layers} layers, own wind 0-10 m/s and tidal phase, T+S, dt=60 s",
            "nodes": int(sum(m["basin"].nkn for m in mems)),
            "elem_layers": int(sum(int(m["basin"].ilhv.sum()) for m in mems)),
            "batched_fp32_value": round(e32, 1), "batched_fp64_value": round(e64, 1),
            "sequential_fp64_value": round(a64, 1),
            "batched_fp32_over_sequential_fp64": round(e32 / a64, 3),
            "batched_fp32_over_batched_fp64": round(e32 / e64, 3),
            "steps": args.ens_steps,
            "settings": {p: {k: params[p].get(k, 0) for k in TUNE_GRID} for p in params},
            "zeta_err_batched_fp32_vs_fp64_m": ensemble_zeta_err(mems, z32, z64),
        }
    if not args.no_serving:
        out["serving"] = run_serving(args)
    if not args.no_c3:
        out["c3"] = c3_run(args, params)
    if not args.no_cpu:
        # all the host threads this process may use (the box's share), and one
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        n, el = cpu_baseline(b, tr, args, threads)
        n1, el1 = cpu_baseline(b, tr, args, 1)
        out["cpu_baseline"] = {"value": round(EL * n / el, 1), "unit": "elem*layer*steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{n} steps of the same workload through the C oracle "
                                         f"(f32 state, ILUT(5,1e-2) + BiCGStab, substeps from "
                                         f"scalar_stability; OpenMP over element columns and conz3d "
                                         f"colours/nodes, serial solve), {el:.1f} s on {threads} threads of "
                                         f"{_cpu_model()} (nproc {os.cpu_count()})",
                               "one_core_value": round(EL * n1 / el1, 1),
                               "one_core_sample": f"{n1} steps in {el1:.1f} s on 1 thread"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
