#!/usr/bin/env python3
"""Throughput of the SHYFEM semi-implicit hydro step + T/S transport on one MI355X.

Workload: the 3D lagoon -- 50k nodes of a seeded jittered-point Delaunay
triangulation (~98k elements) in a 40 x 20 km basin, log-normal depth
(median 1.5 m) with intertidal flats, 12 z-layers of 1 m, M2+K1 tide on the
western boundary, uniform 8 m/s wind, T and S tracers (TVD horizontal and
vertical), dt = 60 s.  Synthetic data; inputs resident on the device before
the timed region.

A "step" is one model time step: set_dry, copy_uvz/copy_depth, hydro (column
solves, zeta assembly, sparse solve, wet/dry loop, final transports, depths,
vertical velocity) and scal3sh for T and S.  Two device paths are timed:
plain FP64, and FP32 state/transport with the FP32 Krylov solve inside FP64
iterative refinement (the metric's headline value).  Prints one JSON line.
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


def _lagoon(args):
    from shyhip import mesh
    b = mesh.lagoon(args.nodes, seed=args.seed, lx=40000.0, ly=20000.0, nlv=args.layers, tile=32)
    tr = dict(temp=mesh.smooth_tracer(b, args.seed + 1, 15.0, 25.0),
              salt=mesh.smooth_tracer(b, args.seed + 2, 25.0, 38.0))
    return b, tr


def _level(t):
    from shyhip import mesh
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD), phases=(0.0, 0.7))


PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, itpre=2, block_rows=256, tri_sweeps=3)
WIND = (8.0, 0.9)


def _kernel_bytes(b, p, ntr):
    """Algorithmic HBM bytes per launch of each kernel group (each listed
    array read or written once at its stored width p)."""
    EL = int(b.ilhv.sum())
    NL = int(b.ilhkv.sum())
    E, N = b.nel, b.nkn
    return {
        # utlov vtlov hdeov read, utlnv vtlnv written, ddxv ddyv (2 per layer) written,
        # visv at the interfaces, per element geometry/levels/summaries
        "hydro_intern": 9 * EL * p + (NL + N) * p + E * (12 + 56 + 6 * p + 4 + 48),
        # utlnv vtlnv utlov vtlov ulnv vlnv read, 3 horizontal terms written, c gathered
        "conz3d_elem": 9 * EL * p + NL * p + E * (12 + 56 + 4),
        # horizontal terms read; c hdknv hdkov wlnv difv read, c written
        "conz3d_node": 3 * EL * p + 6 * NL * p + 2 * N * p,
        # ddxv ddyv read, utlnv vtlnv read+written
        "transports_final": 8 * EL * p + E * (12 + 56 + 3 * p + 2 * p),
        "hydro_vertical": 4 * EL * p + 4 * NL * p + N * p,
        "setdepth": EL * p + NL * p + E * 3 * p,
    }


def run_gpu(b, tr, prec, args):
    import shyhip
    from shyhip import mesh
    m = shyhip.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    m.set_wind(*mesh.wind_stress(*WIND))
    t = 0.0
    for _ in range(args.warmup):
        t += PARAMS["dt"]
        m.step(_level(t))
    m.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t += PARAMS["dt"]
        m.step(_level(t))
    m.sync()
    elapsed = time.perf_counter() - t0
    # kernel timing pass (events on the library's stream), outside the timed region
    m.set_timing(True)
    nt = max(5, min(20, args.steps))
    for _ in range(nt):
        t += PARAMS["dt"]
        m.step(_level(t))
    timing = m.timing()
    m.set_timing(False)
    zn = m.pull("znv")
    stats = m.stats()
    vol = m.volume()
    return dict(model=m, elapsed=elapsed, timing=timing, nt=nt, znv=zn, stats=stats, volume=vol, t=t)


def cpu_baseline(b, tr, args):
    """The oracle (reference precision policy, ILUT + BiCGStab) on the same
    basin, a bounded number of steps, one host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from shyhip import mesh
    om = pyoracle.OracleModel(b, "f32", params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    om.set_wind(*mesh.wind_stress(*WIND))
    t = 0.0
    t0 = time.perf_counter()
    n = 0
    while n < args.cpu_steps or (time.perf_counter() - t0) < args.cpu_seconds and n < 200:
        t += PARAMS["dt"]
        om.step(_level(t))
        n += 1
        if time.perf_counter() - t0 > 4 * args.cpu_seconds:
            break
    el = time.perf_counter() - t0
    return n, el


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
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design note, "
              "not a measured path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    b, tr = _lagoon(args)
    EL = int(b.ilhv.sum())
    res = {}
    for prec in ("fp64", "fp32"):
        res[prec] = run_gpu(b, tr, prec, args)
    thr = {p: EL * args.steps / r["elapsed"] for p, r in res.items()}
    zerr = float(np.max(np.abs(res["fp32"]["znv"] - res["fp64"]["znv"])))

    # roofline of the dominant non-solver kernel of the headline (FP32) path
    head = res["fp32"]
    kb = _kernel_bytes(b, 4, len(tr))
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in head["timing"].items()}
    dom = max((k for k in kb if avg.get(k, 0.0) > 0), key=lambda k: avg[k] * (head["timing"][k][1]))
    achieved = kb[dom] / (avg[dom] * 1e-3) / 1e9
    per_step = {k: v[0] / head["nt"] for k, v in head["timing"].items()}
    step_bytes = sum(kb[k] * head["timing"][k][1] / head["nt"] for k in kb if k in head["timing"])

    out = {
        "metric": "elem*layer*steps/s (FP32 + FP64-refined path; plain FP64 path reported beside it)",
        "value": round(thr["fp32"], 1),
        "unit": "elem*layer*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * res["fp32"]["elapsed"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state/transport + f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"3D lagoon: {b.nkn} nodes, {b.nel} elements, {args.layers} z-layers, T+S (TVD), "
                               f"dt=60 s, M2+K1 tide, 8 m/s wind, seed {args.seed}",
                   "elem_layers": EL, "parallelism": "single GPU"},
        "fp64_value": round(thr["fp64"], 1),
        "fp64_ms_per_step": round(1e3 * res["fp64"]["elapsed"] / args.steps, 4),
        "fp32_over_fp64": round(thr["fp32"] / thr["fp64"], 3),
        "zeta_err_fp32_vs_fp64_m": zerr,
        "solver": {p: res[p]["stats"] for p in res},
        "kernel_ms_per_step_fp32": {k: round(v, 4) for k, v in per_step.items() if v > 0},
        "kernel_ms_per_step_fp64": {k: round(v[0] / res["fp64"]["nt"], 4) for k, v in res["fp64"]["timing"].items()
                                    if v[0] > 0},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": kb[dom], "avg_launch_ms": round(avg[dom], 5)},
        "step_algorithmic_GBps_non_solver": round(step_bytes / (sum(per_step.values()) * 1e-3) / 1e9, 1),
    }
    if not args.no_cpu:
        n, el = cpu_baseline(b, tr, args)
        out["cpu_baseline"] = {"value": round(EL * n / el, 1), "unit": "elem*layer*steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{n} steps of the same workload through the C oracle "
                                         f"(f32 state, ILUT(5,1e-2) + BiCGStab), {el:.1f} s on "
                                         f"{platform.processor() or platform.machine()}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
