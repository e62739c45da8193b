#!/usr/bin/env python3
"""Benchmark of the REEF3D free-surface hot path on one MI355X.

Workload (the 3-D wave tank the headline metric is quoted on): 512x64x128 cells,
dx = dy = dz = 0.02 m, water depth 1.6 m, eta = 0.05 cos(2 pi x / 2), Airy velocities in
the water, air at rest, rho 1000/1.205, FP64.  One "step" is one momentum_FC3 time step:
HJ-WENO5 + TVD-RK3 transport of phi with 9 reinitialisation iterations, WENO5 flux
convection of u, v, w in three RK stages, and three pressure projections solved by
Jacobi-preconditioned BiCGStab to ||b - Ax|| <= 1e-10 ||b||.  Inputs are generated on the
host, uploaded once and stay in HBM; the timed region contains the steps only.

Prints one JSON line: value = interior cells x steps / wall time (Mcell-updates/s).
The roofline object is for the dominant kernel (the largest total time among the solver
kernels), measured with HIP events on the library stream inside the timed region; the
cpu_baseline object times the CPU oracle (a single-threaded restatement of the
reference) on a y-slab of the same tank and scales it to the same metric.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic HBM bytes per interior cell of each solver kernel (FP64): arrays each
# kernel must read or write once; stencil neighbours are served on chip
SOLVER_BYTES = {
    "bicg_p": 48,   # r, p, v, aii in; p, ph out
    "bicg_v": 56,   # ph, 3 face densities, r0, r in; v out
    "bicg_t": 56,   # r, v, aii, 3 face densities in; t out
    "bicg_x": 72,   # x, ph, r, v, aii, t, r0 in; x, r out
}
HBM_PEAK_GBS = 8000.0


def make_domain(cfg):
    from reef3d_amd import Domain
    from reef3d_amd.cases import wave_tank
    f = wave_tank(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["depth"], cfg["amp"], cfg["wavelength"])
    D = Domain(cfg["nx"], cfg["ny"], cfg["nz"], f["XN"], f["YN"], f["ZN"], cfg["dx"], cfg["dx"], jdir=1,
               top_group=3, g=(0.0, 0.0, -9.81), dt=cfg["dt"], N44=-cfg["rtol"], N46=20000)
    for n in ("phi", "u", "v", "w", "press"):
        D.put(n, f[n])
    D.start4("phi", 52)
    D.start1("u", 10)
    D.start2("v", 11)
    D.start3("w", 12)
    D.start4("press", 40)
    D.fluid_update()
    D.sync()
    return D, f


def cpu_baseline(cfg, iters_per_step):
    """CPU oracle on an 8-cell-wide y-slab of the tank: one momentum_FC3 step with the
    pressure solves capped at one iteration, plus standalone BiCGStab iterations on the
    slab's operator; per-step time = non-solver time + iters_per_step x iteration time."""
    from oracle.oracle import Domain as ODomain
    from reef3d_amd.cases import wave_tank
    ny = 8
    f = wave_tank(cfg["nx"], ny, cfg["nz"], cfg["dx"], cfg["depth"], cfg["amp"], cfg["wavelength"])
    kw = dict(nx=cfg["nx"], ny=ny, nz=cfg["nz"], XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=cfg["dx"],
              DZM=cfg["dx"], jdir=1, top_group=3, g=(0.0, 0.0, -9.81), dt=cfg["dt"], N44=-cfg["rtol"], N46=1)
    O = ODomain(**kw)
    phi, u, v, w, press = (f[n].copy() for n in ("phi", "u", "v", "w", "press"))
    ro, visc = np.zeros_like(phi), np.zeros_like(phi)
    O.ghost(phi, 4, 52)
    O.ghost(u, 1, 10)
    O.ghost(v, 2, 11)
    O.ghost(w, 3, 12)
    O.fluid_update(phi, ro, visc)
    t0 = time.perf_counter()
    O.fc3_step(phi, u, v, w, press, ro, visc)
    t_step = time.perf_counter() - t0
    cells = cfg["nx"] * ny * cfg["nz"]
    rhs = np.random.default_rng(2000).standard_normal(cells)
    M = O.poisson(phi, np.zeros_like(phi), rhs)
    x = np.zeros_like(phi)
    n_it = 30
    t0 = time.perf_counter()
    O.bicgstab(x, M, rhs, 1e-300, n_it)
    t_it = (time.perf_counter() - t0) / n_it
    t_nonsolver = max(t_step - 3 * t_it, 0.0)
    t_full = t_nonsolver + iters_per_step * t_it
    return {
        "value": cells / t_full / 1e6,
        "unit": "Mcell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"CPU oracle, 1 thread, {cfg['nx']}x{ny}x{cfg['nz']} y-slab of the same tank: one "
                   f"momentum_FC3 step with 1 BiCGStab iteration per solve ({t_step:.2f} s) plus {n_it} "
                   f"BiCGStab iterations ({t_it * 1e3:.1f} ms each), scaled to {iters_per_step:.0f} "
                   f"iterations per step as the GPU solve needed"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the hot path is single-device); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    from reef3d_amd.cases import C1
    cfg = dict(C1)
    D, _ = make_domain(cfg)
    cells = cfg["nx"] * cfg["ny"] * cfg["nz"]
    for _ in range(args.warmup):
        D.momentum_fc3()
    D.sync()
    D.timing(True)
    D.timing_reset()
    iters = 0
    D.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it, _ = D.momentum_fc3()
        iters += it
    D.sync()
    elapsed = time.perf_counter() - t0
    # per-kernel event timings from the timed region
    fams = {}
    for name in ("bicg_init", "bicg_p", "bicg_v", "bicg_t", "bicg_x", "hj", "reini", "momentum", "ghost",
                 "face_rho", "correct", "poisson_rhs", "heaviside", "copy"):
        ms, n = D.timing_get(name)
        fams[name] = (ms, n)
    D.timing(False)
    dom = max(SOLVER_BYTES, key=lambda k: fams[k][0])
    ms, n = fams[dom]
    avg_s = ms / 1e3 / max(n, 1)
    achieved = SOLVER_BYTES[dom] * cells / avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_" + dom + ".json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    value = cells * args.steps / elapsed / 1e6
    out = {
        "metric": "Mcell-updates/s",
        "value": round(value, 4),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "wave_tank_512x64x128: momentum_FC3 step (WENO5 flux u,v,w; HJ-WENO5 phi + 9 reinit "
                        "iterations; 3 BiCGStab projections to ||r|| <= 1e-10 ||b||)",
            "cells": cells,
            "dt": cfg["dt"],
            "bicgstab_iters_per_step": iters / args.steps,
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_launch_us": round(avg_s * 1e6, 2),
            "bytes_per_cell": SOLVER_BYTES[dom],
        },
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in fams.items()},
    }
    if rank == 0 and not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, iters / args.steps)
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
