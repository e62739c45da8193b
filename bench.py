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
The timed steps replay their launch sequences from hipGraphs (captured in the warm-up).
The roofline object is for the dominant kernel (the largest total time among the solver
kernels), measured with HIP events on the library stream around each launch of one more
step launched one by one; the
cpu_baseline object times the CPU oracle (the C++ restatement of the reference's step,
OpenMP over i slabs with the host's CPU share of threads) on one whole step of the same
tank from the same initial fields.  The
mixed_projection object repeats the timed steps from the same initial state with the
mixed-precision projection (FP32 BiCGStab with FP64 reliable updates, to the same 1e-10
FP64 residual) and compares; mixed_solves (100 standalone solves, FP64 against mixed) and
served (a 16-request sample of the served stream against sequential FP64) add the other
half of the metric; hbm_copy times a 1 GiB device copy (the achievable bandwidth).

--traffic FETCH_DIR WRITE_DIR: no GPU work; turns two rocprofv3 --pmc runs (FETCH_SIZE,
WRITE_SIZE) of `bench.py --steps 1 --warmup 0 --no-cpu --no-solves --no-serve` into
profiles/traffic_<kernel>.json (the roofline's traffic) and a markdown table.

--workload vortex: the level-set vortex deformation benchmark instead: unit cube, 128^3
cells, sphere r = 0.15 at (0.35, 0.35, 0.35), levelset_RK3 + 3 reinit iterations per
step, dt = 1/512 to t = T = 3 (1536 steps); reports Mcell-updates/s, the enclosed volume
against 4/3 pi r^3, and the phi difference to the CPU oracle over the first --check steps.

--workload tanks: 64 independent 2-D (x-z) wave tanks advanced together by the batched
engine (one launch per kernel for all tanks): seed 1003, nx in {256..2048 step 128}, nz in
{64..256 step 32}, dx = dz ~ U[0.01, 0.05], depth = 0.6 nz dz, amplitude ~ U[0.01, 0.1]
depth, wavelength ~ U[1, 4] m, Airy initial fields, per-tank dt = 0.25 dx / max(sqrt(g depth),
max |u|, |w| of the initial fields), mixed-precision projection, each step replayed from
hipGraphs (--no-graphs: launch by launch); --steps momentum_FC3 steps (100 in the full
run).  The same tanks are also run one at a time (a sample of --sequential tanks).

--workload serve: a served request stream: --requests tank requests with seeded Poisson
arrivals (lambda = 200/s, seed 1004), 70% 2-D tanks (nx 256..2048 step 128, nz 64..256
step 32) and 30% 3-D tanks (nx 128..256, ny 32..64, nz 64..128), dx, depth, amplitude and
wavelength drawn as in --workload tanks, each asking for 50..500 momentum_FC3 steps
(divided by --step-div to bound the run); the server (continuous batching, HBM budget
200 GB, mixed-precision projection to 1e-10, the iteration batches of a recurring set of
solving tanks replayed from hipGraphs unless --no-graphs) reports p50 / p99 latency (submit to done)
and solves/s; the same requests are then solved one at a time with the plain FP64
drop-in, and their latencies under the same arrivals follow from FIFO order.

--workload solves: the standalone-solve comparison on the same operator instead: --solves
right-hand sides ~ N(0, 1) per cell (seeds 2000, 2001, ..., mean removed), each solved
in FP64 and in mixed precision from x = 0 to ||b - Ax|| <= 1e-10 ||b||, both with their
8-iteration batches replayed from hipGraphs (--no-graphs: launch by launch); reports
iterations, times, the FP64 residual of both results and max |p - p_FP64|.
"""
import argparse
import math
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
    "bicg_v": 48,   # ph, 3 face densities, r0 in; v out (|r|^2 comes from the x phase)
    "bicg_t": 56,   # r, v, aii, 3 face densities in; t out
    "bicg_x": 72,   # x, ph, r, v, aii, t, r0 in; x, r out
}
# the same kernels of the FP32 inner solve of the mixed-precision projection
# mix_v: the FP32 p update fused with v = A ph (r, p, v, aii, 3 conductances, r0 in; v, p, ph out)
SOLVER_BYTES.update({"mix_v": 44, "bicg32_t": 28, "bicg32_x": 36})
HBM_PEAK_GBS = 8000.0
TIMED_FAMILIES = ("bicg_init", "bicg_p", "bicg_v", "bicg_t", "bicg_x", "mix_init", "mix_v",
                  "bicg32_t", "bicg32_x", "mix_ru", "coef32", "add32", "hj", "reini", "momentum", "ghost",
                  "face_rho", "correct", "poisson_rhs", "heaviside", "copy", "flux")


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


def cpu_threads():
    """CPU threads of the baseline: OMP_NUM_THREADS when set (the GPU host's CPU share),
    else the cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, threads=None):
    """The CPU oracle (the C++ restatement of the reference's step) running one whole
    momentum_FC3 step of the same tank, per-cell loops and BiCGStab sweeps over `threads`
    threads (i slabs), from the same initial fields as the GPU run."""
    from oracle.oracle import Domain as ODomain
    from reef3d_amd.cases import wave_tank
    threads = threads or cpu_threads()
    f = wave_tank(cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["depth"], cfg["amp"], cfg["wavelength"])
    kw = dict(nx=cfg["nx"], ny=cfg["ny"], nz=cfg["nz"], XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=cfg["dx"],
              DZM=cfg["dx"], jdir=1, top_group=3, g=(0.0, 0.0, -9.81), dt=cfg["dt"], N44=-cfg["rtol"], N46=20000)
    O = ODomain(**kw)
    O.set_control("threads", threads)
    phi, u, v, w, press = (f[n].copy() for n in ("phi", "u", "v", "w", "press"))
    ro, visc = np.zeros_like(phi), np.zeros_like(phi)
    O.ghost(phi, 4, 52)
    O.ghost(u, 1, 10)
    O.ghost(v, 2, 11)
    O.ghost(w, 3, 12)
    O.ghost(press, 4, 40)
    O.fluid_update(phi, ro, visc)
    t0 = time.perf_counter()
    iters, _ = O.fc3_step(phi, u, v, w, press, ro, visc)
    t_step = time.perf_counter() - t0
    cells = cfg["nx"] * cfg["ny"] * cfg["nz"]
    return {
        "value": cells / t_step / 1e6,
        "unit": "Mcell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"CPU oracle (C++ restatement of the reference step, g++ -O3 -ffp-contract=off, OpenMP over "
                   f"i slabs), {threads} threads on {cpu_model()}: one whole momentum_FC3 step of the "
                   f"{cfg['nx']}x{cfg['ny']}x{cfg['nz']} tank from the initial fields, {iters} BiCGStab "
                   f"iterations, {t_step:.2f} s"),
        "bicgstab_iters": iters,
        "seconds": round(t_step, 3),
    }


def note(msg):
    """progress line on stderr (the JSON line alone goes to stdout)"""
    print(f"# {msg}", file=sys.stderr, flush=True)


def timed_steps(D, steps, kernel_steps=1, graphs=True):
    """steps momentum_FC3 steps on the wall clock, their launch sequences replayed from
    hipGraphs (segments and iteration batches, captured in the warm-up) with no per-kernel
    events in the timed region; then kernel_steps more steps launched one by one with HIP
    events around each launch for the per-kernel breakdown.  (seconds, iterations,
    families of the instrumented steps, volumes after the timed steps)"""
    D.set_control("graphs", 1 if graphs else 0)
    D.sync()
    iters = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        it, vol = D.momentum_fc3()
        iters += it
    D.sync()
    elapsed = time.perf_counter() - t0
    D.timing(True)
    D.timing_reset()
    for _ in range(kernel_steps):
        D.momentum_fc3()
    D.sync()
    fams = {name: D.timing_get(name) for name in TIMED_FAMILIES}
    D.timing(False)
    return elapsed, iters, fams, vol


# algorithmic bytes per interior cell, averaged over the three stages of one launch family
STENCIL_BYTES = {
    "hj": (40 + 48 + 48) / 3.0,      # phi_s, u, v, w [, phi_n] in; phi out
    "reini": (16 + 24 + 24) / 3.0,   # phi_s [, phi_0] in; phi out
}
# the WENO stencils are bound by VALU issue and load latency, not by HBM: FP64 flops per
# cell (add + mul + 2 fma + trans) counted with SQ_INSTS_VALU_*_F64 on the 512x64x128 tank
# (profiles/wave_tank_512x64x128_stencil_valu.md); the same kernels run here
STENCIL_FP64_FLOPS = {"hj": 554.0, "reini": 733.0, "momentum": 798.0}
FP64_PEAK_TFLOPS = 78.6
MOMENTUM_BYTES = (56 + 64 + 64) / 3.0   # q_in, u, v, w, p, phi [, q_old] in; q out


def stencil_rooflines(fams, cells):
    """HBM and FP64-VALU rates of the WENO stencil kernels from their live launch times."""
    out = {}
    for k, bpc in (("hj", STENCIL_BYTES["hj"]), ("reini", STENCIL_BYTES["reini"]), ("momentum", MOMENTUM_BYTES)):
        ms, n = fams.get(k, (0.0, 0))
        if not n:
            continue
        avg_s = ms / 1e3 / n
        gbs = bpc * cells / avg_s / 1e9
        tf = STENCIL_FP64_FLOPS[k] * cells / avg_s / 1e12
        out[k] = {"avg_launch_us": round(avg_s * 1e6, 2), "hbm_GBs": round(gbs, 1),
                  "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "fp64_TFLOPs": round(tf, 2),
                  "fp64_frac": round(tf / FP64_PEAK_TFLOPS, 4), "bound": "valu"}
    return out


def hbm_copy(nbytes=1 << 30, reps=5):
    """Achievable HBM bandwidth: a device-to-device copy of nbytes (1 GiB), read + written
    bytes over the fastest of `reps` copies timed with HIP events."""
    import torch
    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    ok = bool(torch.equal(dst[:: 1 << 20], src[:: 1 << 20]))
    del src, dst
    torch.cuda.empty_cache()
    gbs = 2.0 * nbytes / (best / 1e3) / 1e9
    return {"bytes": nbytes, "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "best_ms": round(best, 4),
            "checked": ok}


# kernel name -> the family of the bench's SOLVER_BYTES (the traffic tables of the profiles)
TRAFFIC_KERNELS = (("kb_bicg_p<double>", "bicg_p"), ("kb_bicg_v<double, true>", "bicg_v"),
                   ("kb_bicg_t<double, true>", "bicg_t"), ("kb_bicg_x<double>", "bicg_x"),
                   ("kb_mix_v<true>", "mix_v"), ("kb_bicg_t<float, true>", "bicg32_t"),
                   ("kb_bicg_x<float>", "bicg32_x"), ("kb_hj_stage", "hj"), ("kb_reini_stage", "reini"))


def traffic_tables(fetch_dir, write_dir, cells=512 * 64 * 128, out_dir=None):
    """HBM bytes per launch of the solver and stencil kernels from two rocprofv3 --pmc runs
    of `bench.py --steps 1 --warmup 0 --no-cpu --no-solves --no-serve` (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each counter scaled by what it reads for the 1 GiB copy
    of the same run (hbm_copy: 1 GiB read, 1 GiB written).  Writes profiles/traffic_<kernel>.json
    (read by the default run's roofline object) and prints a markdown table."""
    import csv
    import glob

    def per_kernel(d, counter):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        sums, cnt, copy_best = {}, {}, 0.0
        for fn in files:
            with open(fn) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    name, val = row["Kernel_Name"], float(row["Counter_Value"])
                    if "copy" in name.lower() and "kb_" not in name:
                        copy_best = max(copy_best, val)
                    key = (name, row.get("Dispatch_Id", row.get("Correlation_Id")))
                    sums[key] = sums.get(key, 0.0) + val
        fam = {}
        for (name, _), val in sums.items():
            for kname, f in TRAFFIC_KERNELS:
                if kname in name:
                    fam.setdefault(f, []).append(val)
        return {f: sum(v) / len(v) for f, v in fam.items()}, copy_best

    fetch, copy_f = per_kernel(fetch_dir, "FETCH_SIZE")
    write, copy_w = per_kernel(write_dir, "WRITE_SIZE")
    gib_kib = float(1 << 20)
    if copy_f <= 0 or copy_w <= 0:
        raise SystemExit("the 1 GiB copy is missing from the counter runs")
    fs, ws = gib_kib / copy_f, gib_kib / copy_w
    print(f"calibration (1 GiB copy): FETCH_SIZE {copy_f:.0f} KiB -> scale {fs:.4f}; WRITE_SIZE {copy_w:.0f} KiB -> "
          f"scale {ws:.4f}")
    print("| kernel | FETCH_SIZE KiB/disp | WRITE_SIZE KiB/disp | HBM MB/launch (scaled) | algorithmic MB | ratio |")
    print("|---|---:|---:|---:|---:|---:|")
    for _, f in TRAFFIC_KERNELS:
        if f not in fetch or f not in write:
            continue
        hbm = 1024.0 * (fs * fetch[f] + ws * write[f])
        alg = SOLVER_BYTES.get(f)
        alg_mb = alg * cells / 1e6 if alg else None
        print(f"| {f} | {fetch[f]:.0f} | {write[f]:.0f} | {hbm / 1e6:.1f} | "
              f"{'' if alg_mb is None else f'{alg_mb:.1f}'} | {'' if alg_mb is None else f'{hbm / 1e6 / alg_mb:.3f}'} |")
        if alg:
            with open(os.path.join(out_dir or os.path.join(ROOT, "profiles"), "traffic_" + f + ".json"), "w") as fh:
                json.dump({"kernel": f, "bytes_per_launch": int(hbm), "fetch_kib": fetch[f], "write_kib": write[f],
                           "fetch_scale": fs, "write_scale": ws, "calibration": "1 GiB device copy"}, fh, indent=1)


def vortex(args):
    """Level-set vortex deformation benchmark (levelset_RK3 + reini_RK3 on 128^3)."""
    from oracle.oracle import Domain as ODomain
    from reef3d_amd import Domain
    from reef3d_amd.cases import C0, vortex3d
    n, dt = C0["n"], C0["dt"]
    steps = args.steps if args.steps_given else int(round(C0["T"] / dt))
    cells = n ** 3
    f = vortex3d(n)
    kw = dict(nx=n, ny=n, nz=n, XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=1.0 / n, DZM=1.0 / n, jdir=1,
              top_group=21, dt=dt)

    def fresh():
        D = Domain(**kw)
        D.put("phi", f["phi"])
        D.start4("phi", 52)
        for name in ("u", "v", "w", "ro", "visc"):
            D.zero(name)
        return D

    Dw = fresh()
    for s in range(args.warmup):
        Dw.vortex_step(s * dt, volumes=False)
    Dw.sync()
    del Dw
    D = fresh()
    vol0 = D.fluid_update()
    D.sync()
    # the timed run carries no per-kernel events; the breakdown comes from a separate run
    # of the first kernel_steps steps with HIP events around each launch
    t0 = time.perf_counter()
    for s in range(steps):
        D.vortex_step(s * dt, volumes=False)
    D.sync()
    elapsed = time.perf_counter() - t0
    vol_end = D.fluid_update()
    kernel_steps = min(steps, 64)
    Dk = fresh()
    Dk.sync()
    Dk.timing(True)
    Dk.timing_reset()
    for s in range(kernel_steps):
        Dk.vortex_step(s * dt, volumes=False)
    Dk.sync()
    fams = {name: Dk.timing_get(name) for name in TIMED_FAMILIES}
    Dk.timing(False)
    del Dk
    exact = 4.0 / 3.0 * np.pi * 0.15 ** 3
    dom = max(STENCIL_BYTES, key=lambda k: fams[k][0])
    ms, nl = fams[dom]
    avg_s = ms / 1e3 / max(nl, 1)
    achieved = STENCIL_BYTES[dom] * cells / avg_s / 1e9
    out = {
        "metric": "Mcell-updates/s",
        "value": round(cells * steps / elapsed / 1e6, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 4),
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"vortex3d_{n}: levelset_RK3 (HJ-WENO5) + 3 reinit iterations per step, "
                               f"dt = 1/512, t = {steps * dt:g}"},
        "volume": {"initial": float(vol0[0]), "final": float(vol_end[0]), "exact_sphere": exact,
                   "rel_change": float(vol_end[0] / vol0[0] - 1.0),
                   "rel_err_vs_exact": float(vol_end[0] / exact - 1.0)},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_us": round(avg_s * 1e6, 2), "bytes_per_cell": round(STENCIL_BYTES[dom], 2),
                     # what actually bounds it: FP64 VALU issue (see STENCIL_FP64_FLOPS)
                     "valu": stencil_rooflines(fams, cells).get(dom)},
        "kernel_ms_per_step": {k: round(v[0] / kernel_steps, 4) for k, v in fams.items() if v[1]},
    }
    if args.check > 0:
        Dp = fresh()
        O = ODomain(**kw)
        phi = f["phi"].copy()
        O.ghost(phi, 4, 52)
        u, v, w, ro, visc = (np.zeros_like(phi) for _ in range(5))
        worst = 0.0
        t0 = time.perf_counter()
        for s in range(args.check):
            O.vortex_step(phi, u, v, w, ro, visc, s * dt)
        t_cpu = (time.perf_counter() - t0) / args.check
        for s in range(args.check):
            Dp.vortex_step(s * dt, volumes=False)
        got = Dp.get("phi")[3:-3, 3:-3, 3:-3]
        ref = phi[3:-3, 3:-3, 3:-3]
        worst = float(np.abs(got - ref).max() / np.abs(ref).max())
        out["parity"] = {"steps": args.check, "phi_max_rel_err_vs_oracle": worst}
        if not args.no_cpu:
            out["cpu_baseline"] = {"value": round(cells / t_cpu / 1e6, 4), "unit": "Mcell-updates/s", "cores": 1,
                                   "kind": "port",
                                   "sample": f"CPU oracle, 1 thread, the first {args.check} steps of the same "
                                             f"128^3 run ({t_cpu:.2f} s per step)"}
    print(json.dumps(out), flush=True)


def tank_batch(specs, precision, graphs=True, rtol=1e-10):
    from reef3d_amd import Batch
    from reef3d_amd.cases import tank_fields, tank_kwargs
    B = Batch()
    ids = []
    for sp in specs:
        f = tank_fields(sp)
        tid = B.add(**tank_kwargs(sp, f, rtol))
        for n in ("phi", "u", "v", "w", "press"):
            B.put(tid, n, f[n])
        ids.append(tid)
    B.solver_config(precision)
    B.set_control("graphs", 1 if graphs else 0)
    B.ghost_fill("phi", 4, 52)
    B.ghost_fill("u", 1, 10)
    B.ghost_fill("v", 2, 11)
    B.ghost_fill("w", 3, 12)
    B.ghost_fill("press", 4, 40)
    B.fluid_update()
    B.sync()
    return B, ids


def tanks(args):
    """Batched 2-D tanks: lockstep steps replayed from hipGraphs, the same steps with
    iteration-level scheduling, and the same tanks one at a time with the plain FP64
    drop-in."""
    from reef3d_amd.cases import c3_specs
    specs = c3_specs()
    cells = sum(sp["nx"] * sp["nz"] for sp in specs)
    graphs = not args.no_graphs
    # the tile plans: tuned once on the batch, shared by both runs
    B, _ = tank_batch(specs, "mixed", graphs)
    t_tune = time.perf_counter()
    if args.cpt:
        # a fixed chunk length for every size bucket of the batch
        for sp in specs:
            B.set_plan(0, int(math.floor(math.log2(sp["nx"] * sp["nz"]))), 1, 4, args.cpt)
    elif not args.no_tune:
        B.autotune(reps=2)
    t_tune = time.perf_counter() - t_tune
    plans = B.plans()
    B.close()

    def fresh():
        Bx, _ = tank_batch(specs, "mixed", graphs)
        for p_ in plans:
            Bx.set_plan(*p_)
        return Bx

    # lockstep: every step's solves run until the slowest tank has converged
    B = fresh()
    for _ in range(args.warmup):
        B.fc3_step()
    B.sync()
    t0 = time.perf_counter()
    iters = np.zeros(len(specs))
    for q in range(args.steps):
        it, _ = B.fc3_step()
        iters += np.array(it[:len(specs)])
        if (q + 1) % 20 == 0:
            print(f"# tanks lockstep: {q + 1} steps", file=sys.stderr, flush=True)
    B.sync()
    t_lock = time.perf_counter() - t0
    # one more step launch by launch with HIP events around each launch family
    B.timing(True)
    B.fc3_step()
    B.sync()
    fams = {name: B.timing_get(name) for name in TIMED_FAMILIES}
    lock_fields = [B.get(t, "phi") for t in range(len(specs))] if args.steps <= 5 else None
    B.close()
    # iteration-level scheduling: each tank moves through its own steps
    B = fresh()
    B.run(args.warmup)
    B.sync()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        # chunks of 20 steps, each reported when it ends (every tank finishes a chunk
        # before the next one starts)
        n = min(20, args.steps - done)
        B.run(n)
        done += n
        if args.steps > 20:
            print(f"# tanks scheduled: {done} steps", file=sys.stderr, flush=True)
    B.sync()
    t_sched = time.perf_counter() - t0
    same = None
    if lock_fields is not None:
        B.run(1)
        same = all(np.array_equal(B.get(t, "phi"), lock_fields[t]) for t in range(len(specs)))
    B.close()
    # the plain drop-in on the same tanks: one tank at a time, FP64 projection to the same
    # 1e-10 residual, launch by launch (a sample of the tanks, same steps)
    sample = specs[:max(1, min(args.sequential, len(specs)))]
    seq_time, seq_cells = 0.0, 0
    for sp in sample:
        S, _ = tank_batch([sp], "fp64", False)
        for _ in range(args.warmup):
            S.fc3_step()
        S.sync()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            S.fc3_step()
        S.sync()
        seq_time += time.perf_counter() - t1
        seq_cells += sp["nx"] * sp["nz"]
        S.close()
        print(f"# tanks sequential: {sp['nx']}x{sp['nz']} done", file=sys.stderr, flush=True)
    lock_rate = cells * args.steps / t_lock / 1e6
    sched_rate = cells * args.steps / t_sched / 1e6
    seq_rate = seq_cells * args.steps / seq_time / 1e6
    best = max(lock_rate, sched_rate)
    out = {
        "metric": "Mcell-updates/s",
        "value": round(best, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(min(t_lock, t_sched) / args.steps * 1e3, 3),
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f64 (FP32 inner solve)",
        "data": "synthetic",
        "config": {"workload": f"batched_2d_tanks_64: seed 1003, {cells} cells in total, momentum_FC3 steps with "
                               f"mixed-precision projection to ||r|| <= 1e-10 ||b|| per tank",
                   "value_from": "scheduled" if sched_rate >= lock_rate else "lockstep_graphs"},
        "lockstep_graphs": {"Mcell_updates_per_s": round(lock_rate, 3), "ms_per_step": round(t_lock / args.steps * 1e3, 3),
                            "graphs": graphs},
        "scheduled": {"Mcell_updates_per_s": round(sched_rate, 3), "ms_per_step": round(t_sched / args.steps * 1e3, 3),
                      "phi_identical_to_lockstep": same},
        "tank_steps_per_s": round(len(specs) * args.steps / min(t_lock, t_sched), 2),
        "iters_per_step_mean": round(float(iters.mean() / args.steps), 1),
        "iters_per_step_max": round(float(iters.max() / args.steps), 1),
        "sequential": {"kind": "one tank at a time, FP64 projection, launch by launch", "tanks": len(sample),
                       "Mcell_updates_per_s": round(seq_rate, 3),
                       "ms_per_tank_step": round(seq_time / (len(sample) * args.steps) * 1e3, 3)},
        "speedup_vs_sequential": round(best / seq_rate, 3),
        "plans": [dict(jdir=p[0], size_log2=p[1], tile=f"64x{p[2]}x{p[3]}", cpt=p[4]) for p in plans],
        "autotune_s": round(t_tune, 2),
        "kernel_ms_one_lockstep_step": {k: round(v[0], 3) for k, v in fams.items() if v[1]},
    }
    print(json.dumps(out), flush=True)


def serve(args):
    """Served request stream versus the same requests solved one at a time in FP64."""
    print(json.dumps(run_serve(args)), flush=True)


def run_serve(args):
    from reef3d_amd import Batch, Server
    from reef3d_amd.cases import c4_requests, tank_fields
    reqs = c4_requests(args.requests, args.step_div)
    fields = [tank_fields(r) for r in reqs]
    # tile plans of the served tanks: 2-D chunks of --cpt tiles (default 4, the fastest in a
    # sweep of 2, 4 and 8 on this stream) per size bucket, or with --tune the autotuner's
    # plans for the size buckets of the stream (tuned once before serving, each bucket alone)
    t_tune = 0.0
    if not args.tune:
        plans = sorted({(0, int(math.floor(math.log2(r["nx"] * r["nz"]))), 1, 4, args.cpt or 4)
                        for r in reqs if r["jdir"] == 0})
    else:
        t0 = time.perf_counter()
        T = Batch()
        for r, f in zip(reqs, fields):
            t = T.add(nx=r["nx"], ny=r["ny"], nz=r["nz"], XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=r["dx"], DZM=r["dx"],
                      jdir=r["jdir"], top_group=3, g=(0.0, 0.0, -9.81), dt=f["dt"], N44=-1e-10, N46=20000)
            for n in ("phi", "u", "v", "w", "press"):
                T.put(t, n, f[n])
        T.solver_config(args.serve_precision)
        T.autotune()
        plans = T.plans()
        T.close()
        t_tune = time.perf_counter() - t0
    S = Server(hbm_budget_bytes=200e9, precision=args.serve_precision, graphs=not args.no_graphs, plans=plans)
    ids = []
    t_start = time.perf_counter()
    for r, f in zip(reqs, fields):
        delay = r["arrival"] - (time.perf_counter() - t_start)
        if delay > 0:
            time.sleep(delay)
        ids.append(S.submit(r["nx"], r["ny"], r["nz"], f["XN"], f["YN"], f["ZN"], r["dx"], r["dx"], f["dt"], r["steps"],
                            f["phi"], f["u"], f["v"], f["w"], f["press"], jdir=r["jdir"], top_group=3))
    done = 0
    for rid in ids:
        st = S.wait(rid)
        assert st == "done", st
        done += 1
        if done % 8 == 0:
            print(f"# serve: {done}/{len(ids)} done", file=sys.stderr, flush=True)
    info = [S.info(rid) for rid in ids]
    rounds = S.stats()["rounds"]
    S.close()
    lat = np.array([i["t_done"] - i["t_submit"] for i in info])
    makespan = max(i["t_done"] for i in info) - min(i["t_submit"] for i in info)
    # the plain drop-in: one request at a time, FP64 projection, launch by launch
    service = []
    for q, (r, f) in enumerate([] if args.serve_only else zip(reqs, fields)):
        B = Batch()
        t = B.add(nx=r["nx"], ny=r["ny"], nz=r["nz"], XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=r["dx"], DZM=r["dx"],
                  jdir=r["jdir"], top_group=3, g=(0.0, 0.0, -9.81), dt=f["dt"], N44=-1e-10, N46=20000)
        for n in ("phi", "u", "v", "w", "press"):
            B.put(t, n, f[n])
        B.solver_config("fp64")
        B.sync()
        t1 = time.perf_counter()
        B.ghost_fill("phi", 4, 52)
        B.ghost_fill("u", 1, 10)
        B.ghost_fill("v", 2, 11)
        B.ghost_fill("w", 3, 12)
        B.ghost_fill("press", 4, 40)
        B.fluid_update()
        for _ in range(r["steps"]):
            B.fc3_step()
        B.sync()
        service.append(time.perf_counter() - t1)
        B.close()
        if (q + 1) % 8 == 0:
            print(f"# sequential: {q + 1}/{len(reqs)} done", file=sys.stderr, flush=True)
    seq_done, clock = [], 0.0
    for r, sv in zip(reqs, service):
        clock = max(clock, r["arrival"]) + sv
        seq_done.append(clock)
    seq_lat = np.array([d - r["arrival"] for d, r in zip(seq_done, reqs)]) if service else np.zeros(1)
    seq_span = seq_done[-1] - reqs[0]["arrival"] if service else float("nan")
    n2d = sum(r["jdir"] == 0 for r in reqs)
    out = {
        "metric": "served tank solves/s",
        "value": round(len(reqs) / makespan, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f64 (FP32 inner solve)" if args.serve_precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"served_tank_stream: {len(reqs)} requests, Poisson arrivals 200/s (seed 1004), "
                               f"{n2d} 2-D / {len(reqs) - n2d} 3-D, steps 50..500 / {args.step_div}, HBM budget 200 GB, "
                               f"{args.serve_precision} projection to ||r|| <= 1e-10 ||b||",
                   "plans": [list(p_) for p_ in plans], "autotune_s": round(t_tune, 2),
                   "graphs": not args.no_graphs},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                      "max": round(float(lat.max()), 3)},
        "makespan_s": round(makespan, 3),
        "server_rounds": rounds,
        "tank_steps": int(sum(r["steps"] for r in reqs)),
        "sequential_fp64": None if not service else {
            "solves_per_s": round(len(reqs) / seq_span, 4),
            "latency_s": {"p50": round(float(np.percentile(seq_lat, 50)), 3),
                          "p99": round(float(np.percentile(seq_lat, 99)), 3)},
            "service_s_total": round(float(sum(service)), 3)},
        "speedup_vs_sequential": round((len(reqs) / makespan) / (len(reqs) / seq_span), 3) if service else None,
        # per request: nx, ny, nz, steps, served latency (submit to done), FP64 service alone
        "requests": [[r["nx"], r["ny"], r["nz"], r["steps"], round(float(l_), 3),
                      round(float(service[q_]), 3) if service else None]
                     for q_, (r, l_) in enumerate(zip(reqs, lat))],
    }
    return out


def solves(args, cfg):
    """Standalone FP64 vs mixed-precision solves on the wave tank's step-0 operator."""
    print(json.dumps(run_solves(args, cfg)), flush=True)


def run_solves(args, cfg):
    D, _ = make_domain(cfg)
    D.solver_config("fp64", args.delta)   # the reliable-update threshold of the mixed solves
    # both solvers replay their 8-iteration batches as hipGraphs (the same for both)
    D.set_control("graphs", 0 if args.no_graphs else 1)
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    cells = nx * ny * nz
    tol = cfg["rtol"]
    rhs = np.zeros((nz + 6, ny + 6, nx + 6))
    rec = []
    for s in range(args.warmup + args.solves):
        b = np.random.default_rng(2000 + s).standard_normal(cells)
        b -= b.mean()
        bn = float(np.linalg.norm(b))
        rhs[3:-3, 3:-3, 3:-3] = b.reshape(nx, ny, nz).transpose(2, 1, 0)
        D.put("work1", rhs)
        D.zero("work2")
        D.zero("work0")
        D.sync()
        t0 = time.perf_counter()
        it64, _ = D.solver("work2", "work1", -tol, 20000)
        t64 = time.perf_counter() - t0
        t0 = time.perf_counter()
        itm, nru, _ = D.solver_mixed("work0", "work1", -tol, 20000)
        tm = time.perf_counter() - t0
        if s < args.warmup:
            continue
        r64 = D.residual("work2", "work1") / bn
        rm = D.residual("work0", "work1") / bn
        x64 = D.get("work2")[3:-3, 3:-3, 3:-3]
        xm = D.get("work0")[3:-3, 3:-3, 3:-3]
        dmax = float(np.abs(xm - x64).max())
        rec.append((it64, t64, itm, nru, tm, r64, rm, dmax, float(np.abs(x64).max())))
        print(f"solve {s}: fp64 {it64} it {t64 * 1e3:.1f} ms res {r64:.2e} | mixed {itm} it/{nru} upd "
              f"{tm * 1e3:.1f} ms res {rm:.2e} | max|p-p64| {dmax:.2e}", file=sys.stderr, flush=True)
    a = np.array(rec)
    t64, tm = a[:, 1].sum(), a[:, 4].sum()
    out = {
        "metric": "projection solves/s",
        "value": round(len(rec) / tm, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "solves": len(rec),
        "warmup": args.warmup,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f32 inner / f64 defect",
        "data": "synthetic",
        "config": {"workload": f"poisson_solves_{nx}x{ny}x{nz}: wave-tank step-0 operator, b ~ N(0,1) "
                               f"seeds 2000.., x0 = 0, ||b - Ax|| <= {tol:g} ||b||",
                   "graphs": not args.no_graphs},
        "fp64": {"solves_per_s": round(len(rec) / t64, 3), "ms_mean": round(t64 / len(rec) * 1e3, 2),
                 "iters_mean": round(float(a[:, 0].mean()), 1),
                 "final_rel_residual_max": float(a[:, 5].max())},
        "mixed": {"ms_mean": round(tm / len(rec) * 1e3, 2), "iters_mean": round(float(a[:, 2].mean()), 1),
                  "reliable_updates_mean": round(float(a[:, 3].mean()), 2),
                  "final_rel_residual_max": float(a[:, 6].max())},
        "speedup_mixed_vs_fp64": round(t64 / tm, 3),
        "reliable_update_delta": args.delta,
        "max_abs_p_minus_p64": float(a[:, 7].max()),
        "max_rel_p_minus_p64": float((a[:, 7] / a[:, 8]).max()),
    }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison leg")
    ap.add_argument("--no-solves", action="store_true", help="skip the standalone-solve (mixed vs FP64) leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--workload", choices=("fc3", "solves", "vortex", "tanks", "serve"), default="fc3")
    ap.add_argument("--requests", type=int, default=16, help="requests of --workload serve")
    ap.add_argument("--step-div", type=int, default=10, help="--workload serve: steps per request divided by this")
    ap.add_argument("--serve-precision", choices=("fp64", "mixed"), default="mixed", help="--workload serve: projection")
    ap.add_argument("--sequential", type=int, default=8, help="tanks run one at a time in --workload tanks")
    ap.add_argument("--no-graphs", action="store_true", help="launch by launch instead of hipGraph replay")
    ap.add_argument("--no-tune", action="store_true", help="--workload tanks: default tile plans")
    ap.add_argument("--tune", action="store_true", help="--workload serve: autotuned tile plans")
    ap.add_argument("--serve-only", action="store_true", help="--workload serve: skip the sequential FP64 leg")
    ap.add_argument("--cpt", type=int, default=0, help="--workload tanks / serve: this chunk length for every 2-D bucket")
    ap.add_argument("--check", type=int, default=3, help="oracle parity steps of --workload vortex")
    ap.add_argument("--solves", type=int, default=100, help="right-hand sides of --workload solves")
    ap.add_argument("--delta", type=float, default=0.1, help="reliable-update threshold of the mixed solve")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic_*.json from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE)")
    args = ap.parse_args()
    if args.traffic:
        traffic_tables(*args.traffic)
        return
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 3
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the hot path is single-device); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    from reef3d_amd.cases import C1
    cfg = dict(C1)
    if args.workload == "solves":
        solves(args, cfg)
        return
    if args.workload == "vortex":
        vortex(args)
        return
    if args.workload == "tanks":
        tanks(args)
        return
    if args.workload == "serve":
        serve(args)
        return
    note("hbm copy: 1 GiB")
    copy = hbm_copy()
    note("fp64 wave tank: set-up")
    D, _ = make_domain(cfg)
    cells = cfg["nx"] * cfg["ny"] * cfg["nz"]
    vol_init = D.fluid_update()
    D.set_control("graphs", 0 if args.no_graphs else 1)
    for _ in range(args.warmup):
        D.momentum_fc3()
    vol_start = D.fluid_update()
    elapsed, iters, fams, vol_end = timed_steps(D, args.steps, graphs=not args.no_graphs)
    note(f"fp64 wave tank: {args.steps} timed steps, {elapsed / args.steps * 1e3:.1f} ms/step")
    dom = max(("bicg_p", "bicg_v", "bicg_t", "bicg_x"), key=lambda k: fams[k][0])
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
            "graphs": not args.no_graphs,
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
        # achievable bandwidth: a 1 GiB device copy, read + written bytes
        "hbm_copy": copy,
        # one more step launched one by one with HIP events around each launch
        "kernel_ms_per_step": {k: round(v[0], 3) for k, v in fams.items() if v[1]},
        "stencil_rooflines": stencil_rooflines(fams, cells),
        # enclosed water volume (volume1 of fluid_update_fsf) of the tank: at set-up, before
        # and after the timed steps
        "interface_volume": {"initial": float(vol_init[0]), "before_timed": float(vol_start[0]),
                             "after_timed": float(vol_end[0]),
                             "rel_change_timed": float(vol_end[0] / vol_start[0] - 1.0),
                             "rel_change_total": float(vol_end[0] / vol_init[0] - 1.0),
                             "steps_total": args.warmup + args.steps},
    }
    if not args.no_mixed:
        del D
        Dm, _ = make_domain(cfg)
        Dm.solver_config("mixed")
        Dm.set_control("graphs", 0 if args.no_graphs else 1)
        for _ in range(args.warmup):
            Dm.momentum_fc3()
        el_m, it_m, fams_m, vol_m = timed_steps(Dm, args.steps, graphs=not args.no_graphs)
        note(f"mixed wave tank: {args.steps} timed steps, {el_m / args.steps * 1e3:.1f} ms/step")
        _, _, rel = Dm.solver_last()
        dm = max(("mix_v", "bicg32_t", "bicg32_x"), key=lambda k: fams_m[k][0])
        ms_m, n_m = fams_m[dm]
        avg_m = ms_m / 1e3 / max(n_m, 1)
        out["mixed_projection"] = {
            "ms_per_step": round(el_m / args.steps * 1e3, 3),
            "value": round(cells * args.steps / el_m / 1e6, 4),
            "speedup_vs_fp64": round(elapsed / el_m, 3),
            "iters_per_step": it_m / args.steps,
            "last_solve_rel_residual": rel,
            "interface_volume_rel_change_total": float(vol_m[0] / vol_init[0] - 1.0),
            "dominant_kernel": dm,
            "dominant_achieved_GBs": round(SOLVER_BYTES[dm] * cells / avg_m / 1e9, 1),
            "kernel_ms_per_step": {k: round(v[0], 3) for k, v in fams_m.items() if v[1]},
        }
        del Dm
    if not args.no_solves:
        # C2: standalone solves on the step-0 operator, FP64 against mixed precision
        note(f"standalone solves: {args.solves} right-hand sides")
        sv = run_solves(args, cfg)
        out["mixed_solves"] = {k: sv[k] for k in ("solves", "fp64", "mixed", "speedup_mixed_vs_fp64",
                                                  "reliable_update_delta", "max_rel_p_minus_p64")}
        out["mixed_solves"]["workload"] = sv["config"]["workload"]
        out["mixed_solves"]["graphs"] = sv["config"]["graphs"]
    if not args.no_serve:
        # C4: a bounded sample of the served stream against sequential plain FP64 solves
        note("served stream")
        sr = run_serve(args)
        out["served"] = {"solves_per_s": sr["value"], "latency_s": sr["latency_s"], "makespan_s": sr["makespan_s"],
                         "sequential_fp64": sr["sequential_fp64"], "speedup_vs_sequential": sr["speedup_vs_sequential"],
                         "tank_steps": sr["tank_steps"], "workload": sr["config"]["workload"]}
    if rank == 0 and not args.no_cpu:
        note("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(cfg)
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
