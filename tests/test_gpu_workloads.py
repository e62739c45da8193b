"""The benchmark workloads under test on the GPU, not only under the bench.

* The level-set vortex benchmark run to t = T (benchmark_vortex3D.cpp:29-101 with
  levelset_RK3 and reini_RK3) on a 64^3 grid against the oracle: phi and the enclosed
  volume (volume1 of fluid_update_fsf.cpp:50-74) at the end of the run.
* The 64 uneven 2-D tanks of the batched workload (seed 1003) with autotuned tile plans,
  mixed-precision projection and graph replay: the lockstep and the scheduled batch give
  the same bits, a sample of tanks run alone gives the same bits as in the batch, and the
  smallest tank agrees with the FP64 oracle within the mixed solve's tolerance.
* The served request stream (seed 1004, 70 % 2-D / 30 % 3-D, one 256x64x128 request)
  under an HBM budget that forces queueing: every request gives the same bits alone as in
  the mix, admission waits for memory to free up, the live tanks never hold more device
  bytes than the budget, and the smallest request agrees with the FP64 oracle within the
  mixed solve's tolerance.
"""
import time

import numpy as np
import pytest

from tests._cases import inner

pytestmark = pytest.mark.gpu

FIELDS = ("phi", "u", "v", "w", "press", "ro", "visc")
# fields after mixed-precision steps against the FP64 oracle (both solves end at
# ||b - Ax|| <= 1e-10 ||b||; the FP64 steps of the same tanks agree to 1e-8)
MIXED_RTOL = 1e-7


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_vortex_64_to_T_vs_oracle():
    """C0 at 64^3 (dt = 1/256, CFL 0.5 on max|u| = 2): all 768 steps to t = T = 3, the
    oracle on the host's cores (its per-cell loops threaded, results bit-identical to one
    thread).  Three runs beside the oracle's:

    * the device's phi transport and reinitialisation fed with the oracle's velocities at
      every step: bit-identical to the oracle over the whole run;
    * the device's own vortex steps, whose prescribed velocities come from the device's sin
      and cos (a last-bit difference from glibc's);
    * the oracle again with every velocity field one ulp larger at every step, the same
      kind of perturbation.

    The reinitialisation picks one-sided differences by comparisons of nearly equal
    values, so last-bit differences flip a choice somewhere in a long run and the run
    moves by a discrete amount there; the device's own run is held to ten times what the
    one-ulp perturbation does to the reference's algorithm itself.  At 32^3 the sphere
    (4.8 cells in radius) is gone by t = T in every run, so the volume needs the finer
    grid."""
    import os
    from oracle.oracle import Domain as ODomain
    from reef3d_amd import Domain as DDomain
    from reef3d_amd.cases import vortex3d
    n, T = 64, 3.0
    dt = 0.5 / (n * 2.0)
    steps = int(round(T / dt))
    f = vortex3d(n)
    kw = dict(nx=n, ny=n, nz=n, XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=1.0 / n, DZM=1.0 / n, dt=dt)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    O, P = ODomain(**kw), ODomain(**kw)
    O.set_control("threads", threads)
    P.set_control("threads", threads)
    D, X = DDomain(**kw), DDomain(**kw)
    phi = f["phi"].copy()
    O.ghost(phi, 4, 52)
    phip = phi.copy()
    zo = [np.zeros(O.shape) for _ in range(5)]
    zp = [np.zeros(O.shape) for _ in range(5)]
    for dom in (D, X):
        dom.put("phi", phi)
        for name in ("u", "v", "w", "ro", "visc"):
            dom.zero(name)
    vol0_dev = D.fluid_update()
    vol0 = O.fluid_update(phi, zo[3], zo[4])
    print(f"vortex {n}^3: {steps} steps, oracle on {threads} threads", flush=True)
    rec = {}
    for s in range(steps):
        # the velocities the oracle's step transports phi with, before it overwrites them
        X.put("u", zo[0])
        X.put("v", zo[1])
        X.put("w", zo[2])
        X.levelset_rk3()
        X.reini("phi", 3)
        vol_ref = O.vortex_step(phi, *zo, s * dt)
        vol = D.vortex_step(s * dt)
        for a in zp[:3]:
            a[...] = np.nextafter(a, np.inf)
        vol_p = P.vortex_step(phip, *zp, s * dt)
        if s + 1 in (steps // 2, steps):
            rec[s + 1] = dict(bit=_rel(inner(X.get("phi")), inner(phi)), dev=_rel(inner(D.get("phi")), inner(phi)),
                              ulp=_rel(inner(phip), inner(phi)), vol=(vol[0], vol_ref[0], vol_p[0]))
        if (s + 1) % 128 == 0:
            print(f"vortex {n}^3: {s + 1} steps", flush=True)
    exact = 4.0 / 3.0 * np.pi * 0.15 ** 3
    for st, r in rec.items():
        vd, vo, vp = r["vol"]
        print(f"vortex {n}^3 t={st * dt:g}: phi max rel diff to the oracle: fed its velocities {r['bit']:.3e}, own "
              f"steps {r['dev']:.3e}, oracle one ulp off {r['ulp']:.3e}; volume1 device {vd:.9e} oracle {vo:.9e} "
              f"oracle one ulp off {vp:.9e}; initial {vol0[0]:.9e}, exact sphere {exact:.9e}", flush=True)
    # volume sums: the device adds block partials, the oracle adds serially
    assert abs(vol0_dev[0] - vol0[0]) <= 1e-13 * vol0[0]
    for st, r in rec.items():
        vd, vo, vp = r["vol"]
        assert r["bit"] == 0.0, f"t={st * dt:g}: device fed the oracle's velocities differs by {r['bit']:.2e}"
        assert r["dev"] <= 10.0 * r["ulp"] + 1e-12, f"t={st * dt:g}: {r['dev']:.2e} against {r['ulp']:.2e}"
        assert abs(vd - vo) <= 10.0 * abs(vp - vo) + 1e-13 * vol0[0]
    assert rec[steps]["vol"][0] > 0.0


def _c3_batch(specs, fields, plans=None, tune=False):
    from reef3d_amd import Batch
    from reef3d_amd.cases import tank_kwargs
    B = Batch()
    ids = []
    for sp, f in zip(specs, fields):
        t = B.add(**tank_kwargs(sp, f))
        for n in ("phi", "u", "v", "w", "press"):
            B.put(t, n, f[n])
        ids.append(t)
    B.solver_config("mixed")
    B.set_control("graphs", 1)
    for p in plans or ():
        B.set_plan(*p)
    if tune:
        B.autotune(reps=1)
    B.ghost_fill("phi", 4, 52)
    B.ghost_fill("u", 1, 10)
    B.ghost_fill("v", 2, 11)
    B.ghost_fill("w", 3, 12)
    B.ghost_fill("press", 4, 40)
    B.fluid_update()
    return B, ids


def test_c3_64_tanks_batch_scheduled_alone_oracle():
    from oracle.oracle import Domain as ODomain
    from reef3d_amd.cases import c3_specs, tank_fields, tank_kwargs
    specs = c3_specs()
    fields = [tank_fields(sp) for sp in specs]
    steps = 2
    t0 = time.perf_counter()
    B, ids = _c3_batch(specs, fields, tune=True)
    plans = B.plans()
    print(f"c3: 64 tanks added and tuned ({len(plans)} plans) in {time.perf_counter() - t0:.1f} s", flush=True)
    it_lock = [B.fc3_step()[0] for _ in range(steps)]
    lock = [{n: B.get(t, n) for n in FIELDS} for t in ids]
    B.close()
    print("c3: lockstep (graph replay) done", flush=True)
    # iteration-level scheduling: the same bits
    S, sids = _c3_batch(specs, fields, plans)
    S.run(steps)
    for q, t in enumerate(sids):
        for n in FIELDS:
            assert np.array_equal(S.get(t, n), lock[q][n]), f"tank {q} {n}: scheduled != lockstep"
    S.close()
    print("c3: scheduled == lockstep", flush=True)
    # a sample of tanks alone (the smallest, the largest and six more), same plan table
    cells = [sp["nx"] * sp["nz"] for sp in specs]
    sample = sorted({int(np.argmin(cells)), int(np.argmax(cells)), 0, 9, 21, 33, 47, 63})
    for q in sample:
        A, (t,) = _c3_batch([specs[q]], [fields[q]], plans)
        it_alone = [A.fc3_step()[0][t] for _ in range(steps)]
        for n in FIELDS:
            assert np.array_equal(A.get(t, n), lock[q][n]), f"tank {q} {n}: alone != batch"
        assert it_alone == [it[ids[q]] for it in it_lock]
        A.close()
    print(f"c3: tanks {sample} alone == batch", flush=True)
    # the smallest tank against the FP64 oracle
    q = int(np.argmin(cells))
    sp, f = specs[q], fields[q]
    O = ODomain(**tank_kwargs(sp, f))
    st = {n: f[n].copy() for n in ("phi", "u", "v", "w", "press")}
    for n, kind, gcv in (("phi", 4, 52), ("u", 1, 10), ("v", 2, 11), ("w", 3, 12), ("press", 4, 40)):
        O.ghost(st[n], kind, gcv)
    st["ro"], st["visc"] = np.zeros_like(st["phi"]), np.zeros_like(st["phi"])
    O.fluid_update(st["phi"], st["ro"], st["visc"])
    for _ in range(steps):
        O.fc3_step(*(st[n] for n in FIELDS))
    for n in ("phi", "u", "w", "press"):
        err = _rel(inner(lock[q][n]), inner(st[n]))
        print(f"c3: tank {q} ({sp['nx']}x{sp['nz']}) {n} vs oracle {err:.2e}")
        assert err < MIXED_RTOL, f"tank {q} {n}: {err:.2e} against the oracle"


def _c4_stream():
    """Ten requests of the served stream with 1..5 steps each; request 8 (3-D) is made the
    largest 3-D size, 256 x 64 x 128."""
    from reef3d_amd.cases import c4_requests, tank_fields
    reqs = c4_requests(10, step_div=100)
    big = next(q for q, r in enumerate(reqs) if r["jdir"] == 1)
    reqs[big].update(nx=256, ny=64, nz=128)
    reqs[big]["depth"] = 0.6 * 128 * reqs[big]["dx"]
    reqs[big]["amp"] = 0.05 * reqs[big]["depth"]
    return reqs, [tank_fields(r) for r in reqs], big


def _submit(S, r, f):
    return S.submit(r["nx"], r["ny"], r["nz"], f["XN"], f["YN"], f["ZN"], r["dx"], r["dx"], f["dt"], r["steps"],
                    f["phi"], f["u"], f["v"], f["w"], f["press"], jdir=r["jdir"], top_group=3)


def test_c4_served_stream_alone_equals_mix_and_budget():
    from reef3d_amd import Batch, Server
    reqs, fields, big = _c4_stream()
    need = [Batch.tank_bytes(nx=r["nx"], ny=r["ny"], nz=r["nz"], XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=r["dx"],
                             DZM=r["dx"], jdir=r["jdir"], top_group=3) for r, f in zip(reqs, fields)]
    budget = max(max(need), 0.4 * sum(need))
    print(f"c4: {len(reqs)} requests, {sum(need) / 1e9:.2f} GB in all, budget {budget / 1e9:.2f} GB; "
          f"request {big} is {reqs[big]['nx']}x{reqs[big]['ny']}x{reqs[big]['nz']}", flush=True)
    cfg = dict(hbm_budget_bytes=budget, precision="mixed", graphs=True)
    S = Server(**cfg)
    ids, t_start = [], time.perf_counter()
    for r, f in zip(reqs, fields):
        delay = r["arrival"] - (time.perf_counter() - t_start)
        if delay > 0:
            time.sleep(delay)
        ids.append(_submit(S, r, f))
    for rid in ids:
        assert S.wait(rid, 300.0) == "done"
    mix = [{n: S.result(rid, n) for n in FIELDS} for rid in ids]
    info = [S.info(rid) for rid in ids]
    stats = S.stats()
    S.close()
    print(f"c4: served in {max(i['t_done'] for i in info):.2f} s, peak {stats['peak_bytes'] / 1e9:.2f} GB", flush=True)
    assert [i["steps"] for i in info] == [r["steps"] for r in reqs]
    assert 0 < stats["peak_bytes"] <= budget
    # the budget held requests back: some request was admitted only after another finished
    first_done = min(i["t_done"] for i in info)
    assert any(i["t_admit"] >= first_done for i in info)
    # each request alone on a server of the same configuration
    for q, (r, f) in enumerate(zip(reqs, fields)):
        A = Server(**cfg)
        rid = _submit(A, r, f)
        assert A.wait(rid, 300.0) == "done"
        for n in FIELDS:
            assert np.array_equal(A.result(rid, n), mix[q][n]), f"request {q} {n}: alone != mix"
        assert A.info(rid)["iters"] == info[q]["iters"]
        A.close()
    print("c4: every request alone == in the mix", flush=True)
    # the smallest served request against the FP64 oracle, step for step as served
    from oracle.oracle import Domain as ODomain
    from reef3d_amd.cases import tank_kwargs
    q = int(np.argmin([r["nx"] * r["ny"] * r["nz"] for r in reqs]))
    r, f = reqs[q], fields[q]
    O = ODomain(**tank_kwargs(r, f))
    st = {n: f[n].copy() for n in ("phi", "u", "v", "w", "press")}
    for n, kind, gcv in (("phi", 4, 52), ("u", 1, 10), ("v", 2, 11), ("w", 3, 12), ("press", 4, 40)):
        O.ghost(st[n], kind, gcv)
    st["ro"], st["visc"] = np.zeros_like(st["phi"]), np.zeros_like(st["phi"])
    O.fluid_update(st["phi"], st["ro"], st["visc"])
    for _ in range(r["steps"]):
        O.fc3_step(*(st[n] for n in FIELDS))
    for n in ("phi", "u", "w", "press"):
        err = _rel(inner(mix[q][n]), inner(st[n]))
        print(f"c4: request {q} ({r['nx']}x{r['ny']}x{r['nz']}, {r['steps']} steps) {n} vs oracle {err:.2e}",
              flush=True)
        assert err < MIXED_RTOL, f"request {q} {n}: {err:.2e} against the oracle"
