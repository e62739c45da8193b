"""Batched tanks on the GPU: many tanks advanced by the same launches.

A tank's arithmetic depends on the tank alone (its chunks and reduction order are fixed
by its own size), so every field of every tank after a batched momentum_FC3 step must be
bit-identical to the same tank run alone, in FP64 and in the mixed-precision projection;
and a tank run alone matches the CPU oracle within the solver tolerance.
"""
import numpy as np
import pytest

from tests._cases import inner

pytestmark = pytest.mark.gpu

FIELDS = ("phi", "u", "v", "w", "press", "ro", "visc")
SOLVED_RTOL = 1e-8


def _tank(nx, ny, nz, dx, depth, amp, wavelength, dt, jdir):
    from reef3d_amd.cases import wave_tank
    f = wave_tank(nx, ny, nz, dx, depth, amp, wavelength)
    kw = dict(nx=nx, ny=ny, nz=nz, XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=dx, DZM=dx, jdir=jdir, top_group=3,
              g=(0.0, 0.0, -9.81), dt=dt, N44=-1e-10, N46=20000)
    return kw, f


TANKS = [
    _tank(96, 9, 40, 0.02, 0.5, 0.02, 1.0, 0.004, 1),
    _tank(300, 1, 64, 0.025, 0.96, 0.05, 1.5, 0.005, 0),
    _tank(70, 5, 24, 0.03, 0.45, 0.03, 0.8, 0.003, 1),
    _tank(130, 1, 40, 0.04, 0.96, 0.08, 2.0, 0.006, 0),
]


def _initial(B, tid, kw, f):
    B.put(tid, "phi", f["phi"])
    for n in ("u", "v", "w", "press"):
        B.put(tid, n, f[n])


def _prepare(B):
    B.ghost_fill("phi", 4, 52)
    B.ghost_fill("u", 1, 10)
    B.ghost_fill("v", 2, 11)
    B.ghost_fill("w", 3, 12)
    B.ghost_fill("press", 4, 40)
    B.fluid_update()


def _run(tanks, steps, precision, graphs=False, every_step=False, d20=0):
    from reef3d_amd import Batch
    B = Batch()
    ids = [B.add(**kw) for kw, _ in tanks]
    for tid, (kw, f) in zip(ids, tanks):
        _initial(B, tid, kw, f)
    B.solver_config(precision)
    B.set_control("graphs", 1 if graphs else 0)
    B.set_control("D20", d20)
    _prepare(B)
    iters, history = [], []
    for _ in range(steps):
        it, _ = B.fc3_step()
        iters.append(it)
        if every_step:
            history.append([{n: B.get(tid, n) for n in FIELDS} for tid in ids])
    out = [{n: B.get(tid, n) for n in FIELDS} for tid in ids]
    B.close()
    return (history if every_step else out), iters


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_results_identical_to_alone(precision):
    mixed, it_mixed = _run(TANKS, 2, precision)
    for q, t in enumerate(TANKS):
        alone, it_alone = _run([t], 2, precision)
        for n in FIELDS:
            assert np.array_equal(mixed[q][n], alone[0][n]), f"tank {q} {n} differs between batch and alone"
        assert [s[q] for s in it_mixed] == [s[0] for s in it_alone]


def test_batch_with_explicit_diffusion_identical_to_alone():
    """D20 = 1 (ediff2) on a batch: each tank's steps are the bits of the tank run alone,
    and the control reaches the batched launches (the results differ from D20 = 0)."""
    mixed, _ = _run(TANKS, 2, "fp64", graphs=True, d20=1)
    plain, _ = _run(TANKS, 2, "fp64", graphs=True)
    for q, t in enumerate(TANKS):
        alone, _ = _run([t], 2, "fp64", d20=1)
        for n in FIELDS:
            assert np.array_equal(mixed[q][n], alone[0][n]), f"tank {q} {n} differs between batch and alone"
        assert not np.array_equal(mixed[q]["u"], plain[q]["u"]), f"tank {q}: D20 = 1 changed nothing"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_identical(precision):
    """Steps replayed from captured hipGraphs (segments and iteration batches) give the
    same bits as launch-by-launch steps, on every replay; the host reads all fields back
    between the replays."""
    eager, it_eager = _run(TANKS[:3], 3, precision, graphs=False, every_step=True)
    graph, it_graph = _run(TANKS[:3], 3, precision, graphs=True, every_step=True)
    assert it_eager == it_graph
    for step in range(3):
        for q in range(3):
            for n in FIELDS:
                assert np.array_equal(eager[step][q][n], graph[step][q][n]), \
                    f"step {step} tank {q} {n}: graph replay differs"


def test_batched_tank_matches_oracle():
    from oracle.oracle import Domain as ODomain
    res, _ = _run(TANKS, 1, "fp64")
    for q, (kw, f) in enumerate(TANKS):
        O = ODomain(**kw)
        st = {n: f[n].copy() for n in ("phi", "u", "v", "w", "press")}
        for n, kind, gcv in (("phi", 4, 52), ("u", 1, 10), ("v", 2, 11), ("w", 3, 12), ("press", 4, 40)):
            O.ghost(st[n], kind, gcv)
        st["ro"], st["visc"] = np.zeros_like(st["phi"]), np.zeros_like(st["phi"])
        O.fluid_update(st["phi"], st["ro"], st["visc"])
        O.fc3_step(*(st[n] for n in FIELDS))
        for n in ("phi", "u", "w", "press"):
            ref = inner(st[n])
            err = np.abs(inner(res[q][n]) - ref).max() / (np.abs(ref).max() or 1.0)
            assert err < SOLVED_RTOL, f"tank {q} {n}: rel err {err:.2e}"


def test_batch_slots_and_budget():
    from reef3d_amd import Batch, ReefError
    kw0, f0 = TANKS[0]
    # the grid with its boundary groups (the ghost programs of a free-surface top differ
    # from a wall's)
    need = Batch.tank_bytes(**{k: kw0[k] for k in ("nx", "ny", "nz", "XN", "YN", "ZN", "DYM", "DZM", "jdir",
                                                   "top_group")})
    assert need > 0
    B = Batch(hbm_budget_bytes=2.5 * need)
    a = B.add(**kw0)
    b = B.add(**kw0)
    with pytest.raises(ReefError):
        B.add(**kw0)
    B.remove(a)
    c = B.add(**kw0)
    assert c == a                      # the freed slot is reused
    slots, active, used = B.info()
    assert slots == 2 and active == 2 and used <= 2.5 * need
    # tank_bytes is exactly what a tank takes from the device pool, ghost programs included
    assert used == 2 * need
    # a tank added after others were stepped is advanced only once made active with them
    for t in (b, c):
        _initial(B, t, kw0, f0)
    _prepare(B)
    B.set_active([c])
    it, _ = B.fc3_step()
    assert it[c] > 0 and it[b] == 0
    B.close()


def test_autotune_plans_and_identity():
    """The autotuner leaves every tank's state unchanged, records a valid plan for each
    size bucket, and with that plan table a tank's step is the same alone and in the batch."""
    from reef3d_amd import Batch
    B = Batch()
    ids = [B.add(**kw) for kw, _ in TANKS]
    for tid, (kw, f) in zip(ids, TANKS):
        _initial(B, tid, kw, f)
    B.solver_config("mixed")
    _prepare(B)
    before = [{n: B.get(tid, n) for n in FIELDS} for tid in ids]
    nb = B.autotune(reps=1)
    plans = B.plans()
    assert nb == len(plans) >= 2
    for jdir, lg, ty, tz, cpt in plans:
        assert (ty, tz) in ((4, 1), (2, 2), (1, 4)) and (jdir == 1 or ty == 1) and 1 <= cpt <= 64
    for q, tid in enumerate(ids):
        for n in FIELDS:
            assert np.array_equal(before[q][n], B.get(tid, n)), f"autotune changed tank {q} {n}"
    B.fc3_step()
    batch = [{n: B.get(tid, n) for n in FIELDS} for tid in ids]
    B.close()
    for q, (kw, f) in enumerate(TANKS):
        A = Batch()
        for p in plans:
            A.set_plan(*p)
        t = A.add(**kw)
        _initial(A, t, kw, f)
        A.solver_config("mixed")
        _prepare(A)
        A.fc3_step()
        for n in FIELDS:
            assert np.array_equal(batch[q][n], A.get(t, n)), f"tank {q} {n}: tuned plan, alone != batch"
        A.close()


@pytest.mark.parametrize("precision,graphs", [("fp64", False), ("mixed", False), ("mixed", True)])
def test_scheduled_run_identical_to_lockstep(precision, graphs):
    """Iteration-level scheduling (each tank moves through its own steps, no lockstep at the
    pressure solves) gives the same bits and iteration counts as lockstep steps, also when
    the batches of recurring solving sets replay from captured graphs."""
    from reef3d_amd import Batch
    lock, it_lock = _run(TANKS, 2, precision)
    B = Batch()
    ids = [B.add(**kw) for kw, _ in TANKS]
    for tid, (kw, f) in zip(ids, TANKS):
        _initial(B, tid, kw, f)
    B.solver_config(precision)
    B.set_control("graphs", 1 if graphs else 0)
    _prepare(B)
    it = B.run(2)
    sched = [{n: B.get(tid, n) for n in FIELDS} for tid in ids]
    B.close()
    assert [it[t] for t in ids] == [it_lock[-1][q] for q in range(len(TANKS))]
    for q in range(len(TANKS)):
        for n in FIELDS:
            assert np.array_equal(sched[q][n], lock[q][n]), f"tank {q} {n}: scheduled != lockstep"


def test_batch_refuses_differing_shared_parameters():
    """Only dt and F45 may differ between the tanks of a batch; the shared values stay
    those of the first tank added to the empty batch, also after that tank is removed."""
    from reef3d_amd import Batch, ReefError
    kw0, _ = TANKS[0]
    kw2, _ = TANKS[2]
    B = Batch()
    a = B.add(**kw0)
    b = B.add(**dict(kw2, dt=0.001, F45=1.5))
    for bad in (dict(rho=(998.0, 1.205)), dict(g=(0.0, 0.0, -9.8)), dict(N44=-1e-9), dict(F43=0.5),
                dict(N46=100), dict(visc=(2.0e-6, 1.4938e-5))):
        with pytest.raises(ReefError):
            B.add(**dict(kw2, **bad))
    B.remove(a)
    with pytest.raises(ReefError):
        B.add(**dict(kw2, rho=(998.0, 1.205)))
    B.remove(b)
    B.add(**dict(kw2, rho=(998.0, 1.205)))   # an empty batch takes new shared values
    B.close()
