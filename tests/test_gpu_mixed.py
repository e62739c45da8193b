"""Mixed-precision projection on the GPU: FP32 BiCGStab with FP64 reliable updates.

The mixed solve has no reference counterpart, so its bar is the one the FP64 solve is
held to: the FP64 residual ||b - A x||_2 / ||b||_2 of its result, recomputed by the CPU
oracle with the reference operator (poisson_pcorr assembly, res_calc order), must meet
the 1e-10 criterion; the device-reported residual must agree with the oracle's; and the
solution must agree with the plain FP64 solve of the same system to what that criterion
allows.
"""
import numpy as np
import pytest

from tests._cases import inner

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _system(nx, ny, nz, jdir, seed, dx=0.02):
    from oracle.oracle import Domain as ODomain
    from reef3d_amd import Domain as DDomain
    from reef3d_amd.cases import wave_tank
    f = wave_tank(nx, ny, nz, dx, depth=0.6 * nz * dx, amp=0.05 * 0.6 * nz * dx, wavelength=1.0)
    kw = dict(nx=nx, ny=ny, nz=nz, XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=dx, DZM=dx, jdir=jdir, top_group=3,
              g=(0.0, 0.0, -9.81), dt=0.005, N44=-TOL, N46=20000)
    O, D = ODomain(**kw), DDomain(**kw)
    phi = f["phi"].copy()
    O.ghost(phi, 4, 52)
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    b = rng.standard_normal(n)
    b -= b.mean()
    M = O.poisson(phi, np.zeros(O.shape), b.copy())
    D.put("phi", phi)
    rhs = np.zeros(O.shape)
    inner(rhs)[...] = b.reshape(nx, ny, nz).transpose(2, 1, 0)
    D.put("work1", rhs)
    return O, D, M, b


@pytest.mark.parametrize("shape", [(150, 21, 40, 1), (300, 1, 96, 0)])
def test_mixed_solve_reaches_fp64_residual(shape):
    nx, ny, nz, jdir = shape
    O, D, M, b = _system(nx, ny, nz, jdir, seed=2000 + nx)
    bn = np.linalg.norm(b)
    D.zero("work0")
    iters, nru, rel = D.solver_mixed("work0", "work1", -TOL, 20000)
    x = D.get("work0")
    res = O.residual(x, M, b.copy())
    print(f"mixed {shape}: iterations {iters} reliable updates {nru} rel(dev) {rel:.3e} rel(oracle) {res / bn:.3e}")
    assert nru >= 1 and iters > 0
    assert res <= TOL * bn
    assert abs(rel - res / bn) <= 1e-6 * (res / bn)
    assert abs(D.residual("work0", "work1") - res) <= 1e-6 * res
    # plain FP64 solve of the same system
    D.zero("work2")
    it64, _ = D.solver("work2", "work1", -TOL, 20000)
    x64 = D.get("work2")
    res64 = O.residual(x64, M, b.copy())
    dx = np.abs(inner(x) - inner(x64)).max() / np.abs(inner(x64)).max()
    print(f"fp64 {shape}: iters {it64} rel(oracle) {res64 / bn:.3e}; max|x - x64|/max|x64| {dx:.3e}")
    assert dx < 1e-6


def test_mixed_solve_second_call_and_absolute_stop():
    """The FP32 buffers are reused across calls; an absolute stop (>= 0) is the reference's
    ||r||_2 / N criterion."""
    nx, ny, nz = 96, 9, 32
    O, D, M, b = _system(nx, ny, nz, 1, seed=7)
    bn = np.linalg.norm(b)
    D.zero("work0")
    D.solver_mixed("work0", "work1", -TOL, 20000)
    first = D.get("work0")
    D.zero("work0")
    D.solver_mixed("work0", "work1", -TOL, 20000)
    assert np.array_equal(D.get("work0"), first)
    stop = 1e-6 * bn / (nx * ny * nz)
    D.zero("work0")
    _, nru, rel = D.solver_mixed("work0", "work1", stop, 20000)
    res = O.residual(D.get("work0"), M, b.copy())
    assert res / (nx * ny * nz) < stop
    assert nru >= 1


def test_fc3_step_with_mixed_projection_vs_oracle():
    """A momentum_FC3 step whose three projections use the mixed solve stays within the
    solver tolerance of the oracle's FP64 step."""
    from oracle.oracle import Domain as ODomain
    from reef3d_amd import Domain as DDomain
    from reef3d_amd.cases import wave_tank
    nx, ny, nz, dx = 120, 10, 48, 0.02
    f = wave_tank(nx, ny, nz, dx, depth=0.6, amp=0.03, wavelength=1.0)
    kw = dict(nx=nx, ny=ny, nz=nz, XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=dx, DZM=dx, jdir=1, top_group=3,
              g=(0.0, 0.0, -9.81), dt=0.005, N44=-TOL, N46=20000)
    D, O = DDomain(**kw), ODomain(**kw)
    st = {n: f[n].copy() for n in ("phi", "u", "v", "w", "press")}
    for n, kind, gcv in (("phi", 4, 52), ("u", 1, 10), ("v", 2, 11), ("w", 3, 12), ("press", 4, 40)):
        O.ghost(st[n], kind, gcv)
    st["ro"], st["visc"] = np.zeros_like(st["phi"]), np.zeros_like(st["phi"])
    O.fluid_update(st["phi"], st["ro"], st["visc"])
    for n, a in st.items():
        D.put(n, a)
    D.solver_config("mixed")
    it, _ = D.momentum_fc3()
    _, nru, rel = D.solver_last()
    O.fc3_step(*(st[n] for n in ("phi", "u", "v", "w", "press", "ro", "visc")))
    print(f"fc3 mixed: iterations {it}, last solve: {nru} reliable updates, rel {rel:.2e}")
    assert rel <= TOL
    for n in ("phi", "u", "v", "w", "press"):
        got, ref = inner(D.get(n)), inner(st[n])
        err = np.abs(got - ref).max() / (np.abs(ref).max() or 1.0)
        print(f"  {n}: rel err vs FP64 oracle {err:.2e}")
        assert err < 1e-7, f"{n}: rel err {err:.2e}"
