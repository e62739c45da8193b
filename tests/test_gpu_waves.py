"""Wave generation and absorption by relaxation zones (iowave, B98 = 2, B99 = 1, 2) on the
device against the reference's own runs (tests/golden/wavebox, tankwave, tank2dwave).

* The set-up (wave parameters, interface width psi of a wave run) is host FP64 with the C
  library: bit-identical to the reference.
* The relaxation operators after wavegen_precalc: the device evaluates the wave (cos, cosh,
  sinh, sin) and the band's sin with its own functions, which may differ from glibc's in
  the last bit, so the fields agree to 1e-13 of their scale.
* Three driver-loop steps with the generator (prologue + momentum_FC3 with its
  relaxations): within the solver tolerance, as every step downstream of a BiCGStab solve.
* A wave tank gives the same bits alone and in a batch beside a tank without waves.
"""
import numpy as np
import pytest

from tests._cases import CASES, WAVES, domain_args, inner, load

pytestmark = pytest.mark.gpu

SOLVED_RTOL = 1e-8
FIELDS = ("phi", "u", "v", "w", "press", "ro", "visc")


def _desc(case):
    from reef3d_amd.hotpath import wave_desc
    return wave_desc(**WAVES[case])


def _dev(case, d, phi):
    from reef3d_amd import Domain
    D = Domain(**domain_args(case, d))
    D.put("phi", phi)
    return D, D.wave_set(_desc(case))


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("case,key,phi", [("wavebox", "wv_scalars", "wv0_phi"), ("tankwave", "tk_wave", "tk_phi_0"),
                                          ("tank2dwave", "tk_wave", "tk_phi_0")])
def test_wave_parameters_match_reference(case, key, phi):
    d = load(case)
    _, params = _dev(case, d, d[phi])
    assert params == tuple(d[key][:5])


def test_wave_relaxation_matches_reference():
    d = load("wavebox")
    D, _ = _dev("wavebox", d, d["wv0_phi"])
    for q in (0, 1):
        t = f"wv{q}_"
        D.put("phi", d[t + "phi"])
        D.put("ro", d[t + "ro"])
        for n in ("u", "v", "w", "press"):
            D.put(n, d[t + n + "_in"])
        D.put("work0", d[t + "phi"])
        D.wave_relax(float(d[t + "time"][0]), u="u", v="v", w="w", press="press", f="work0")
        for n, dev_name in (("u", "u"), ("v", "v"), ("w", "w"), ("press", "press"), ("f", "work0")):
            got, ref = D.get(dev_name), d[t + n + "_out"]
            err = _rel(got, ref)
            print(f"wavebox {t}{n}: max rel diff {err:.2e}")
            assert err < 1e-13, f"{t}{n}: {err:.2e}"
            # outside the zones nothing moves
            same_in = d[t + n + "_in"] if n != "f" else d[t + "phi"]
            untouched = ref == same_in
            assert np.array_equal(got[untouched], ref[untouched])


@pytest.mark.parametrize("case", ["tankwave", "tank2dwave"])
def test_wave_tank_steps_match_reference(case):
    d = load(case)
    D, _ = _dev(case, d, d["tk_phi_0"])
    dt = CASES[case]["dt"]
    for s in (1, 2, 3):
        D.set_control("simtime", (s - 1) * dt)
        for n in FIELDS:
            D.put(n, d[f"tk_{n}_{s - 1}"])
        it, vol = D.momentum_fc3()
        assert it > 0
        for n in ("phi", "u", "v", "w", "press"):
            err = _rel(inner(D.get(n)), inner(d[f"tk_{n}_{s}"]))
            assert err < SOLVED_RTOL, f"{case} step {s} {n}: rel err {err:.2e}"
        # the volumes sum the Heaviside of phi, which follows the solve to its tolerance
        np.testing.assert_allclose(vol, d[f"tk_vol_{s}"][:2], rtol=1e-9)


def test_wave_tank_alone_equals_batch():
    """A 3-D wave tank beside the 2-D tank of the other fixture, without waves, two steps:
    the wave tank's fields are bit-identical to the tank run alone."""
    from reef3d_amd import Batch
    d = load("tankwave")
    d2 = load("tank2d")
    kw = domain_args("tankwave", d)
    kw2 = domain_args("tank2d", d2)
    for k in ("F49", "visc", "D20"):
        kw.pop(k, None)
        kw2.pop(k, None)
    D, _ = _dev("tankwave", d, d["tk_phi_0"])
    B = Batch()
    a = B.add(**kw)
    b = B.add(**kw2)
    for n in FIELDS:
        B.put(a, n, d[f"tk_{n}_0"])
        B.put(b, n, d2[f"tk_{n}_0"])
        D.put(n, d[f"tk_{n}_0"])
    B.wave_set(a, _desc("tankwave"))
    for _ in range(2):
        D.momentum_fc3()
        B.fc3_step()
    for n in ("phi", "u", "v", "w", "press", "ro"):
        assert np.array_equal(B.get(a, n), D.get(n)), n
    B.close()
    D.close()
