"""How far the solver-dependent fields may move when only the rounding of BiCGStab's dot
products changes, which is all a device reduction changes.  This bounds the tolerance
the GPU FC3-step tests use for fields downstream of a solve (SOLVED_RTOL in
tests/test_gpu_parity.py).

The oracle with its dot products accumulated in long double ("dot_ext") follows the
reference's arithmetic everywhere else.  On the tankfsf fixture this alone moves phi by
1.51e-9 relative after one step (a borderline restart / early-exit decision of the solve
goes the other way, within the ||r||/N < 1e-14 criterion); the device shows the same
deviation to the printed digits.
"""
import numpy as np
import pytest

from tests._cases import domain_args, inner, load

SOLVED_RTOL = 1e-8


@pytest.mark.parametrize("case", ["tank", "tank2d", "tankfsf"])
def test_dot_rounding_moves_solved_fields_within_tolerance(case):
    from oracle.oracle import Domain
    d = load(case)
    O = Domain(**domain_args(case, d))
    O.set_control("dot_ext", 1)
    worst = 0.0
    for s in (1, 2, 3):
        f = {n: d[f"tk_{n}_{s - 1}"].copy() for n in ("phi", "u", "v", "w", "press", "ro", "visc")}
        O.fc3_step(f["phi"], f["u"], f["v"], f["w"], f["press"], f["ro"], f["visc"])
        for n in ("phi", "u", "v", "w", "press"):
            ref = inner(d[f"tk_{n}_{s}"])
            worst = max(worst, float(np.abs(inner(f[n]) - ref).max() / (np.abs(ref).max() or 1.0)))
    # the perturbation matters (it is not a no-op) and stays inside the tolerance
    assert worst < SOLVED_RTOL / 5
    if case == "tankfsf":
        assert worst > 1e-10
