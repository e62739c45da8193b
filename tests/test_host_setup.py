"""Host-side set-up of the HIP library (grid arrays, WENO tables, compiled ghost programs)
against the reference's golden vectors; runs without a GPU."""
import numpy as np
import pytest

from tests._cases import CASES, ghost_cases, load


def desc(case, d):
    from reef3d_amd.hotpath import grid_desc
    c = CASES[case]
    gs = d["grid_scalars"]
    return grid_desc(c["nx"], c["ny"], c["nz"], d["XN"], d["YN"], d["ZN"], gs[1], gs[2], c["jdir"], c["top_group"])


@pytest.mark.parametrize("case", ["box", "nonuni", "vortex", "tank", "tank2d"])
def test_grid_arrays_and_tables_bitexact(case):
    from reef3d_amd.hotpath import host_grid_arrays
    d = load(case)
    gd, keep = desc(case, d)
    c = CASES[case]
    for ax, (a, n) in enumerate(zip("XYZ", (c["nx"], c["ny"], c["nz"]))):
        P, DN, DP, qf, cf, isf, sc = host_grid_arrays(gd, ax, n)
        assert np.array_equal(P, d[a + "P"][:n + 10])
        assert np.array_equal(DN, d["D" + a + "N"][:n + 10])
        assert np.array_equal(DP, d["D" + a + "P"][:n + 10])
        x = a.lower()
        assert np.array_equal(qf, d["qf" + x]) and np.array_equal(cf, d["cf" + x])
        assert np.array_equal(isf, d["isf" + x])
        gs = d["grid_scalars"]
        assert sc[0] == gs[0] and sc[3] == gs[3]


@pytest.mark.parametrize("case", ["box", "nonuni"])
def test_ghost_programs_bitexact(case):
    from reef3d_amd.hotpath import host_ghost_apply
    d = load(case)
    gd, keep = desc(case, d)
    for kind, gcv, kin, kout in ghost_cases(d):
        got = host_ghost_apply(gd, kind, gcv, d[kin])
        assert np.array_equal(got, d[kout]), f"start{kind}(gcv={gcv})"


def test_unsupported_geometry_is_rejected():
    from reef3d_amd.hotpath import ReefError, grid_desc, host_ghost_apply
    from tests._cases import uniform_nodes
    gd, keep = grid_desc(8, 1, 8, uniform_nodes(8, 0.1), uniform_nodes(1, 0.1), uniform_nodes(8, 0.1), 0.1, 0.1,
                         jdir=1)
    with pytest.raises(ReefError):
        host_ghost_apply(gd, 4, 52, np.zeros((14, 7, 14)))
