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


def test_field_size_limit_of_32bit_offsets():
    """The solver walks address FP64 arrays through 32-bit byte offsets: a field of 2^29 or
    more doubles (here a 1024 x 512 x 1024 tank, which would fit in HBM) is refused at
    set-up instead of addressing the wrong memory."""
    from reef3d_amd.hotpath import Batch, ReefError, grid_desc, host_grid_arrays
    from tests._cases import uniform_nodes
    n = (1024, 512, 1024)
    nodes = [uniform_nodes(m, 0.01) for m in n]
    gd, keep = grid_desc(*n, *nodes, 0.01, 0.01, jdir=1, top_group=3)
    with pytest.raises(ReefError, match="too large"):
        host_grid_arrays(gd, 0, n[0])
    assert Batch.tank_bytes(*n, *nodes, 0.01, 0.01, jdir=1, top_group=3) == 0
    # just below the limit the set-up is accepted
    n = (1000, 500, 1000)
    nodes = [uniform_nodes(m, 0.01) for m in n]
    gd, keep = grid_desc(*n, *nodes, 0.01, 0.01, jdir=1, top_group=3)
    P = host_grid_arrays(gd, 0, n[0])[0]
    assert P.shape == (n[0] + 10,)


def test_tank_bytes_include_ghost_programs():
    """reef_batch_tank_bytes counts the compiled ghost programs (16 bytes per op, about
    5-6 % of a 3-D tank), not only the field arrays."""
    from reef3d_amd.hotpath import Batch
    from tests._cases import uniform_nodes
    nx, ny, nz = 128, 32, 64
    nodes = [uniform_nodes(m, 0.02) for m in (nx, ny, nz)]
    total = Batch.tank_bytes(nx, ny, nz, *nodes, 0.02, 0.02, jdir=1, top_group=3)
    sx = (nx + 6 + 31) // 32 * 32
    nalloc = (32 + sx * (ny + 6) * (nz + 6) + 64 + 63) // 64 * 64
    fields = 38 * nalloc * 8 + 13 * nalloc * 4
    # every ghost program writes at least one ghost layer over the 6 faces
    faces = 2 * (nx * ny + ny * nz + nx * nz)
    assert total >= fields + 13 * faces * 16
    assert total < 1.2 * fields
