"""REEF3D case files at the boundary, host only (no GPU): the DIVEMesh grid reader and the
ctrl.txt hot-path keys, against the files the reference's own run read (the fixture
harness stores them beside the grid arrays the reference built from them)."""
import numpy as np
import pytest

from tests._cases import CASES, load

HEADER_BYTES = 74 * 4 + 22 * 8          # read_grid.cpp:76-295: 74 ints and 22 doubles


def _write(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(bytes(data))
    return str(p)


@pytest.mark.parametrize("case", ["box", "nonuni", "vortex", "tank", "tank2d", "tankfsf"])
def test_grid_file_gives_the_references_grid(case, tmp_path):
    from reef3d_amd.hotpath import read_grid_file
    d = load(case)
    c = CASES[case]
    g = read_grid_file(_write(tmp_path, "grid.dat", d["gridfile"]))
    assert (g["nx"], g["ny"], g["nz"], g["jdir"]) == (c["nx"], c["ny"], c["nz"], c["jdir"])
    # the node arrays the reference read from the same file (lexer::XN, ...), bit for bit
    for key, n in (("XN", c["nx"]), ("YN", c["ny"]), ("ZN", c["nz"])):
        assert np.array_equal(g[key], d[key][:n + 11]), key
    gs = d["grid_scalars"]
    assert (g["DYM"], g["DZM"]) == (gs[1], gs[2])
    assert g["group"][1:] == [21, 21, 21, 21, 21, c["top_group"]]


def test_grid_file_refusals(tmp_path):
    from reef3d_amd import ReefError
    from reef3d_amd.hotpath import read_grid_file
    raw = bytearray(load("box")["gridfile"].tobytes())
    solid = bytearray(raw)
    solid[HEADER_BYTES:HEADER_BYTES + 4] = np.int32(-1).tobytes()     # flag4 of cell (0,0,0)
    with pytest.raises(ReefError, match="solid"):
        read_grid_file(_write(tmp_path, "solid.dat", solid))
    with pytest.raises(ReefError, match="truncated"):
        read_grid_file(_write(tmp_path, "short.dat", raw[:HEADER_BYTES + 100]))
    topo = bytearray(raw)
    off = (4 + 6 + 1 + 20 + 6 + 6 + 6 + 9 + 3 + 1 + 1) * 4 + 16 * 8     # toporead
    assert np.frombuffer(bytes(topo[off:off + 4]), np.int32)[0] == 0
    topo[off:off + 4] = np.int32(1).tobytes()
    with pytest.raises(ReefError, match="topography"):
        read_grid_file(_write(tmp_path, "topo.dat", topo))


@pytest.mark.parametrize("case", ["box", "tank2d", "vortex"])
def test_ctrl_file_hot_path_keys(case, tmp_path):
    from reef3d_amd.hotpath import read_ctrl_file
    d = load(case)
    c = CASES[case]
    prm, ctrl = read_ctrl_file(_write(tmp_path, "ctrl.txt", d["ctrlfile"]))
    assert (prm.W1, prm.W2, prm.W3, prm.W4) == (1000.0, 1.0e-6, 1.205, 1.4938e-5)
    assert (prm.W20, prm.W21, prm.W22) == c["g"]
    assert (prm.dt, prm.N44) == (c["dt"], c["N44"])
    assert (prm.F43, prm.F45) == (0.55, 2.1)        # ini.cpp defaults, not in the file
    assert ctrl["F49"] == 1 and ctrl["N48"] == 0 and ctrl["N40"] == (13 if case == "vortex" else 3)


def test_ctrl_file_refuses_other_schemes(tmp_path):
    from reef3d_amd import ReefError
    from reef3d_amd.hotpath import read_ctrl_file
    base = "N 10 1\nD 20 0\nN 40 3\n"
    prm, ctrl = read_ctrl_file(_write(tmp_path, "ok.txt", (base + "/ comment\nN 48 1\nN 47 0.2\n").encode()))
    assert ctrl["N48"] == 1 and ctrl["N47"] == 0.2 and prm.N46 == 250 and prm.W1 == 998.2
    # without a D 20 line the default idiff2_FS (D20 = 2, ini.cpp:391) is read
    assert read_ctrl_file(_write(tmp_path, "d.txt", b"N 10 1\nN 40 3\n"))[1]["D20"] == 2
    for text, key in (("D 20 0\nN 40 3\n", "N 10"), ("N 10 1\nN 40 3\nD 20 4\n", "D 20"),
                      (base + "D 10 5\n", "D 10"), (base + "N 40 2\n", "N 40")):
        with pytest.raises(ReefError, match=key):
            read_ctrl_file(_write(tmp_path, "bad.txt", text.encode()))


def test_ctrl_file_diffusion(tmp_path):
    """D 20 1 (ediff2) and D 20 2 (idiff2_FS) are read from the fixture runs' ctrl.txt with
    their viscosities; D 20 1 is refused together with N 48 1, D 20 3 is refused; single
    keys (N 43, the wave keys) come from reef_ctrl_key."""
    from reef3d_amd import ReefError
    from reef3d_amd.hotpath import read_ctrl_file, read_ctrl_key
    prm, ctrl = read_ctrl_file(_write(tmp_path, "ctrl.txt", load("tankvisc")["ctrlfile"]))
    assert ctrl["D20"] == 1 and (prm.W2, prm.W4) == (0.01, 0.002)
    prm, ctrl = read_ctrl_file(_write(tmp_path, "ci.txt", load("tankidiff")["ctrlfile"]))
    assert ctrl["D20"] == 2 and (prm.W2, prm.W4) == (0.01, 0.002)
    assert read_ctrl_file(_write(tmp_path, "c0.txt", load("tank")["ctrlfile"]))[1]["D20"] == 0
    for text in ("N 10 1\nN 40 3\nD 20 1\nN 48 1\n", "N 10 1\nN 40 3\nD 20 3\n"):
        with pytest.raises(ReefError, match="D 20"):
            read_ctrl_file(_write(tmp_path, "bad.txt", text.encode()))
    path = _write(tmp_path, "keys.txt", b"N 43 1e-7\nB 96 0.5 0.75\nB 96 0.25 1.5\n/ B 98 2\n")
    assert read_ctrl_key(path, "N 43") == [1e-7]
    assert read_ctrl_key(path, "B 96") == [0.25, 1.5]
    assert read_ctrl_key(path, "B 98") == []


# ---- the reference's FP32 state file (cfd_state)
STATE_CASES = ["tank", "tank2d"]


def _state_args(d):
    fields = [d["st_" + n] for n in ("u", "v", "w", "press", "phi")]
    m = d["st_meta"]
    return fields, int(m[0]), int(m[1]), tuple(m[2:8])


@pytest.mark.parametrize("case", STATE_CASES)
def test_cfd_state_file_is_the_references(case, tmp_path):
    """The state file written from the reference's final tank fields and print counters is
    byte for byte the file cfd_state::write wrote for them in the fixture run."""
    from reef3d_amd.hotpath import write_cfd_state_host
    d = load(case)
    fields, count, printcount, times = _state_args(d)
    path = str(tmp_path / "state.r3d")
    write_cfd_state_host(path, fields, count, printcount, times)
    got = np.fromfile(path, dtype=np.uint8)
    ref = d["state_file"]
    assert got.size == ref.size
    assert np.array_equal(got, ref), f"first difference at byte {int(np.argmax(got != ref))}"


@pytest.mark.parametrize("case", STATE_CASES)
def test_cfd_state_read_of_the_references_file(case, tmp_path):
    """Reading the reference's state file gives the FP32-rounded fields on the loop cells
    (ULOOP / VLOOP / WLOOP leave the last staggered row, LOOP every cell) and the counters."""
    from reef3d_amd.hotpath import read_cfd_state_host
    d = load(case)
    c = CASES[case]
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    fields, count, printcount, times = _state_args(d)
    path = _write(tmp_path, "state.r3d", d["state_file"])
    got, gc, gp, gt = read_cfd_state_host(path, nx, ny, nz)
    assert (gc, gp) == (count, printcount) and gt == times
    drop = ((0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 0, 0), (0, 0, 0))   # (k, j, i) rows left out
    for q, (dk, dj, di) in enumerate(drop):
        box = (slice(3, 3 + nz - dk), slice(3, 3 + ny - dj), slice(3, 3 + nx - di))
        want = fields[q][box].astype(np.float32).astype(np.float64)
        assert np.array_equal(got[q][box], want), q
        mask = np.ones(got[q].shape, bool)
        mask[box] = False
        assert not got[q][mask].any(), q                              # untouched cells stay zero


def test_cfd_state_refusals(tmp_path):
    from reef3d_amd import ReefError
    from reef3d_amd.hotpath import read_cfd_state_host
    d = load("tank")
    path = _write(tmp_path, "state.r3d", d["state_file"])
    c = CASES["tank"]
    with pytest.raises(ReefError, match="needs"):
        read_cfd_state_host(path, c["nx"] + 1, c["ny"], c["nz"])
    with pytest.raises(ReefError, match="cannot open"):
        read_cfd_state_host(str(tmp_path / "missing.r3d"), c["nx"], c["ny"], c["nz"])
