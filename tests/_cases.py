"""Shared helpers of the tests: golden fixture loading and domain construction."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# physical / control values the fixture harness used (oracle/ref_harness.cpp: write_ctrl)
CASES = {
    "box": dict(nx=12, ny=10, nz=8, jdir=1, top_group=21, g=(0.0, 0.0, 0.0), dt=0.01, N44=1e-10),
    "nonuni": dict(nx=10, ny=9, nz=11, jdir=1, top_group=21, g=(0.0, 0.0, 0.0), dt=0.01, N44=1e-10),
    "vortex": dict(nx=16, ny=16, nz=16, jdir=1, top_group=21, g=(0.0, 0.0, 0.0), dt=0.5 / 32.0, N44=1e-10),
    "tank": dict(nx=16, ny=6, nz=12, jdir=1, top_group=3, g=(0.0, 0.0, -9.81), dt=0.005, N44=1e-14),
    "tank2d": dict(nx=24, ny=1, nz=12, jdir=0, top_group=3, g=(0.0, 0.0, -9.81), dt=0.005, N44=1e-14),
    # the tank with reinidisc_fsf (F49 = 0)
    "tankfsf": dict(nx=16, ny=6, nz=12, jdir=1, top_group=3, g=(0.0, 0.0, -9.81), dt=0.005, N44=1e-14, F49=0),
    # the tank with explicit diffusion (D20 = 1, ediff2), viscosity 0.01 in water and 0.002 in air
    "tankvisc": dict(nx=16, ny=6, nz=12, jdir=1, top_group=3, g=(0.0, 0.0, -9.81), dt=0.005, N44=1e-14,
                     visc=(0.01, 0.002), D20=1),
    "tank2dvisc": dict(nx=24, ny=1, nz=12, jdir=0, top_group=3, g=(0.0, 0.0, -9.81), dt=0.005, N44=1e-14,
                       visc=(0.01, 0.002), D20=1),
    # operator-level implicit-diffusion fixtures on a stretched 2-D (x-z) box
    "box2d": dict(nx=14, ny=1, nz=9, jdir=0, top_group=21, g=(0.0, 0.0, 0.0), dt=0.01, N44=1e-10),
    # the tanks with implicit diffusion (D20 = 2, idiff2_FS / idiff2_FS_2D, N43 = 1e-5)
    "tankidiff": dict(nx=16, ny=6, nz=12, jdir=1, top_group=3, g=(0.0, 0.0, -9.81), dt=0.005, N44=1e-14,
                      visc=(0.01, 0.002), D20=2),
    "tank2didiff": dict(nx=24, ny=1, nz=12, jdir=0, top_group=3, g=(0.0, 0.0, -9.81), dt=0.005, N44=1e-14,
                        visc=(0.01, 0.002), D20=2),
}


def load(case):
    return np.load(os.path.join(GOLDEN, case + ".npz"))


def domain_args(case, d):
    c = CASES[case]
    gs = d["grid_scalars"]
    return dict(nx=c["nx"], ny=c["ny"], nz=c["nz"], XN=d["XN"], YN=d["YN"], ZN=d["ZN"], DYM=gs[1], DZM=gs[2],
                jdir=c["jdir"], top_group=c["top_group"], g=c["g"], dt=c["dt"], N44=c["N44"], N46=5000,
                F49=c.get("F49", 1), visc=c.get("visc", (1.0e-6, 1.4938e-5)), D20=c.get("D20", 0))


def inner(a):
    return a[3:-3, 3:-3, 3:-3]


def uniform_nodes(n, d, x0=0.0):
    return x0 + d * np.arange(-5, n + 6, dtype=np.float64)


def ghost_cases(d):
    out = []
    for key in sorted(d.files):
        if key.startswith("gc") and key.endswith("_in"):
            kind, gcv = map(int, key[2:-3].split("_"))
            out.append((kind, gcv, key, key[:-3] + "_out"))
    return out
