"""The work decomposition of the grid kernels (reef3d_amd/csrc/tiling.h), checked on the
host before any kernel uses it: every interior cell exactly once, nothing out of range,
the block groups partition the items."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_tiling_covers_every_cell_once(tmp_path):
    exe = tmp_path / "tiling_check"
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(HERE, "cpp", "tiling_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
