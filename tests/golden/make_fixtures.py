"""Regenerate the golden fixtures from the reference's own C++ hot path.

Runs only where the reference sources exist (/root/reference): builds the static
archive and the harness with oracle/Makefile.ref, runs oracle/_ref/ref_harness for
every case and packs each case's .npy dumps into tests/golden/<case>.npz.
The fixtures are data (inputs and the reference's outputs); nothing here ships.

    python tests/golden/make_fixtures.py [case ...]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
CASES = ["box", "nonuni", "vortex", "tank", "tank2d", "tankfsf", "tankvisc", "tank2dvisc", "box2d", "tankidiff",
         "tank2didiff", "wavebox", "tankwave", "tank2dwave"]


def main(cases):
    subprocess.run(["make", "-f", "oracle/Makefile.ref", "-j8"], cwd=ROOT, check=True)
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    for case in cases:
        with tempfile.TemporaryDirectory() as tmp:
            out = os.path.join(tmp, case)
            subprocess.run([harness, case, out], cwd=tmp, check=True,
                           stdout=subprocess.DEVNULL)
            arrays = {f[:-4]: np.load(os.path.join(out, f))
                      for f in sorted(os.listdir(out)) if f.endswith(".npy")}
        dst = os.path.join(ROOT, "tests", "golden", case + ".npz")
        np.savez_compressed(dst, **arrays)
        print(f"{case}: {len(arrays)} arrays -> {os.path.getsize(dst) / 1024:.0f} KiB")


if __name__ == "__main__":
    main(sys.argv[1:] or CASES)
