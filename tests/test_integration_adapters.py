"""The reference-side adapters (reef3d_amd/integration/reef3d_hip_adapters.h) compile
against the reference's own headers.  Runs only where the reference sources exist."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_adapters_compile_against_reference_headers(tmp_path):
    src = tmp_path / "use_adapters.cpp"
    src.write_text('#include "reef3d_hip_adapters.h"\n'
                   "convection* make_flux(lexer* p) { return new reef_adapt::hip_weno_flux(p); }\n"
                   "convection* make_hj(lexer* p) { return new reef_adapt::hip_weno_hj(p); }\n"
                   "reini* make_reini(lexer* p) { return new reef_adapt::hip_reini(p); }\n"
                   "momentum* make_mom(lexer* p) { return new reef_adapt::hip_momentum_FC3(p); }\n"
                   "solver* make_solv(lexer* p) { return new reef_adapt::hip_bicgstab(p); }\n"
                   "pressure* make_press(lexer* p) { return new reef_adapt::hip_pjm_corr(p); }\n"
                   "timestep* make_ts(lexer* p) { return new reef_adapt::hip_ietimestep(p); }\n")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-w", "-I/opt/conda/include", f"-I{REF}",
           f"-I{os.path.join(ROOT, 'reef3d_amd', 'integration')}", f"-I{os.path.join(ROOT, 'include')}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
