"""The C-ABI library: it loads without a GPU and exports exactly what include/*.h declares."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(reef_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_the_boundary():
    names = declared()
    for required in ("reef_ctx_create", "reef_ghost_fill", "reef_convection", "reef_reini",
                     "reef_poisson_solve", "reef_pjm_corr", "reef_fc3_step", "reef_vortex_step"):
        assert required in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "reef3d_amd", "libreef3d_hip.so")
    if not os.path.exists(path):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    from reef3d_amd import hotpath
    lib = hotpath.library()
    for n in declared():
        assert getattr(lib, n).argtypes is not None or n == "reef_last_error", n


def test_context_rejects_bad_arguments_without_a_device():
    from reef3d_amd import hotpath
    lib = hotpath.library()
    assert lib.reef_ctx_create(None, None, None) == -1
    assert b"null" in lib.reef_last_error()
    assert lib.reef_sync(None) == -1


def test_server_rejects_bad_arguments_without_a_device():
    from reef3d_amd import hotpath
    lib = hotpath.library()
    assert lib.reef_server_create(None, None) == -1
    assert b"null" in lib.reef_server_last_error()
    assert lib.reef_server_wait(None, 1, 0.0, None) == -1
