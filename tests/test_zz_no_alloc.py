"""The allocation-free fluid update (reef_fluid_update_noalloc) against reef_fluid_update.

Two contexts in a row, the second on a smaller grid, each called through the C-ABI with
both entry points from the same phi: REEF_RO, REEF_VISC (ghosts included) and volume1,
volume2 must be bit-identical.  The second, smaller call reuses the static scratch the
first one filled, so stale partials would show up as a different volume.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tank(nx, ny, nz, dx, depth, amp, wavelength):
    from reef3d_amd import Domain
    from reef3d_amd.cases import wave_tank
    f = wave_tank(nx, ny, nz, dx, depth=depth, amp=amp, wavelength=wavelength)
    kw = dict(nx=nx, ny=ny, nz=nz, XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=dx, DZM=dx, jdir=1, top_group=3,
              g=(0.0, 0.0, -9.81), dt=0.005)
    return Domain(**kw), f


@pytest.mark.parametrize("shape", [(48, 10, 32, 0.05, 0.9, 0.04, 1.2), (20, 4, 14, 0.05, 0.4, 0.02, 0.7)],
                         ids=["first", "second_smaller"])
def test_noalloc_fluid_update_matches(shape):
    nx, ny, nz = shape[:3]
    print(f"no-alloc fluid update on {nx}x{ny}x{nz}", flush=True)
    D, f = _tank(*shape)
    D.put("phi", f["phi"])
    D.start4("phi", 52)
    vol_a = D.fluid_update_noalloc()
    ro_a, visc_a = D.get("ro"), D.get("visc")
    D.zero("ro")
    D.zero("visc")
    vol_b = D.fluid_update()
    ro_b, visc_b = D.get("ro"), D.get("visc")
    assert np.array_equal(ro_a, ro_b)
    assert np.array_equal(visc_a, visc_b)
    assert vol_a[0] == vol_b[0] and vol_a[1] == vol_b[1], (vol_a, vol_b)
    assert vol_a[0] > 0.0 and vol_a[1] > 0.0
    # again after the library's own call: the scratch's old contents do not leak in
    vol_c = D.fluid_update_noalloc()
    assert vol_c[0] == vol_b[0] and vol_c[1] == vol_b[1]
    D.close()
