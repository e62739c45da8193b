"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  Arrays use the GPU field layout:
numpy shape (nz+6, ny+6, nx+6), float64, C order (i fastest), 3 ghost layers.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Params(ctypes.Structure):
    _fields_ = [("rho_water", ctypes.c_double), ("visc_water", ctypes.c_double),
                ("rho_air", ctypes.c_double), ("visc_air", ctypes.c_double),
                ("gi", ctypes.c_double), ("gj", ctypes.c_double), ("gk", ctypes.c_double),
                ("F43", ctypes.c_double), ("F45", ctypes.c_double),
                ("dt", ctypes.c_double), ("N44", ctypes.c_double), ("N46", ctypes.c_int)]


class WaveDesc(ctypes.Structure):
    """orc_wave_desc (oracle/reef_oracle.h)."""
    _fields_ = [("B92", ctypes.c_int), ("B91", ctypes.c_int), ("B91_1", ctypes.c_double), ("B91_2", ctypes.c_double),
                ("B93", ctypes.c_int), ("B93_1", ctypes.c_double), ("B93_2", ctypes.c_double),
                ("B96_1", ctypes.c_double), ("B96_2", ctypes.c_double), ("B98", ctypes.c_int), ("B99", ctypes.c_int),
                ("B101", ctypes.c_int), ("B102", ctypes.c_double), ("B105_1", ctypes.c_double),
                ("B120", ctypes.c_double), ("F60", ctypes.c_double), ("W50", ctypes.c_double)]


def build():
    import subprocess
    root = os.path.dirname(_HERE)
    subprocess.run(["make", "-s", "-f", "oracle/Makefile"], cwd=root, check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 4 + [dp] * 3 + [ctypes.c_double] * 2 + \
            [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Params)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_grid_array",):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp]
        L.orc_tables.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
        L.orc_scalars.argtypes = [ctypes.c_void_p, dp]
        L.orc_ghost.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int]
        L.orc_hj.argtypes = [ctypes.c_void_p] + [dp] * 5
        L.orc_flux.argtypes = [ctypes.c_void_p, ctypes.c_int] + [dp] * 5
        L.orc_reini.argtypes = [ctypes.c_void_p, dp, ctypes.c_int]
        L.orc_reini_gcv.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int]
        L.orc_fluid_update.argtypes = [ctypes.c_void_p] + [dp] * 4
        L.orc_poisson.argtypes = [ctypes.c_void_p] + [dp] * 4
        L.orc_set_control.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        L.orc_set_control.restype = ctypes.c_int
        L.orc_residual.argtypes = [ctypes.c_void_p, dp, dp, dp, dp]
        L.orc_residual.restype = ctypes.c_double
        L.orc_bicgstab.argtypes = [ctypes.c_void_p, dp, dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.orc_bicgstab.restype = ctypes.c_int
        L.orc_pjm.argtypes = [ctypes.c_void_p] + [dp] * 6 + [ctypes.c_double]
        L.orc_pjm.restype = ctypes.c_int
        L.orc_fc3_step.argtypes = [ctypes.c_void_p] + [dp] * 8
        L.orc_fc3_step.restype = ctypes.c_int
        L.orc_bicgstab_var.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.orc_bicgstab_var.restype = ctypes.c_int
        L.orc_veclen.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_veclen.restype = ctypes.c_long
        L.orc_idiff_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double] + [dp] * 9
        L.orc_idiff.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double] + [dp] * 8
        L.orc_idiff.restype = ctypes.c_int
        L.orc_vortex_step.argtypes = [ctypes.c_void_p] + [dp] * 6 + [ctypes.c_double, dp]
        L.orc_phi_rk3.argtypes = [ctypes.c_void_p] + [dp] * 4
        L.orc_timestep.argtypes = [ctypes.c_void_p] + [dp] * 3 + [ctypes.c_double, ctypes.c_int, dp]
        L.orc_wave_set.argtypes = [ctypes.c_void_p, ctypes.POINTER(WaveDesc), dp]
        L.orc_wave_params.argtypes = [ctypes.c_void_p, dp]
        L.orc_set_simtime.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_wave_relax.argtypes = [ctypes.c_void_p, ctypes.c_double] + [dp] * 7
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Domain:
    """One box domain of the oracle (orc_create)."""

    def __init__(self, nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir=1, top_group=21,
                 rho=(1000.0, 1.205), visc=(1.0e-6, 1.4938e-5), g=(0.0, 0.0, 0.0),
                 F43=0.55, F45=2.1, dt=0.01, N44=1e-10, N46=5000, F49=1, D20=0, N43=None):
        self.nx, self.ny, self.nz, self.jdir = nx, ny, nz, jdir
        self.shape = (nz + 6, ny + 6, nx + 6)
        groups = (ctypes.c_int * 7)(0, 21, 21, 21, 21, 21, top_group)
        self.prm = Params(rho[0], visc[0], rho[1], visc[1], g[0], g[1], g[2], F43, F45, dt, N44, N46)
        self._xn = [np.ascontiguousarray(a, dtype=np.float64) for a in (XN, YN, ZN)]
        self.h = lib().orc_create(nx, ny, nz, jdir, _p(self._xn[0]), _p(self._xn[1]), _p(self._xn[2]),
                                  DYM, DZM, groups, ctypes.byref(self.prm))
        if F49 != 1 and lib().orc_set_control(self.h, b"F49", float(F49)) != 0:
            raise ValueError(f"F49={F49} not supported")
        if D20 != 0 and lib().orc_set_control(self.h, b"D20", float(D20)) != 0:
            raise ValueError(f"D20={D20} not supported")
        if N43 is not None and lib().orc_set_control(self.h, b"N43", float(N43)) != 0:
            raise ValueError(f"N43={N43} not supported")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def zeros(self):
        return np.zeros(self.shape)

    def grid_array(self, which, axis):
        n = (self.nx, self.ny, self.nz)[axis] + 10
        out = np.zeros(n)
        lib().orc_grid_array(self.h, which, axis, _p(out))
        return out

    def tables(self, axis):
        n = (self.nx, self.ny, self.nz)[axis] + 2
        qf, cf, isf = np.zeros((n, 2, 6, 2)), np.zeros((n, 2, 6)), np.zeros((n, 2, 6, 3))
        lib().orc_tables(self.h, axis, _p(qf), _p(cf), _p(isf))
        return qf, cf, isf

    def scalars(self):
        out = np.zeros(5)
        lib().orc_scalars(self.h, _p(out))
        return dict(DXM=out[0], DYM=out[1], DZM=out[2], psi=out[3], cellnum=int(out[4]))

    def ghost(self, f, kind, gcv):
        lib().orc_ghost(self.h, _p(f), kind, gcv)

    def hj(self, phi, u, v, w, L):
        lib().orc_hj(self.h, _p(phi), _p(u), _p(v), _p(w), _p(L))

    def flux(self, ipol, q, u, v, w, acc):
        lib().orc_flux(self.h, ipol, _p(q), _p(u), _p(v), _p(w), _p(acc))

    def reini(self, phi, iters, gcv=52):
        lib().orc_reini_gcv(self.h, _p(phi), iters, gcv)

    def fluid_update(self, phi, ro, visc):
        vol = np.zeros(2)
        lib().orc_fluid_update(self.h, _p(phi), _p(ro), _p(visc), _p(vol))
        return vol

    def poisson(self, phi, pcorr, rhs):
        n = self.nx * self.ny * self.nz
        M = np.zeros(7 * n)
        lib().orc_poisson(self.h, _p(phi), _p(pcorr), _p(M), _p(rhs))
        return M.reshape(7, n)

    def bicgstab(self, x, M, rhs, stop, maxiter):
        res = np.zeros(1)
        M = np.ascontiguousarray(M)
        it = lib().orc_bicgstab(self.h, _p(x), _p(M), _p(rhs), stop, maxiter, _p(res))
        return it, res[0]

    def bicgstab_var(self, var, x, M, rhs, stop, maxiter):
        """bicgstab_ijk::start with var 1..5 on M (7 rows of veclen(var)) and rhs (vec order)."""
        res = np.zeros(1)
        M = np.ascontiguousarray(M, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        it = lib().orc_bicgstab_var(self.h, int(var), _p(x), _p(M), _p(rhs), stop, maxiter, _p(res))
        return it, res[0]

    def idiff_assemble(self, comp, alpha, vin, u, v, w, visc, au=None, av=None):
        """idiff2_FS(_2D) diff_u/v/w operator of component comp: (M 7 x n, rhs n)."""
        n = self.veclen(comp)
        M, rhs = np.zeros(7 * n), np.zeros(n)
        au = u if au is None else au
        av = v if av is None else av
        lib().orc_idiff_assemble(self.h, int(comp), float(alpha), _p(vin), _p(u), _p(v), _p(w), _p(au), _p(av),
                                 _p(visc), _p(M), _p(rhs))
        return M.reshape(7, n), rhs

    def idiff(self, comp, alpha, diff, vin, u, v, w, visc, au=None, av=None):
        """the whole diff_u/v/w call into diff; returns the iterations"""
        au = u if au is None else au
        av = v if av is None else av
        return lib().orc_idiff(self.h, int(comp), float(alpha), _p(diff), _p(vin), _p(u), _p(v), _p(w), _p(au), _p(av),
                               _p(visc))

    def veclen(self, var):
        return lib().orc_veclen(self.h, int(var))

    def set_control(self, key, value):
        if lib().orc_set_control(self.h, key.encode(), float(value)) != 0:
            raise ValueError(f"control {key}={value} not supported")

    def residual(self, x, M, rhs, r=None):
        """||rhs - A x||_2 in the order of res_calc; r (field layout) receives the residual."""
        M = np.ascontiguousarray(M)
        return lib().orc_residual(self.h, _p(x), _p(M), _p(rhs), _p(r) if r is not None else None)

    def pjm(self, phi, u, v, w, press, pcorr, alpha):
        return lib().orc_pjm(self.h, _p(phi), _p(u), _p(v), _p(w), _p(press), _p(pcorr), alpha)

    def fc3_step(self, phi, u, v, w, press, ro, visc):
        vol = np.zeros(2)
        it = lib().orc_fc3_step(self.h, _p(phi), _p(u), _p(v), _p(w), _p(press), _p(ro), _p(visc), _p(vol))
        return it, vol

    def vortex_step(self, phi, u, v, w, ro, visc, simtime):
        vol = np.zeros(2)
        lib().orc_vortex_step(self.h, _p(phi), _p(u), _p(v), _p(w), _p(ro), _p(visc), simtime, _p(vol))
        return vol

    def phi_rk3(self, phi, u, v, w):
        lib().orc_phi_rk3(self.h, _p(phi), _p(u), _p(v), _p(w))

    def wave_set(self, desc, phi):
        """iowave set-up from a WaveDesc (fields as reef_wave_desc) and psi from phi (inipsi
        with B90 > 0); returns (wT, wL, wk, ww, wA)."""
        lib().orc_wave_set(self.h, ctypes.byref(desc), _p(np.ascontiguousarray(phi)))
        out = np.zeros(5)
        lib().orc_wave_params(self.h, _p(out))
        return tuple(float(x) for x in out)

    def set_simtime(self, t):
        lib().orc_set_simtime(self.h, float(t))

    def wave_relax(self, simtime, phi, ro, u, v, w, press, f):
        """wavegen_precalc at simtime, then u/v/w_relax, p_relax and phi_relax in place."""
        lib().orc_wave_relax(self.h, float(simtime), _p(phi), _p(ro), _p(u), _p(v), _p(w), _p(press), _p(f))

    def timestep(self, u, v, w, N47=0.3, ini=False):
        """ietimestep::start / ::ini: (dt, umax, vmax, wmax)."""
        out = np.zeros(4)
        lib().orc_timestep(self.h, _p(u), _p(v), _p(w), float(N47), 1 if ini else 0, _p(out))
        return tuple(float(x) for x in out)
