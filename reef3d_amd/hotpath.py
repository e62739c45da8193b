"""Python host mirror of the reference's hot-path interfaces over the C-ABI.

The reference drives this path through C++ strategy objects on lexer/fdm/ghostcell:
convection::start (convection.h:37), reini::start (reini.h:39), solver::start
(solver.h:42), pressure::start (pressure.h:41), ghostcell::start1..4 (ghostcell.h:58-61),
momentum::start (momentum.h) and freesurface/benchmark steps.  `Domain` owns one
device-resident domain (reef_ctx) and exposes the same operations with the same
argument meaning; fields are addressed by name ("u", "v", "w", "press", "phi", "ro",
"visc", "F", "G", "H", "L", "pcorr", "work0".."work3").

Arrays exchanged with the host are float64 with shape (nz+6, ny+6, nx+6) (i fastest,
3 ghost layers), or the reference's own k-fastest field storage with layout="ref".
Every call goes to the HIP library; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libreef3d_hip.so")
_LIB = None

FIELDS = {"u": 0, "v": 1, "w": 2, "press": 3, "phi": 4, "ro": 5, "visc": 6, "F": 7, "G": 8,
          "H": 9, "L": 10, "pcorr": 11, "work0": 12, "work1": 13, "work2": 14, "work3": 15}


class ReefError(RuntimeError):
    pass


class GridDesc(ctypes.Structure):
    _fields_ = [("knox", ctypes.c_int), ("knoy", ctypes.c_int), ("knoz", ctypes.c_int),
                ("j_dir", ctypes.c_int), ("XN", ctypes.POINTER(ctypes.c_double)),
                ("YN", ctypes.POINTER(ctypes.c_double)), ("ZN", ctypes.POINTER(ctypes.c_double)),
                ("DYM", ctypes.c_double), ("DZM", ctypes.c_double), ("group", ctypes.c_int * 7)]


class Params(ctypes.Structure):
    _fields_ = [("W1", ctypes.c_double), ("W2", ctypes.c_double), ("W3", ctypes.c_double),
                ("W4", ctypes.c_double), ("W20", ctypes.c_double), ("W21", ctypes.c_double),
                ("W22", ctypes.c_double), ("F43", ctypes.c_double), ("F45", ctypes.c_double),
                ("dt", ctypes.c_double), ("N44", ctypes.c_double), ("N46", ctypes.c_int)]


class WaveDesc(ctypes.Structure):
    """reef_wave_desc: the iowave keys of ctrl.txt (B91..B120, F60, W50)."""
    _fields_ = [("B92", ctypes.c_int), ("B91", ctypes.c_int), ("B91_1", ctypes.c_double), ("B91_2", ctypes.c_double),
                ("B93", ctypes.c_int), ("B93_1", ctypes.c_double), ("B93_2", ctypes.c_double),
                ("B96_1", ctypes.c_double), ("B96_2", ctypes.c_double), ("B98", ctypes.c_int), ("B99", ctypes.c_int),
                ("B101", ctypes.c_int), ("B102", ctypes.c_double), ("B105_1", ctypes.c_double),
                ("B120", ctypes.c_double), ("F60", ctypes.c_double), ("W50", ctypes.c_double)]


def wave_desc(F60, height, length=None, period=None, gen=0.0, beach=0.0, B99=2, B101=0, B102=1.0, angle=0.0,
              B120=-90.0, W50=0.0, B98=2):
    """A reef_wave_desc of linear waves (B92 = 2) of the given height and either length
    (B91) or period (B93) over still water at F60; gen / beach: the zone half widths B96."""
    if (length is None) == (period is None):
        raise ValueError("give either the wave length or the period")
    return WaveDesc(2, 1 if length is not None else 0, height, length or 0.0, 1 if period is not None else 0, height,
                    period or 0.0, gen, beach, B98, B99, B101, B102, angle, B120, F60, W50)


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_double), ("max_live", ctypes.c_int), ("precision", ctypes.c_int),
                ("delta", ctypes.c_double), ("graphs", ctypes.c_int), ("F49", ctypes.c_int), ("shared", Params),
                ("plans", ctypes.POINTER(ctypes.c_int)), ("nplans", ctypes.c_int)]


def library():
    """Loads libreef3d_hip.so; raises if it was not built (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(_LIBPATH):
        raise ImportError(f"{_LIBPATH} missing: build it with `python reef3d_amd/build.py`")
    L = ctypes.CDLL(_LIBPATH)
    dp, ip, vp = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p
    sig = {
        "reef_ctx_create": ([ctypes.POINTER(GridDesc), ctypes.POINTER(Params), ctypes.POINTER(vp)], ctypes.c_int),
        "reef_ctx_destroy": ([vp], None),
        "reef_set_control": ([vp, ctypes.c_char_p, ctypes.c_double], ctypes.c_int),
        "reef_last_error": ([], ctypes.c_char_p),
        "reef_case_last_error": ([], ctypes.c_char_p),
        "reef_grid_read": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(GridDesc)], ctypes.c_int),
        "reef_grid_file_free": ([vp], None),
        "reef_ctrl_read": ([ctypes.c_char_p, ctypes.POINTER(Params), dp], ctypes.c_int),
        "reef_ctrl_key": ([ctypes.c_char_p, ctypes.c_char, ctypes.c_int, dp, ctypes.c_int, ip], ctypes.c_int),
        "reef_field_put": ([vp, ctypes.c_int, dp, ctypes.c_int], ctypes.c_int),
        "reef_field_get": ([vp, ctypes.c_int, dp, ctypes.c_int], ctypes.c_int),
        "reef_field_zero": ([vp, ctypes.c_int], ctypes.c_int),
        "reef_ghost_fill": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "reef_convection": ([vp] + [ctypes.c_int] * 5, ctypes.c_int),
        "reef_reini": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "reef_reini_ini": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "reef_fluid_update": ([vp, dp], ctypes.c_int),
        "reef_fluid_update_noalloc": ([vp, dp], ctypes.c_int),
        "reef_poisson_solve": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ip, dp], ctypes.c_int),
        "reef_poisson_solve_mixed": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ip, ip, dp],
                                     ctypes.c_int),
        "reef_poisson_residual": ([vp, ctypes.c_int, ctypes.c_int, dp], ctypes.c_int),
        "reef_solver_start": ([vp, ctypes.c_int, ctypes.POINTER(dp), dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                               ip, dp], ctypes.c_int),
        "reef_solver_veclen": ([vp, ctypes.c_int], ctypes.c_long),
        "reef_solver_config": ([vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "reef_solver_last": ([vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long), dp], ctypes.c_int),
        "reef_pjm_corr": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ip], ctypes.c_int),
        "reef_fc3_step": ([vp, ip, dp], ctypes.c_int),
        "reef_vortex_step": ([vp, ctypes.c_double, dp], ctypes.c_int),
        "reef_timestep": ([vp, ctypes.c_double, ctypes.c_int, dp], ctypes.c_int),
        "reef_set_timestep": ([vp, ctypes.c_double], ctypes.c_int),
        "reef_checkpoint_write": ([vp, ctypes.c_char_p], ctypes.c_int),
        "reef_checkpoint_read": ([vp, ctypes.c_char_p], ctypes.c_int),
        "reef_cfd_state_write": ([vp, ctypes.c_char_p, dp], ctypes.c_int),
        "reef_cfd_state_read": ([vp, ctypes.c_char_p, dp], ctypes.c_int),
        "reef_cfd_state_write_host": ([ctypes.c_int] * 3 + [ctypes.c_char_p, ctypes.POINTER(dp), dp], ctypes.c_int),
        "reef_cfd_state_read_host": ([ctypes.c_int] * 3 + [ctypes.c_char_p, ctypes.POINTER(dp), dp], ctypes.c_int),
        "reef_phi_rk3": ([vp], ctypes.c_int),
        "reef_sync": ([vp], ctypes.c_int),
        "reef_timing_enable": ([vp, ctypes.c_int], ctypes.c_int),
        "reef_timing_get": ([vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "reef_timing_reset": ([vp], ctypes.c_int),
        "reef_ctx_sizes": ([vp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "reef_batch_create": ([ctypes.c_double, ctypes.POINTER(vp)], ctypes.c_int),
        "reef_batch_destroy": ([vp], None),
        "reef_batch_tank_bytes": ([ctypes.POINTER(GridDesc)], ctypes.c_long),
        "reef_batch_add": ([vp, ctypes.POINTER(GridDesc), ctypes.POINTER(Params), ip], ctypes.c_int),
        "reef_batch_remove": ([vp, ctypes.c_int], ctypes.c_int),
        "reef_batch_wave_set": ([vp, ctypes.c_int, ctypes.POINTER(WaveDesc), dp], ctypes.c_int),
        "reef_wave_set": ([vp, ctypes.POINTER(WaveDesc), dp], ctypes.c_int),
        "reef_wave_relax": ([vp, ctypes.c_double] + [ctypes.c_int] * 5, ctypes.c_int),
        "reef_batch_set_active": ([vp, ip, ctypes.c_int], ctypes.c_int),
        "reef_batch_info": ([vp, ip, ip, dp], ctypes.c_int),
        "reef_batch_put": ([vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int], ctypes.c_int),
        "reef_batch_get": ([vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int], ctypes.c_int),
        "reef_batch_ghost_fill": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "reef_batch_fluid_update": ([vp], ctypes.c_int),
        "reef_batch_control": ([vp, ctypes.c_char_p, ctypes.c_double], ctypes.c_int),
        "reef_batch_solver_config": ([vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "reef_batch_fc3_step": ([vp, ctypes.POINTER(ctypes.c_long), dp], ctypes.c_int),
        "reef_batch_autotune": ([vp, ctypes.c_int, ip], ctypes.c_int),
        "reef_batch_plan_set": ([vp] + [ctypes.c_int] * 5, ctypes.c_int),
        "reef_batch_plans": ([vp, ip, ctypes.c_int, ip], ctypes.c_int),
        "reef_batch_sync": ([vp], ctypes.c_int),
        "reef_batch_run": ([vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "reef_batch_timestep": ([vp, ctypes.c_double, ctypes.c_int, dp], ctypes.c_int),
        "reef_batch_timing_enable": ([vp, ctypes.c_int], ctypes.c_int),
        "reef_batch_timing_get": ([vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "reef_server_create": ([ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)], ctypes.c_int),
        "reef_server_destroy": ([vp], None),
        "reef_server_last_error": ([], ctypes.c_char_p),
        "reef_server_submit": ([vp, ctypes.POINTER(GridDesc), ctypes.c_double, ctypes.c_int] + [dp] * 5
                               + [ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "reef_server_wait": ([vp, ctypes.c_long, ctypes.c_double, ip], ctypes.c_int),
        "reef_server_result": ([vp, ctypes.c_long, ctypes.c_int, dp], ctypes.c_int),
        "reef_server_request_info": ([vp, ctypes.c_long, dp, ip, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "reef_server_release": ([vp, ctypes.c_long], ctypes.c_int),
        "reef_server_stats": ([vp, ip, ip, ctypes.POINTER(ctypes.c_long), dp], ctypes.c_int),
        "reef_server_peak_bytes": ([vp, dp], ctypes.c_int),
        "reef_host_grid_arrays": ([ctypes.POINTER(GridDesc), ctypes.c_double, ctypes.c_int] + [dp] * 7, ctypes.c_int),
        "reef_host_ghost_apply": ([ctypes.POINTER(GridDesc), ctypes.c_double, ctypes.c_int, ctypes.c_int, dp],
                                  ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        msg = library().reef_last_error().decode(errors="replace")
        raise ReefError(f"reef call failed ({rc}): {msg}")


def _case_check(rc):
    if rc != 0:
        msg = library().reef_case_last_error().decode(errors="replace")
        raise ReefError(f"reef case file ({rc}): {msg}")


def read_grid_file(path):
    """lexer::read_grid of a DIVEMesh grid file (reef_grid_read): a dict with nx, ny, nz,
    jdir, XN, YN, ZN, DYM, DZM and group (boundary group of sides 1..6)."""
    L = library()
    h, gd = ctypes.c_void_p(), GridDesc()
    _case_check(L.reef_grid_read(os.fsencode(path), ctypes.byref(h), ctypes.byref(gd)))
    try:
        out = dict(nx=gd.knox, ny=gd.knoy, nz=gd.knoz, jdir=gd.j_dir, DYM=gd.DYM, DZM=gd.DZM,
                   group=[gd.group[q] for q in range(7)])
        for key, ptr, n in (("XN", gd.XN, gd.knox), ("YN", gd.YN, gd.knoy), ("ZN", gd.ZN, gd.knoz)):
            out[key] = np.ctypeslib.as_array(ptr, shape=(n + 11,)).copy()
    finally:
        L.reef_grid_file_free(h)
    return out


def read_ctrl_file(path):
    """The hot-path keys of a ctrl.txt (reef_ctrl_read): (Params, dict F49, N48, N47, N40, D20)."""
    prm, c5 = Params(), np.zeros(5)
    _case_check(library().reef_ctrl_read(os.fsencode(path), ctypes.byref(prm), _dp(c5)))
    return prm, dict(F49=int(c5[0]), N48=int(c5[1]), N47=float(c5[2]), N40=int(c5[3]), D20=int(c5[4]))


def read_ctrl_key(path, key):
    """The values of one ctrl.txt key such as "N 43" or "B 96" (reef_ctrl_key): a list,
    empty when the file does not set the key."""
    letter, num = key.split()
    vals, n = np.zeros(16), ctypes.c_int(0)
    _case_check(library().reef_ctrl_key(os.fsencode(path), letter.encode(), int(num), _dp(vals), 16, ctypes.byref(n)))
    return [float(x) for x in vals[:min(n.value, 16)]]


def cfd_state_name(num, rank=1):
    """File name of cfd_state's result file (cfd_state_filename.cpp:27)."""
    return os.path.join("REEF3D_CFD_STATE", "REEF3D-CFD-State-%08i-%06i.r3d" % (num, rank))


def _state_meta(count, printcount, times):
    meta = np.zeros(8)
    meta[0], meta[1] = int(count), int(printcount)
    meta[2:8] = np.asarray(times, dtype=np.float64).reshape(6)
    return meta


def write_cfd_state_host(path, fields, count=0, printcount=0, times=(0.0,) * 6):
    """cfd_state::write_result from host fields (reef_cfd_state_write_host): fields = u, v,
    w, press, phi in the i-fast layout (nz + 6, ny + 6, nx + 6); times = simtime, printtime,
    sedprinttime, fsfprinttime, probeprinttime, stateprinttime."""
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    nz, ny, nx = (d - 6 for d in fs[0].shape)
    arr = (ctypes.POINTER(ctypes.c_double) * 5)(*[_dp(f) for f in fs])
    _case_check(library().reef_cfd_state_write_host(nx, ny, nz, os.fsencode(path), arr,
                                                    _dp(_state_meta(count, printcount, times))))


def read_cfd_state_host(path, nx, ny, nz):
    """cfd_state::read into zeroed host fields (reef_cfd_state_read_host): (u, v, w, press,
    phi), count, printcount, times."""
    fs = [np.zeros((nz + 6, ny + 6, nx + 6)) for _ in range(5)]
    meta = np.zeros(8)
    arr = (ctypes.POINTER(ctypes.c_double) * 5)(*[_dp(f) for f in fs])
    _case_check(library().reef_cfd_state_read_host(nx, ny, nz, os.fsencode(path), arr, _dp(meta)))
    return tuple(fs), int(meta[0]), int(meta[1]), tuple(float(x) for x in meta[2:])


def grid_desc(nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir=1, top_group=21):
    """reef_grid_desc of a box; the node arrays are returned too (they must stay alive)."""
    xn = [np.ascontiguousarray(a, dtype=np.float64) for a in (XN, YN, ZN)]
    gd = GridDesc(nx, ny, nz, jdir, _dp(xn[0]), _dp(xn[1]), _dp(xn[2]), DYM, DZM,
                  (ctypes.c_int * 7)(0, 21, 21, 21, 21, 21, top_group))
    return gd, xn


def host_grid_arrays(gd, axis, n, F45=2.1):
    """Grid arrays, WENO tables and (DXM, DYM, DZM, psi) the library builds on the host."""
    P, DN, DP = np.zeros(n + 10), np.zeros(n + 10), np.zeros(n + 10)
    qf, cf, isf = np.zeros((n + 2, 2, 6, 2)), np.zeros((n + 2, 2, 6)), np.zeros((n + 2, 2, 6, 3))
    sc = np.zeros(4)
    _check(library().reef_host_grid_arrays(ctypes.byref(gd), F45, axis, _dp(P), _dp(DN), _dp(DP), _dp(qf), _dp(cf),
                                           _dp(isf), _dp(sc)))
    return P, DN, DP, qf, cf, isf, sc


def host_ghost_apply(gd, kind, gcv, field, F45=2.1):
    """Applies the library's compiled ghost program to a host field (i-fast layout)."""
    f = np.ascontiguousarray(field, dtype=np.float64).copy()
    _check(library().reef_host_ghost_apply(ctypes.byref(gd), F45, kind, gcv, _dp(f)))
    return f


class Domain:
    """One device-resident box domain (lexer + fdm + ghostcell of the reference).

    nx, ny, nz: interior cells; XN/YN/ZN: node coordinates with 5 ghost nodes per side;
    top_group: boundary group of the +z side (21 wall, 3 free surface); the other five
    sides are walls.  Physical/control values default to the reference's ini.cpp
    defaults except where the hot-path configs set them.
    """

    def __init__(self, nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir=1, top_group=21,
                 rho=(1000.0, 1.205), visc=(1.0e-6, 1.4938e-5), g=(0.0, 0.0, 0.0),
                 F43=0.55, F45=2.1, dt=0.01, N44=1e-10, N46=5000, F49=1, D20=0):
        L = library()
        self.nx, self.ny, self.nz, self.jdir = nx, ny, nz, jdir
        self.shape = (nz + 6, ny + 6, nx + 6)
        self._xn = [np.ascontiguousarray(a, dtype=np.float64) for a in (XN, YN, ZN)]
        for a, n in zip(self._xn, (nx, ny, nz)):
            if a.size != n + 11:
                raise ValueError("node arrays need n + 11 entries (5 ghost nodes per side)")
        gd = GridDesc(nx, ny, nz, jdir, _dp(self._xn[0]), _dp(self._xn[1]), _dp(self._xn[2]),
                      DYM, DZM, (ctypes.c_int * 7)(0, 21, 21, 21, 21, 21, top_group))
        self.params = Params(rho[0], visc[0], rho[1], visc[1], g[0], g[1], g[2], F43, F45, dt, N44, N46)
        h = ctypes.c_void_p()
        _check(L.reef_ctx_create(ctypes.byref(gd), ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h
        if F49 != 1:
            self.set_control("F49", F49)
        if D20 != 0:
            self.set_control("D20", D20)

    @classmethod
    def from_case(cls, case_dir):
        """A domain from a REEF3D case directory: DIVEMesh_Grid/grid-000001.dat and ctrl.txt
        (the hot-path keys; N48 = 1 turns on the adaptive time step of every step)."""
        g = read_grid_file(os.path.join(case_dir, "DIVEMesh_Grid", "grid-000001.dat"))
        prm, ctrl = read_ctrl_file(os.path.join(case_dir, "ctrl.txt"))
        if g["group"][1:6] != [21] * 5:
            raise ReefError(f"reef case: sides 1..5 must be walls (group 21), got {g['group'][1:6]}")
        D = cls(g["nx"], g["ny"], g["nz"], g["XN"], g["YN"], g["ZN"], g["DYM"], g["DZM"], jdir=g["jdir"],
                top_group=g["group"][6], rho=(prm.W1, prm.W3), visc=(prm.W2, prm.W4),
                g=(prm.W20, prm.W21, prm.W22), F43=prm.F43, F45=prm.F45, dt=prm.dt, N44=prm.N44, N46=prm.N46,
                F49=ctrl["F49"], D20=ctrl["D20"])
        if ctrl["D20"] == 2:
            n43 = read_ctrl_key(os.path.join(case_dir, "ctrl.txt"), "N 43")
            D.set_control("N43", n43[0] if n43 else 1.0e-5)   # ini.cpp:589
        if ctrl["N48"]:
            D.set_control("N47", ctrl["N47"])
            D.set_control("N48", 1)
        D.ctrl = ctrl
        return D

    def set_control(self, key, value):
        """lexer control key of a hot-path variant (reef_set_control), e.g. F49 = 0."""
        _check(library().reef_set_control(self.h, key.encode(), float(value)))

    def close(self):
        if getattr(self, "h", None):
            library().reef_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- fields (fdm)
    def put(self, name, arr, layout="ifast"):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != int(np.prod(self.shape)):
            raise ValueError(f"field {name}: expected {self.shape} values")
        _check(library().reef_field_put(self.h, FIELDS[name], _dp(a), 0 if layout == "ifast" else 1))

    def get(self, name, layout="ifast"):
        out = np.empty(self.shape if layout == "ifast" else (self.nx + 6, self.ny + 6, self.nz + 6))
        _check(library().reef_field_get(self.h, FIELDS[name], _dp(out), 0 if layout == "ifast" else 1))
        return out

    def zero(self, name):
        _check(library().reef_field_zero(self.h, FIELDS[name]))

    def sync(self):
        _check(library().reef_sync(self.h))

    # ---------------------------------------------------------------- ghostcell
    def start1(self, name, gcv):
        _check(library().reef_ghost_fill(self.h, FIELDS[name], 1, gcv))

    def start2(self, name, gcv):
        _check(library().reef_ghost_fill(self.h, FIELDS[name], 2, gcv))

    def start3(self, name, gcv):
        _check(library().reef_ghost_fill(self.h, FIELDS[name], 3, gcv))

    def start4(self, name, gcv):
        _check(library().reef_ghost_fill(self.h, FIELDS[name], 4, gcv))

    # ---------------------------------------------------------------- convection / reini
    def convection(self, q, ipol, u="u", v="v", w="w"):
        """convection::start(p, a, q, ipol, u, v, w): += into F/G/H (ipol 1-3) or L (4)."""
        _check(library().reef_convection(self.h, ipol, FIELDS[q], FIELDS[u], FIELDS[v], FIELDS[w]))

    def reini(self, name="phi", iters=3):
        """reini::start with p->count > 0 and p->F44 = iters."""
        _check(library().reef_reini(self.h, FIELDS[name], iters))

    def reini_ini(self, name="phi", iters=None, gcv=50, maxlength=None, F43=0.55, DXM=None):
        """reini::start with p->count == 0: start4(phi, gcv), then 2*int(maxlength/(F43*DXM))
        pseudo steps (or iters) with the ghost rule gcv (reini_RK3.cpp:88-101)."""
        if iters is None:
            iters = 2 * int(maxlength / (F43 * DXM))
        _check(library().reef_reini_ini(self.h, FIELDS[name], iters, gcv))

    def fluid_update(self):
        vol = np.zeros(2)
        _check(library().reef_fluid_update(self.h, _dp(vol)))
        return vol

    # ---------------------------------------------------------------- ioflow = iowave
    def wave_set(self, desc):
        """Relaxation-zone wave generation and beach (reef_wave_set) from a WaveDesc
        (wave_desc(...)); returns (wT, wL, wk, ww, wA)."""
        out = np.zeros(5)
        _check(library().reef_wave_set(self.h, ctypes.byref(desc), _dp(out)))
        return tuple(float(x) for x in out)

    def wave_relax(self, simtime, u=None, v=None, w=None, press=None, f=None):
        """wavegen_precalc at simtime and iowave's relaxations of the named fields."""
        fid = lambda n: -1 if n is None else FIELDS[n]
        _check(library().reef_wave_relax(self.h, float(simtime), fid(u), fid(v), fid(w), fid(press), fid(f)))

    def fluid_update_noalloc(self):
        """fluid_update with no device allocation (reef_fluid_update_noalloc): the reduction
        memory is the library's static device scratch; volume1, volume2 returned."""
        vol = np.zeros(2)
        _check(library().reef_fluid_update_noalloc(self.h, _dp(vol)))
        return vol

    # ---------------------------------------------------------------- projection
    def solver(self, x, rhs, stop, maxiter):
        """solver::start(var=5) on the operator assembled from phi; returns (iters, residual)."""
        it, res = ctypes.c_int(), np.zeros(1)
        _check(library().reef_poisson_solve(self.h, FIELDS[x], FIELDS[rhs], stop, maxiter,
                                            ctypes.byref(it), _dp(res)))
        return it.value, float(res[0])

    def solver_mixed(self, x, rhs, stop, maxiter):
        """Mixed-precision solve (FP32 BiCGStab with FP64 reliable updates) of the same
        system; returns (iterations, reliable updates, final relative FP64 residual)."""
        it, nru, rel = ctypes.c_int(), ctypes.c_int(), np.zeros(1)
        _check(library().reef_poisson_solve_mixed(self.h, FIELDS[x], FIELDS[rhs], stop, maxiter,
                                                  ctypes.byref(it), ctypes.byref(nru), _dp(rel)))
        return it.value, nru.value, float(rel[0])

    def solver_start(self, var, M7, rhs, x, stop, maxiter):
        """solver::start(p, a, pgc, f, rhsvec, var) on the caller's matrix_diag: M7 rows
        (M.p, M.n, M.s, M.w, M.e, M.t, M.b) and rhs in vec order (the FLEXLOOP cells of var,
        i outermost); field x holds the initial guess and receives the solution.  Returns
        (iterations, final ||r||_2 / cells)."""
        n = library().reef_solver_veclen(self.h, int(var))
        rows = [np.ascontiguousarray(M7[q], dtype=np.float64) for q in range(7)]
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        if any(r.shape != (n,) for r in rows) or b.shape != (n,):
            raise ValueError(f"operator rows and rhs must have {n} entries for var {var}")
        ptrs = (ctypes.POINTER(ctypes.c_double) * 7)(*(_dp(r) for r in rows))
        it, res = ctypes.c_int(), np.zeros(1)
        _check(library().reef_solver_start(self.h, int(var), ptrs, _dp(b), FIELDS[x], float(stop), int(maxiter),
                                           ctypes.byref(it), _dp(res)))
        return it.value, float(res[0])

    def residual(self, x, rhs):
        """||rhs - A x||_2 in FP64 on the device (operator assembled from phi)."""
        out = np.zeros(1)
        _check(library().reef_poisson_residual(self.h, FIELDS[x], FIELDS[rhs], _dp(out)))
        return float(out[0])

    def solver_config(self, precision="fp64", delta=0.1):
        """Solver used by pressure() and momentum_fc3(): "fp64" or "mixed"."""
        prec = {"fp64": 0, "mixed": 1}[precision]
        _check(library().reef_solver_config(self.h, prec, delta))

    def solver_last(self):
        """(iterations, reliable updates, relative residual) of the last mixed solve."""
        a, b, r = ctypes.c_long(), ctypes.c_long(), np.zeros(1)
        _check(library().reef_solver_last(self.h, ctypes.byref(a), ctypes.byref(b), _dp(r)))
        return a.value, b.value, float(r[0])

    def pressure(self, u="u", v="v", w="w", alpha=1.0):
        """pressure::start with pjm_corr; returns the BiCGStab iteration count."""
        it = ctypes.c_int()
        _check(library().reef_pjm_corr(self.h, FIELDS[u], FIELDS[v], FIELDS[w], alpha, ctypes.byref(it)))
        return it.value

    # ---------------------------------------------------------------- time steps
    def momentum_fc3(self):
        """momentum_FC3::start: one two-phase time step; returns (iterations, volumes)."""
        it, vol = ctypes.c_int(), np.zeros(2)
        _check(library().reef_fc3_step(self.h, ctypes.byref(it), _dp(vol)))
        return it.value, vol

    def timestep(self, N47=0.3, ini=False):
        """ietimestep::start (ini=False) / ::ini: sets the next step's dt on the device;
        returns (dt, umax, vmax, wmax)."""
        out = np.zeros(4)
        _check(library().reef_timestep(self.h, float(N47), 1 if ini else 0, _dp(out)))
        return tuple(float(x) for x in out)

    def checkpoint(self, path):
        """FP64 checkpoint of every public field and the step size (reef_checkpoint_write)."""
        _check(library().reef_checkpoint_write(self.h, os.fsencode(path)))

    def restore(self, path):
        """Restores a checkpoint of a domain with the same grid (reef_checkpoint_read)."""
        _check(library().reef_checkpoint_read(self.h, os.fsencode(path)))

    def write_cfd_state(self, path, count=0, printcount=0, times=(0.0,) * 6):
        """The reference's FP32 state file of u, v, w, press, phi (reef_cfd_state_write);
        times = simtime, printtime, sedprinttime, fsfprinttime, probeprinttime,
        stateprinttime."""
        _check(library().reef_cfd_state_write(self.h, os.fsencode(path), _dp(_state_meta(count, printcount, times))))

    def read_cfd_state(self, path):
        """Loads a state file (reef_cfd_state_read): u, v, w, press, phi with their ghost
        layers refilled; returns (count, printcount, times)."""
        meta = np.zeros(8)
        _check(library().reef_cfd_state_read(self.h, os.fsencode(path), _dp(meta)))
        return int(meta[0]), int(meta[1]), tuple(float(x) for x in meta[2:])

    def set_timestep(self, dt):
        """dt of the following steps (p->dt)."""
        _check(library().reef_set_timestep(self.h, float(dt)))

    def vortex_step(self, simtime, volumes=True):
        """levelset_RK3 + reinit + benchmark_vortex3D step; volumes=False skips the
        volume read-back (and the stream synchronisation it needs)."""
        if not volumes:
            _check(library().reef_vortex_step(self.h, simtime, None))
            return None
        vol = np.zeros(2)
        _check(library().reef_vortex_step(self.h, simtime, _dp(vol)))
        return vol

    def levelset_rk3(self):
        _check(library().reef_phi_rk3(self.h))

    # ---------------------------------------------------------------- timing
    def timing(self, on=True):
        _check(library().reef_timing_enable(self.h, 1 if on else 0))

    def timing_reset(self):
        _check(library().reef_timing_reset(self.h))

    def timing_get(self, name):
        ms, n = np.zeros(1), ctypes.c_long()
        _check(library().reef_timing_get(self.h, name.encode(), _dp(ms), ctypes.byref(n)))
        return float(ms[0]), n.value

    def sizes(self):
        out = (ctypes.c_long * 5)()
        _check(library().reef_ctx_sizes(self.h, out))
        return list(out)


class Batch:
    """Many independent tanks advanced by the same launches (reef_batch).

    Tanks are added with the same arguments as Domain; each keeps its own fields, grid and
    time step.  fc3_step() advances every active tank by one momentum_FC3 step.  A tank's
    results do not depend on the other tanks of the batch.
    """

    def __init__(self, hbm_budget_bytes=0.0):
        h = ctypes.c_void_p()
        _check(library().reef_batch_create(float(hbm_budget_bytes), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            library().reef_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def tank_bytes(nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir=1, top_group=21):
        gd, keep = grid_desc(nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir, top_group)
        return library().reef_batch_tank_bytes(ctypes.byref(gd))

    def add(self, nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir=1, top_group=21, rho=(1000.0, 1.205),
            visc=(1.0e-6, 1.4938e-5), g=(0.0, 0.0, 0.0), F43=0.55, F45=2.1, dt=0.01, N44=1e-10, N46=5000):
        """Adds a tank; returns its id (slot)."""
        gd, keep = grid_desc(nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir, top_group)
        prm = Params(rho[0], visc[0], rho[1], visc[1], g[0], g[1], g[2], F43, F45, dt, N44, N46)
        tid = ctypes.c_int()
        _check(library().reef_batch_add(self.h, ctypes.byref(gd), ctypes.byref(prm), ctypes.byref(tid)))
        self._keep[tid.value] = ((nz + 6, ny + 6, nx + 6), keep)
        return tid.value

    def wave_set(self, tank, desc):
        """reef_batch_wave_set: a wave generator on one tank; returns (wT, wL, wk, ww, wA)."""
        out = np.zeros(5)
        _check(library().reef_batch_wave_set(self.h, tank, ctypes.byref(desc), _dp(out)))
        return tuple(float(x) for x in out)

    def remove(self, tank):
        _check(library().reef_batch_remove(self.h, tank))
        self._keep.pop(tank, None)

    def set_active(self, tanks=()):
        arr = (ctypes.c_int * max(len(tanks), 1))(*tanks)
        _check(library().reef_batch_set_active(self.h, arr, len(tanks)))

    def info(self):
        slots, active, used = ctypes.c_int(), ctypes.c_int(), np.zeros(1)
        _check(library().reef_batch_info(self.h, ctypes.byref(slots), ctypes.byref(active), _dp(used)))
        return slots.value, active.value, float(used[0])

    def put(self, tank, name, arr, layout="ifast"):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(library().reef_batch_put(self.h, tank, FIELDS[name], _dp(a), 0 if layout == "ifast" else 1))

    def get(self, tank, name, layout="ifast"):
        out = np.zeros(self._keep[tank][0])
        _check(library().reef_batch_get(self.h, tank, FIELDS[name], _dp(out), 0 if layout == "ifast" else 1))
        return out

    def ghost_fill(self, name, kind, gcv):
        _check(library().reef_batch_ghost_fill(self.h, FIELDS[name], kind, gcv))

    def fluid_update(self):
        _check(library().reef_batch_fluid_update(self.h))

    def set_control(self, key, value):
        _check(library().reef_batch_control(self.h, key.encode(), float(value)))

    def timestep(self, N47=0.3, ini=False):
        """ietimestep for every active tank: {slot: (dt, umax, vmax, wmax)}."""
        slots = self.info()[0]
        out = np.zeros(4 * max(slots, 1))
        _check(library().reef_batch_timestep(self.h, float(N47), 1 if ini else 0, _dp(out)))
        return {q: tuple(float(x) for x in out[4 * q:4 * q + 4]) for q in range(slots)}

    def solver_config(self, precision="fp64", delta=0.1):
        _check(library().reef_batch_solver_config(self.h, {"fp64": 0, "mixed": 1}[precision], delta))

    def autotune(self, reps=2):
        """Tunes the tile plan of each size bucket of the active tanks that has none;
        returns the number of buckets tuned."""
        n = ctypes.c_int()
        _check(library().reef_batch_autotune(self.h, int(reps), ctypes.byref(n)))
        return n.value

    def plans(self):
        """The plan table: list of (jdir, size_log2, ty, tz, cpt)."""
        n = ctypes.c_int()
        _check(library().reef_batch_plans(self.h, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_int * max(5 * n.value, 1))()
        _check(library().reef_batch_plans(self.h, buf, n.value, ctypes.byref(n)))
        return [tuple(buf[5 * q:5 * q + 5]) for q in range(n.value)]

    def set_plan(self, jdir, size_log2, ty, tz, cpt):
        _check(library().reef_batch_plan_set(self.h, jdir, size_log2, ty, tz, cpt))

    def fc3_step(self):
        """One momentum_FC3 step of every active tank; returns (iterations, volumes) per slot."""
        slots = self.info()[0]
        it = (ctypes.c_long * max(slots, 1))()
        vol = np.zeros(2 * max(slots, 1))
        _check(library().reef_batch_fc3_step(self.h, it, _dp(vol)))
        return list(it)[:slots], vol[:2 * slots].reshape(-1, 2)

    def run(self, nsteps):
        """nsteps momentum_FC3 steps of every active tank, iteration-level scheduled;
        returns the iterations of each slot's last step."""
        slots = self.info()[0]
        it = (ctypes.c_long * max(slots, 1))()
        _check(library().reef_batch_run(self.h, int(nsteps), it))
        return list(it)[:slots]

    def sync(self):
        _check(library().reef_batch_sync(self.h))

    def timing(self, on=True):
        _check(library().reef_batch_timing_enable(self.h, 1 if on else 0))

    def timing_get(self, name):
        ms, n = np.zeros(1), ctypes.c_long()
        _check(library().reef_batch_timing_get(self.h, name.encode(), _dp(ms), ctypes.byref(n)))
        return float(ms[0]), n.value


def _scheck(rc):
    if rc != 0:
        msg = library().reef_server_last_error().decode(errors="replace")
        raise ReefError(f"reef server call failed ({rc}): {msg}")


class Server:
    """Continuous-batching server of tank requests (reef_server, include/reef3d_serve.h).

    A worker thread in the library admits requests in arrival order by the HBM budget,
    advances every live tank one momentum_FC3 step per round in the same launches and
    retires finished tanks.  All requests share the physical and solver parameters given
    here; each has its own grid, dt and step count.
    """

    STATES = ("queued", "live", "done", "failed")

    def __init__(self, hbm_budget_bytes=0.0, max_live=0, precision="mixed", delta=0.1, graphs=False, F49=1,
                 rho=(1000.0, 1.205), visc=(1.0e-6, 1.4938e-5), g=(0.0, 0.0, -9.81), F43=0.55, F45=2.1,
                 N44=-1e-10, N46=20000, plans=()):
        flat = [int(x) for p in plans for x in p]
        self._plans = (ctypes.c_int * max(len(flat), 1))(*flat)
        cfg = ServerConfig(float(hbm_budget_bytes), int(max_live), {"fp64": 0, "mixed": 1}[precision], float(delta),
                           1 if graphs else 0, int(F49),
                           Params(rho[0], visc[0], rho[1], visc[1], g[0], g[1], g[2], F43, F45, 0.0, N44, N46),
                           self._plans, len(plans))
        h = ctypes.c_void_p()
        _scheck(library().reef_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._shapes = {}

    def close(self):
        if getattr(self, "h", None):
            library().reef_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, nx, ny, nz, XN, YN, ZN, DYM, DZM, dt, steps, phi, u, v, w, press, jdir=1, top_group=21):
        """Queues a tank; returns the request id.  The arrays are copied."""
        gd, keep = grid_desc(nx, ny, nz, XN, YN, ZN, DYM, DZM, jdir, top_group)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi, u, v, w, press)]
        shape = (nz + 6, ny + 6, nx + 6)
        for a in arrs:
            if a.shape != shape:
                raise ValueError(f"field shape {a.shape}, expected {shape}")
        rid = ctypes.c_long()
        _scheck(library().reef_server_submit(self.h, ctypes.byref(gd), float(dt), int(steps), *(_dp(a) for a in arrs),
                                             ctypes.byref(rid)))
        self._shapes[rid.value] = shape
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """Waits for a request; returns its state ("queued", "live", "done", "failed")."""
        st = ctypes.c_int()
        _scheck(library().reef_server_wait(self.h, rid, float(timeout), ctypes.byref(st)))
        return self.STATES[st.value]

    def result(self, rid, name):
        out = np.zeros(self._shapes[rid])
        _scheck(library().reef_server_result(self.h, rid, FIELDS[name], _dp(out)))
        return out

    def info(self, rid):
        t = np.zeros(3)
        steps, iters = ctypes.c_int(), ctypes.c_long()
        _scheck(library().reef_server_request_info(self.h, rid, _dp(t), ctypes.byref(steps), ctypes.byref(iters)))
        return dict(t_submit=t[0], t_admit=t[1], t_done=t[2], steps=steps.value, iters=iters.value)

    def release(self, rid):
        _scheck(library().reef_server_release(self.h, rid))
        self._shapes.pop(rid, None)

    def stats(self):
        q, lv, st, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_long(), np.zeros(1)
        _scheck(library().reef_server_stats(self.h, ctypes.byref(q), ctypes.byref(lv), ctypes.byref(st), _dp(b)))
        pk = np.zeros(1)
        _scheck(library().reef_server_peak_bytes(self.h, _dp(pk)))
        return dict(queued=q.value, live=lv.value, rounds=st.value, bytes_in_use=float(b[0]), peak_bytes=float(pk[0]))
