"""Synthetic inputs of the hot-path workloads (no datasets exist for this path).

wave_tank(): the two-phase wave tank of the benchmark configurations: phi = signed
distance to the free surface z = d + a cos(k x) (Newton closest point, positive in
water), linear (Airy) velocities in the water at the staggered face positions, air at
rest, hydrostatic pressure integrated from the top.  vortex3d(): the level-set vortex
benchmark (benchmark_vortex3D.cpp): sphere r = 0.15 at (0.35, 0.35, 0.35) in the unit
cube, zero initial velocity.

Arrays have shape (nz+6, ny+6, nx+6) (i fastest, 3 ghost layers).
"""
import math

import numpy as np


def nodes(n, d, x0=0.0):
    """Node coordinates with 5 ghost nodes per side (read_grid.cpp:407-432)."""
    return x0 + d * np.arange(-5, n + 6, dtype=np.float64)


def airy_omega(k, depth, g=9.81):
    return math.sqrt(g * k * math.tanh(k * depth))


def surface_distance(x, z, depth, amp, k, iters=20):
    """Signed distance from (x, z) to z = depth + amp cos(k x), positive below."""
    xs = np.array(x, dtype=np.float64, copy=True)
    for _ in range(iters):
        eta = depth + amp * np.cos(k * xs)
        d1 = -amp * k * np.sin(k * xs)
        d2 = -amp * k * k * np.cos(k * xs)
        f = (xs - x) + (eta - z) * d1
        fp = 1.0 + d1 * d1 + (eta - z) * d2
        xs = xs - f / fp
    eta = depth + amp * np.cos(k * xs)
    dist = np.sqrt((xs - x) ** 2 + (eta - z) ** 2)
    below = z < depth + amp * np.cos(k * x)
    return np.where(below, dist, -dist)


def wave_tank(nx, ny, nz, dx, depth, amp, wavelength, g=9.81, rho=(1000.0, 1.205)):
    """Fields of the wave-tank workload; returns dict(XN, YN, ZN, phi, u, v, w, press)."""
    XN, YN, ZN = nodes(nx, dx), nodes(ny, dx), nodes(nz, dx)
    XP = 0.5 * (XN[:-1] + XN[1:])
    ZP = 0.5 * (ZN[:-1] + ZN[1:])
    # cell centres / faces of the 3-ghost field: index m <-> cell m-3 <-> XP[m+2]
    xc = XP[2:2 + nx + 6]
    zc = ZP[2:2 + nz + 6]
    xf = XN[3:3 + nx + 6]          # u faces: XN[i+1]
    zf = ZN[3:3 + nz + 6]          # w faces: ZN[k+1]
    k = 2.0 * math.pi / wavelength
    om = airy_omega(k, depth, g)
    shape = (nz + 6, ny + 6, nx + 6)
    X, Z = np.meshgrid(xc, zc)                        # (nz+6, nx+6)
    phi2 = surface_distance(X, Z, depth, amp, k)
    Xu, Zu = np.meshgrid(xf, zc)
    u2 = np.where(Zu <= depth + amp * np.cos(k * Xu),
                  amp * om * np.cosh(k * Zu) / math.sinh(k * depth) * np.cos(k * Xu), 0.0)
    Xw, Zw = np.meshgrid(xc, zf)
    w2 = np.where(Zw <= depth + amp * np.cos(k * Xw),
                  amp * om * np.sinh(k * Zw) / math.sinh(k * depth) * np.sin(k * Xw), 0.0)
    phi = np.broadcast_to(phi2[:, None, :], shape).copy()
    u = np.broadcast_to(u2[:, None, :], shape).copy()
    w = np.broadcast_to(w2[:, None, :], shape).copy()
    v = np.zeros(shape)
    # hydrostatic pressure from the density of the sharp interface, integrated downwards
    ro = np.where(phi2 >= 0.0, rho[0], rho[1])
    dz = np.diff(ZN)[2:2 + nz + 6]
    col = ro * g * dz[:, None]
    p2 = np.cumsum(col[::-1], axis=0)[::-1] - 0.5 * col
    press = np.broadcast_to(p2[:, None, :], shape).copy()
    return dict(XN=XN, YN=YN, ZN=ZN, phi=phi, u=u, v=v, w=w, press=press)


def vortex3d(n):
    """Level-set vortex benchmark inputs on the unit cube with n^3 cells."""
    dx = 1.0 / n
    XN = nodes(n, dx)
    XP = 0.5 * (XN[:-1] + XN[1:])
    c = XP[2:2 + n + 6]
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    phi = 0.15 - np.sqrt((X - 0.35) ** 2 + (Y - 0.35) ** 2 + (Z - 0.35) ** 2)
    return dict(XN=XN, YN=XN.copy(), ZN=XN.copy(), phi=phi)


# workload definitions of the benchmark configurations
C0 = dict(name="vortex3d_128", n=128, T=3.0, dt=1.0 / 512.0)
C1 = dict(name="wave_tank_512x64x128", nx=512, ny=64, nz=128, dx=0.02, depth=1.6, amp=0.05,
          wavelength=2.0, dt=0.005, steps=20, rtol=1e-10)


def c3_specs(n=64, seed=1003):
    """Seeded 2-D tank set of the batched workload: nx in {256..2048 step 128}, nz in
    {64..256 step 32}, dx = dz ~ U[0.01, 0.05], depth = 0.6 nz dz, amplitude ~ U[0.01, 0.1]
    depth, wavelength ~ U[1, 4] m."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        nx = int(rng.choice(np.arange(256, 2049, 128)))
        nz = int(rng.choice(np.arange(64, 257, 32)))
        dx = float(rng.uniform(0.01, 0.05))
        depth = 0.6 * nz * dx
        amp = float(rng.uniform(0.01, 0.1)) * depth
        lam = float(rng.uniform(1.0, 4.0))
        specs.append(dict(nx=nx, ny=1, nz=nz, jdir=0, dx=dx, depth=depth, amp=amp, wavelength=lam))
    return specs


def c4_requests(n, step_div=1, seed=1004, rate=200.0):
    """Seeded request stream of the served workload: Poisson arrivals (rate per second),
    70 % 2-D tanks (nx 256..2048 step 128, nz 64..256 step 32), 30 % 3-D tanks (nx
    128..256 step 32, ny 32..64 step 8, nz 64..128 step 16), dx, depth, amplitude and
    wavelength as in c3_specs, 50..500 steps each (divided by step_div, at least 1)."""
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        if rng.random() < 0.7:
            nx, ny, nz, jdir = int(rng.choice(np.arange(256, 2049, 128))), 1, int(rng.choice(np.arange(64, 257, 32))), 0
        else:
            nx, ny, nz, jdir = (int(rng.choice(np.arange(128, 257, 32))), int(rng.choice(np.arange(32, 65, 8))),
                                int(rng.choice(np.arange(64, 129, 16))), 1)
        dx = float(rng.uniform(0.01, 0.05))
        depth = 0.6 * nz * dx
        amp = float(rng.uniform(0.01, 0.1)) * depth
        lam = float(rng.uniform(1.0, 4.0))
        steps = int(rng.integers(50, 501))
        reqs.append(dict(arrival=t, nx=nx, ny=ny, nz=nz, jdir=jdir, dx=dx, depth=depth, amp=amp, wavelength=lam,
                         steps=max(1, steps // step_div), steps_full=steps))
    return reqs


def tank_fields(spec, g=9.81):
    """wave_tank fields of a c3_specs / c4_requests entry plus its fixed time step
    dt = 0.25 dx / max(sqrt(g depth), max |u|, |w| of the initial fields)."""
    f = wave_tank(spec["nx"], spec["ny"], spec["nz"], spec["dx"], spec["depth"], spec["amp"], spec["wavelength"], g)
    umax = max(float(np.abs(f["u"]).max()), float(np.abs(f["w"]).max()))
    f["dt"] = 0.25 * spec["dx"] / max(math.sqrt(g * spec["depth"]), umax)
    return f


def tank_kwargs(spec, f, rtol=1e-10, N46=20000):
    """Batch.add / Domain keyword arguments of a c3_specs / c4_requests tank."""
    return dict(nx=spec["nx"], ny=spec["ny"], nz=spec["nz"], XN=f["XN"], YN=f["YN"], ZN=f["ZN"], DYM=spec["dx"],
                DZM=spec["dx"], jdir=spec["jdir"], top_group=3, g=(0.0, 0.0, -9.81), dt=f["dt"], N44=-rtol, N46=N46)
