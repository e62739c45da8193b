"""Closed-form and invariant checks of the CPU oracle, independent of the fixtures.

They pin what the golden vectors cannot isolate: the textbook limits of the non-uniform
WENO tables on a uniform grid, exactness of the schemes on linear data, the upwind side
of the HJ reconstruction, the reinitialisation fixed point, the Poisson operator against
a dense assembly and BiCGStab against a direct solve.  Each one fails for a plausible
slip (a swapped table row, a sign, a wrong stencil offset, a transposed axis).
"""
import numpy as np
import pytest

from tests._cases import inner, uniform_nodes


def dom(nx=12, ny=9, nz=10, dx=0.1, top=21, jdir=1, **kw):
    from oracle.oracle import Domain
    return Domain(nx, ny, nz, uniform_nodes(nx, dx), uniform_nodes(ny, dx), uniform_nodes(nz, dx), dx, dx,
                  jdir=jdir, top_group=top, **kw)


def coords(O, dx=0.1):
    k, j, i = np.meshgrid(np.arange(O.nz + 6) - 3, np.arange(O.ny + 6) - 3, np.arange(O.nx + 6) - 3, indexing="ij")
    return (i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx


def test_uniform_tables_are_textbook_weno5():
    """On a uniform grid the non-uniform tables reduce to Jiang & Peng's WENO5
    (SIAM J. Sci. Comput. 21, 2000): candidate stencils v1/3 - 7v2/6 + 11v3/6,
    -v2/6 + 5v3/6 + v4/3, v3/3 + 5v4/6 - v5/6, linear weights 1/10, 6/10, 3/10, and
    IS = 13/12 (second difference)^2 + 1/4 (...)^2, which in the first differences used
    by the reference is 4/3 a^2 + 11/3 ab + 10/3 b^2 (outer) and 4/3, 5/3, 4/3 (centre)."""
    O = dom()
    for ax in range(3):
        qf, cf, isf = O.tables(ax)
        for uf in (0, 1):
            for r in range(1, len(qf) - 1):
                assert np.allclose(cf[r, uf], [0.3, 0.6, 0.1, 0.1, 0.6, 0.3], rtol=0, atol=1e-14)
                assert np.allclose(qf[r, uf].ravel(), [1 / 3, 1 / 6, 1 / 3, 1 / 6, 1 / 3, 11 / 6,
                                                       11 / 6, 1 / 3, 1 / 3, 1 / 6, 1 / 3, 1 / 6], rtol=0, atol=1e-14)
                outer, centre = [4 / 3, 11 / 3, 10 / 3], [4 / 3, 5 / 3, 4 / 3]
                assert np.allclose(isf[r, uf], [outer, centre, outer, outer, centre, outer], rtol=0, atol=1e-13)


def test_hj_gradient_is_exact_for_linear_phi():
    """Any convex combination of the candidate stencils is exact for linear data."""
    O = dom(nx=14, ny=11, nz=12)
    x, y, z = coords(O)
    phi = 0.7 * x - 0.3 * y + 1.1 * z + 0.2
    u = np.full(O.shape, 0.5)
    v = np.full(O.shape, -0.25)
    w = np.full(O.shape, 0.125)
    L = np.zeros(O.shape)
    O.hj(phi, u, v, w, L)
    expect = -(0.5 * 0.7 + (-0.25) * (-0.3) + 0.125 * 1.1)
    assert np.allclose(inner(L), expect, atol=1e-12)


def test_hj_weno_converges_at_high_order():
    """-u dphi/dx of a smooth profile: the interior L1 error falls ~h^5."""
    errs = []
    for n in (32, 64, 128):
        O = dom(nx=n, ny=4, nz=4, dx=1.0 / n)
        x, _, _ = coords(O, 1.0 / n)
        phi = np.sin(2 * np.pi * x) + 0.3 * np.cos(6 * np.pi * x)
        u = np.ones(O.shape)
        zero = np.zeros(O.shape)
        L = np.zeros(O.shape)
        O.hj(phi, u, zero, zero, L)
        xc = inner(x)[1, 1]
        exact = -(2 * np.pi * np.cos(2 * np.pi * xc) - 0.3 * 6 * np.pi * np.sin(6 * np.pi * xc))
        errs.append(np.mean(np.abs(inner(L)[1, 1] - exact)[n // 8:-n // 8]))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 4.0), orders


def test_hj_upwind_side_follows_velocity_sign():
    """A kink seen only from one side: the upwind side decides which slope is used."""
    O = dom(nx=16, ny=5, nz=5)
    x, _, _ = coords(O)
    phi = np.where(x < 0.8, x, 0.8 + 3.0 * (x - 0.8))   # slope 1 left of 0.8, 3 right
    zero = np.zeros(O.shape)
    for sgn, slope_far in ((1.0, 1.0), (-1.0, 3.0)):
        u = np.full(O.shape, sgn)
        L = np.zeros(O.shape)
        O.hj(phi, u, zero, zero, L)
        i0 = 3 + 3                     # cell i = 3 (x = 0.35), far left of the kink
        i1 = 3 + 12                    # cell i = 12 (x = 1.25), far right
        assert np.isclose(L[5, 5, i0], -sgn * 1.0, atol=1e-12)
        assert np.isclose(L[5, 5, i1], -sgn * 3.0, atol=1e-12)


def test_flux_form_conserves_and_is_exact_for_constant_velocity():
    O = dom(nx=14, ny=10, nz=12)
    x, y, z = coords(O)
    c = (0.3, -0.2, 0.4)
    u = np.full(O.shape, c[0])
    v = np.full(O.shape, c[1])
    w = np.full(O.shape, c[2])
    # q linear: flux divergence of a linear field with constant velocity = -(c . grad q)
    q = 2.0 * x + 0.5 * y - 1.5 * z
    acc = np.zeros(O.shape)
    O.flux(1, q, u, v, w, acc)
    got = inner(acc)[:, :, :O.nx - 1]
    assert np.allclose(got, -(c[0] * 2.0 + c[1] * 0.5 + c[2] * -1.5), atol=1e-11)


def test_reinit_keeps_a_plane_distance_function_inside():
    O = dom(nx=16, ny=6, nz=16, top=21)
    _, _, z = coords(O)
    phi = 0.8 - z
    before = phi.copy()
    O.ghost(phi, 4, 52)
    O.reini(phi, 3)
    # away from the walls the signed distance of a plane is a fixed point
    assert np.allclose(inner(phi)[4:-4, 1:-1, 4:-4], inner(before)[4:-4, 1:-1, 4:-4], atol=1e-12)
    # the sign of phi never flips
    assert np.all(np.sign(inner(phi)) == np.sign(inner(before)))


def test_reinit_drives_gradient_magnitude_to_one():
    O = dom(nx=20, ny=6, nz=20)
    x, _, z = coords(O)
    phi = 2.5 * (0.55 * 2.0 - np.sqrt((x - 1.0) ** 2 + (z - 1.0) ** 2))    # |grad| = 2.5
    O.ghost(phi, 4, 52)
    O.reini(phi, 40)
    g = np.gradient(inner(phi), 0.1)
    mag = np.sqrt(g[0] ** 2 + g[2] ** 2)
    band = np.abs(inner(phi)) < 0.3
    assert abs(np.median(mag[band]) - 1.0) < 0.05


def dense_operator(O, phi):
    n = O.nx * O.ny * O.nz
    M = O.poisson(phi, np.zeros(O.shape), np.zeros(n))
    A = np.zeros((n, n))
    idx = np.arange(n).reshape(O.nx, O.ny, O.nz)      # LOOP order (k fastest)
    for (i, j, k), m in np.ndenumerate(idx):
        A[m, m] = M[0, m]
        for coef, (di, dj, dk) in zip(M[1:], ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))):
            ii, jj, kk = i + di, j + dj, k + dk
            if 0 <= ii < O.nx and 0 <= jj < O.ny and 0 <= kk < O.nz:
                A[m, idx[ii, jj, kk]] = coef[m]
            else:
                assert coef[m] == 0.0
    return M, A


def test_poisson_operator_is_the_variable_density_laplacian():
    O = dom(nx=5, ny=4, nz=6, dx=0.1)
    x, y, z = coords(O)
    phi = 0.3 - z + 0.05 * x
    M, A = dense_operator(O, phi)
    # constant pressure in the interior: only the boundary-kept diagonal terms remain, so
    # row sums equal the folded-out couplings (Dirichlet-0 correction at the walls)
    assert np.all(np.diag(A) > 0) and np.all(A - np.diag(np.diag(A)) <= 0)
    assert np.all(A.sum(axis=1) >= -1e-9)
    # symmetric up to rounding on a uniform grid (rho_face is shared by both cells)
    assert np.allclose(A, A.T, rtol=1e-12, atol=0)
    # density jump: coefficient ratio water/air faces ~ 1/830
    off = -np.array([c for c in M[5] if c != 0.0])
    assert off.max() / off.min() > 500


def test_bicgstab_matches_direct_solve():
    O = dom(nx=6, ny=5, nz=7, dx=0.1)
    x, y, z = coords(O)
    phi = 0.35 - z
    M, A = dense_operator(O, phi)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0])
    xs = np.linalg.solve(A, b)
    xf = np.zeros(O.shape)
    it, res = O.bicgstab(xf, M, b.copy(), -1e-13, 2000)
    got = inner(xf).transpose(2, 1, 0).ravel()      # field (k,j,i) -> LOOP order (i,j,k)
    assert res * O.nx * O.ny * O.nz <= 1e-13 * np.linalg.norm(b) * 1.0001
    assert np.allclose(got, xs, rtol=0, atol=1e-9 * np.abs(xs).max())


def test_residual_is_rhs_minus_dense_product():
    """orc_residual: r = b - A x per cell with the assembled operator (a sign or a
    transposed neighbour fails it against the dense matrix)."""
    O = dom(nx=5, ny=6, nz=4, dx=0.1)
    x, y, z = coords(O)
    phi = 0.2 - z + 0.03 * y
    M, A = dense_operator(O, phi)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(A.shape[0])
    xf = np.zeros(O.shape)
    inner(xf)[...] = rng.standard_normal((O.nz, O.ny, O.nx))
    xv = inner(xf).transpose(2, 1, 0).ravel()
    r = np.zeros(O.shape)
    norm = O.residual(xf, M, b.copy(), r)
    want = b - A @ xv
    got = inner(r).transpose(2, 1, 0).ravel()
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(A).max() * np.abs(xv).max())
    assert abs(norm - np.linalg.norm(want)) <= 1e-12 * np.linalg.norm(want)
    assert O.residual(np.zeros(O.shape), M, b.copy()) == pytest.approx(np.linalg.norm(b), rel=1e-14)


def test_heaviside_limits_and_volume():
    O = dom(nx=8, ny=8, nz=8, dx=0.1)
    psi = O.scalars()["psi"]
    _, _, z = coords(O)
    phi = 0.4 - z
    ro, visc = np.zeros(O.shape), np.zeros(O.shape)
    vol = O.fluid_update(phi, ro, visc)
    r = inner(ro)
    pin = inner(phi)
    assert np.all(r[pin > psi] == 1000.0) and np.all(r[pin < -psi] == 1.205)
    # smoothed Heaviside is symmetric: the water volume equals the geometric one
    assert np.isclose(vol[0], 0.8 * 0.8 * 0.4, rtol=1e-12)
    assert np.isclose(vol[0] + vol[1], 0.8 ** 3, rtol=1e-12)


@pytest.mark.parametrize("gcv,expect", [(10, "orth"), (7, "noslip")])
def test_velocity_wall_ghosts(gcv, expect):
    O = dom(nx=6, ny=5, nz=5, dx=0.1)
    u = np.random.default_rng(1).standard_normal(O.shape)
    O.ghost(u, 1, gcv)
    # -x wall: the u face at the wall ghost layer is zero, the next ones extrapolate
    # linearly (2-point Lagrange, wall value 0) or are zero (no-slip)
    g1, g2, g3, f0 = u[4, 4, 2], u[4, 4, 1], u[4, 4, 0], u[4, 4, 3]
    assert g1 == 0.0
    if expect == "orth":
        assert np.isclose(g2, -f0) and np.isclose(g3, -2.0 * f0)
    else:
        assert g2 == 0.0 and g3 == 0.0
