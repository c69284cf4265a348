"""Pins the CPU oracle to the reference's known answers and to the mathematics.

Each check is chosen so that a plausible restatement mistake (a wrong index,
sign, stencil offset or upwind branch) fails it:
  - FD tables: every row of the reference's rational tables
    (src/mod_coeffs.f90:598-772) against exact Lagrange weights computed here
    with fractions;
  - coordinates, cdg: closed forms on uniform grids (src/mod_coeffs.f90:1048-1258);
  - stencils: polynomial exactness, gradient of a linear field, divergence of a
    constant field, the upwind branch on a one-point pulse;
  - Poisson: one JOR sweep against a matrix assembled from the cdg arrays,
    BiCGstab against its own residual, and the projection closure
    D(u_new) = b - (D G) p across separately restated routines.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O


def lagrange_weights(offsets, order):
    n = len(offsets)
    A = [[Fraction(x) ** p for x in offsets] + [Fraction(math.factorial(order) if p == order else 0)]
         for p in range(n)]
    for col in range(n):
        piv = next(r for r in range(col, n) if A[r][col] != 0)
        A[col], A[piv] = A[piv], A[col]
        for r in range(n):
            if r != col and A[r][col] != 0:
                f = A[r][col] / A[col][col]
                A[r] = [a - f * b for a, b in zip(A[r], A[col])]
    return [A[r][n] / A[r][r] for r in range(n)]


# rows of the reference table whose weights are not a derivative stencil
# (src/mod_coeffs.f90:766 and :769 list -225 where exact weights need -255)
KNOWN_TABLE_TYPOS = {(7, 2, -1.0), (7, 2, -5.0)}


def table_rows():
    rows = []
    for n in range(2, 8):
        for abl in range(0, 3):
            for twice in range(-2 * (n - 1) - 2, 3):
                dx = twice / 2
                t = O.exact_table(abl, n, dx)
                if t is not None:
                    rows.append((n, abl, dx, t))
    return rows


def test_table_has_all_rows():
    assert len(table_rows()) == 118


def test_tables_are_exact_lagrange_weights():
    for n, abl, dx, t in table_rows():
        offs = [Fraction(int(round(2 * dx)), 2) + q for q in range(n)]
        w = np.array([float(x) for x in lagrange_weights(offs, abl)])
        if (n, abl, dx) in KNOWN_TABLE_TYPOS:
            assert not np.array_equal(w, t)
            # the typo rows do not annihilate constants
            assert abs(t.sum() - 30.0 / 180.0) < 1e-14
        else:
            assert np.array_equal(w, t), (n, abl, dx)


@pytest.mark.parametrize("bc", [(1, 1, 1, 1, 1, 1), (-1, -1, 1, 1, -1, -1)])
@pytest.mark.parametrize("stretch", [(0, 0, 0), (0, 1.5, 1.5)])
def test_second_derivative_rows_annihilate_constants(bc, stretch):
    # a row built from a typo table row would sum to 30/180/dxi^2 != 0
    c = O.OracleCase((17, 13, 11), bc, (2, 1, 1), stretch)
    for name, N in (("cp11", 17), ("cu11", 17), ("cp22", 13), ("cv22", 13), ("cp33", 11), ("cw33", 11)):
        a = c.coeff(name).reshape(N + 1, 7)
        scale = np.abs(a).max(axis=1) + 1e-300
        assert np.all(np.abs(a.sum(axis=1)) / scale < 1e-12), name


def test_uniform_coordinates_and_mirror():
    c = O.OracleCase((17, 9, 9), (1, 1, 1, 1, -1, -1), (2.0, 1.0, 1.0))
    h = 2.0 / 16
    xp = c.coords(0, 0)
    xu = c.coords(0, 1)
    i = np.arange(-3, 17 + 4)
    np.testing.assert_allclose(xp, (i - 1) * h, atol=1e-15)
    np.testing.assert_allclose(xu, (i - 0.5) * h, atol=1e-15)
    zp = c.coords(2, 0)
    k = np.arange(-3, 9 + 4)
    np.testing.assert_allclose(zp, (k - 1) * (1.0 / 8), atol=1e-15)


def test_stretched_coordinates_symmetric_and_clustered():
    c = O.OracleCase((9, 33, 9), (1, 1, 1, 1, 1, 1), (1, 1, 1), (0, 1.5, 0))
    y = c.coords(1, 0)[4:4 + 33]
    assert y[0] == 0.0 and abs(y[-1] - 1.0) < 1e-14
    np.testing.assert_allclose(y + y[::-1], 1.0, atol=1e-14)
    dy = np.diff(y)
    # the atanh map of coord_tanh refines towards the middle of the extent
    assert dy[0] > dy[15] and np.all(dy > 0)


def test_cdg_closed_form():
    c = O.OracleCase((17, 9, 12), (1, 1, 1, 1, -1, -1), (2.0, 1.0, 1.0))
    h1 = 2.0 / 16
    g1 = c.coeff("cdg1").reshape(17, 3)
    inv = 1.0 / h1 ** 2
    np.testing.assert_allclose(g1[1:-1], np.tile([inv, -2 * inv, inv], (15, 1)), rtol=1e-13)
    np.testing.assert_allclose(g1[0], [0.0, -2 * inv, 2 * inv], rtol=1e-13)
    np.testing.assert_allclose(g1[-1], [2 * inv, -2 * inv, 0.0], rtol=1e-13)
    h3 = 1.0 / 11
    g3 = c.coeff("cdg3").reshape(12, 3)
    np.testing.assert_allclose(g3[:11], np.tile([1, -2, 1], (11, 1)) / h3 ** 2, rtol=1e-12)


def _field(c, fn_of_xyz, grid):
    """fn evaluated at (x, y, z) of the given grid ('p', 'u', 'v', 'w')."""
    xs = []
    for d in range(3):
        which = 1 if grid == "uvw"[d] else 0
        xs.append(c.coords(d, which))
    K, J, I = np.meshgrid(np.arange(c.N[2] + 7), np.arange(c.N[1] + 7), np.arange(c.N[0] + 7), indexing="ij")
    return np.ascontiguousarray(fn_of_xyz(xs[0][I], xs[1][J], xs[2][K]))


def _box(lo, hi):
    return tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_interior_stencils_exact_for_polynomials(deg):
    # uniform grid, periodic-free walls; interior stencils are 6-7 points wide
    c = O.OracleCase((25, 21, 19), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0))
    f = lambda x, y, z: x ** deg + 0.5 * y ** deg - 0.25 * z ** deg
    df = [lambda x, y, z: deg * x ** (deg - 1), lambda x, y, z: 0.5 * deg * y ** (deg - 1),
          lambda x, y, z: -0.25 * deg * z ** (deg - 1)]
    p = _field(c, f, "p")
    for m in (1, 2, 3):
        g = c.zeros()
        c.gradient(m, p.copy(), g)
        exact = _field(c, df[m - 1], "uvw"[m - 1])
        # interior, away from the one-sided wall stencils (3 points)
        lo = [c.SI[m - 1][d] + 3 for d in range(3)]
        hi = [c.NI[m - 1][d] - 3 for d in range(3)]
        s = _box(lo, hi)
        np.testing.assert_allclose(g[s], exact[s], rtol=1e-9, atol=1e-9)


def test_divergence_of_linear_velocity():
    c = O.OracleCase((17, 13, 11), (1, 1, 1, 1, 1, 1), (2, 1, 1), (0, 1.5, 1.5))
    vel = c.zeros(3)
    vel[0] = _field(c, lambda x, y, z: 2.0 * x + 1.0, "u")
    vel[1] = _field(c, lambda x, y, z: -3.0 * y, "v")
    vel[2] = _field(c, lambda x, y, z: 0.5 * z - 4.0, "w")
    div = c.zeros()
    c.divergence2(vel, div)
    s = _box(c.Sp, c.Np)
    # the wall ghosts are zero, so only points whose stencils stay inside are exact
    s_in = _box([c.Sp[d] + 2 for d in range(3)], [c.Np[d] - 2 for d in range(3)])
    np.testing.assert_allclose(div[s_in], -0.5, atol=1e-11)
    assert np.isfinite(div[s]).all()


def test_upwind_branch_on_a_pulse():
    # pp >= 0 selects cN..U (offsets -3..2): a pulse 3 points downstream of i
    # (at i+3) is invisible to it; pp < 0 selects cN..D (offsets -2..3), which sees it.
    c = O.OracleCase((25, 9, 9), (-1, -1, -1, -1, -1, -1))
    vel = c.zeros(3)
    i0, j0, k0 = 12, 4, 4
    vel[0][k0 + 3, j0 + 3, i0 + 3] = 1.0
    for sign, expect_zero in ((1.0, True), (-1.0, False)):
        work = c.zeros(3)
        work[0][:] = sign  # advection velocity u (interpolated to the u grid stays sign)
        nl = c.zeros(3)
        pp = c.zeros()
        c.nonlinear(vel, work, nl, pp)
        val = nl[0][k0 + 3, j0 + 3, i0 - 3 + 3]
        assert (val == 0.0) == expect_zero, (sign, val)


def test_helmholtz_of_quadratic_is_constant():
    c = O.OracleCase((21, 19, 17), (-1, -1, -1, -1, -1, -1), (1, 1, 1))
    vel = c.zeros(3)
    # periodic grid: use a trigonometric field with a known Laplacian
    vel[0] = _field(c, lambda x, y, z: np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y), "u")
    nl = c.zeros(3)
    c.helmholtz(1.0, vel, nl)
    lap = -8 * np.pi ** 2 * vel[0]
    s = _box(c.SI[0], c.NI[0])
    # nl = -multL * Lap(u): 6th-order accurate at 20 points per period
    np.testing.assert_allclose(nl[0][s], -lap[s], atol=2e-3 * np.abs(lap).max())


def _assemble_cdg(c):
    """Sparse cdg operator on the pressure box from the coefficient arrays."""
    import scipy.sparse as sp
    n = [c.Np[d] - c.Sp[d] + 1 for d in range(3)]
    g = [c.coeff(f"cdg{d + 1}").reshape(c.N[d], 3) for d in range(3)]
    idx = lambda i, j, k: (i - c.Sp[0]) + n[0] * ((j - c.Sp[1]) + n[1] * (k - c.Sp[2]))
    rows, cols, vals = [], [], []
    for k in range(c.Sp[2], c.Np[2] + 1):
        for j in range(c.Sp[1], c.Np[1] + 1):
            for i in range(c.Sp[0], c.Np[0] + 1):
                r = idx(i, j, k)
                diag = g[0][i - 1, 1] + g[1][j - 1, 1] + g[2][k - 1, 1]
                rows.append(r); cols.append(r); vals.append(diag)
                for d, (a, b, e) in enumerate(((i, j, k),)):
                    pass
                for d in range(3):
                    pos = [i, j, k]
                    for s, tap in ((-1, 0), (1, 2)):
                        q = list(pos)
                        q[d] += s
                        coef = g[d][pos[d] - 1, tap]
                        if coef == 0.0:
                            continue
                        if q[d] < c.Sp[d]:
                            q[d] = c.Np[d]
                        if q[d] > c.Np[d]:
                            q[d] = c.Sp[d]
                        rows.append(r); cols.append(idx(*q)); vals.append(coef)
    N = n[0] * n[1] * n[2]
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


@pytest.mark.parametrize("bc", [(1, 1, 1, 1, 1, 1), (-1, -1, 1, 1, -1, -1)])
def test_one_jor_sweep_matches_assembled_matrix(bc):
    c = O.OracleCase((11, 9, 8), bc, (1.5, 1, 1), (0, 1.3, 0))
    A = _assemble_cdg(c)
    rng = np.random.default_rng(1)
    rel = np.ascontiguousarray(rng.standard_normal(c.shape))
    bb = np.ascontiguousarray(rng.standard_normal(c.shape))
    s = _box(c.Sp, c.Np)
    x0 = rel[s].transpose(2, 1, 0).reshape(-1, order="F").copy()
    b0 = bb[s].transpose(2, 1, 0).reshape(-1, order="F").copy()
    omega = 0.8
    comp = c.zeros()
    c.jor(1, omega, c.Sp, c.Np, rel, bb, comp)
    D = A.diagonal()
    x1 = x0 + omega * (b0 - A @ x0) / D
    got = rel[s].transpose(2, 1, 0).reshape(-1, order="F")
    np.testing.assert_allclose(got, x1, rtol=1e-12, atol=1e-12 * np.abs(x1).max())


def test_bicgstab_solves_its_operator():
    c = O.OracleCase((13, 11, 9), (1, 1, 1, 1, 1, 1), (1.2, 1, 1), (0, 1.2, 1.2))
    rng = np.random.default_rng(2)
    xt = np.ascontiguousarray(rng.standard_normal(c.shape))
    bb = c.zeros()
    c.product_div_grad(xt, bb, c.zeros())
    phi = c.zeros()
    st = c.bicgstab(1e-10 * np.abs(bb).max(), 2000, c.Sp, c.Np, bb, phi)
    assert st.exit_reason == 0 and st.iterations > 5
    lap = c.zeros()
    c.product_div_grad(phi, lap, c.zeros())
    s = _box(c.Sp, c.Np)
    assert np.abs(lap[s] - bb[s]).max() <= 1e-10 * np.abs(bb).max() * 1.0001


def test_reduction_modes_agree():
    c = O.OracleCase((37, 21, 19), (1, 1, 1, 1, 1, 1))
    rng = np.random.default_rng(3)
    a = np.ascontiguousarray(rng.standard_normal(c.shape))
    b = np.ascontiguousarray(rng.standard_normal(c.shape))
    c.set_reduction_mode(0)
    d0 = c.dot(c.Sp, c.Np, a, b)
    c.set_reduction_mode(1)
    d1 = c.dot(c.Sp, c.Np, a, b)
    s = _box(c.Sp, c.Np)
    ref = math.fsum((a[s] * b[s]).ravel())
    assert abs(d0 - ref) < 1e-12 * abs(ref) + 1e-12
    assert abs(d1 - ref) < 1e-12 * abs(ref) + 1e-12


def _config1_fields(c, rng, amp=0.01):
    vel = c.zeros(3)
    vel[0] = _field(c, lambda x, y, z: 16 * y * (1 - y) * z * (1 - z), "u")
    for q in range(3):
        vel[q] += rng.normal(0.0, amp, size=vel[q].shape)
    return np.ascontiguousarray(vel)


def test_projection_closure():
    """After a step, D u = b - (D G) p: the velocity divergence equals the
    Poisson residual of BiCGstab, which ties product_div_grad to gradient
    and divergence2 restated separately."""
    c = O.OracleCase((17, 17, 17), (1, 1, 1, 1, 1, 1))
    rng = np.random.default_rng(4)
    vel = _config1_fields(c, rng)
    pre = c.zeros()
    nl = c.zeros(3)
    p = O.StepParams(Re=100.0, dtime=1e-3, jor_sweeps=50, jor_omega=0.8, bicg_rel_tol=1e-10, bicg_max_it=2000,
                     level_pressure=1, inflow_kind=1, inflow_amp=1.0, mask3=None)
    st = c.timestep(p, 0, 0.0, vel, pre, nl)
    assert all(n > 0 for n in st.bicg_iterations)
    assert abs(st.div_max_after - st.bicg_final[2]) <= 1e-6 * st.bicg_final[2] + 1e-15


def test_tube_mask_counts():
    N = (33, 65, 65)
    c = O.OracleCase(N, (1, 1, 1, 1, 1, 1), (1, 1, 1))
    R = 0.3
    r = np.full(N[0] + 7, R)
    m = c.tube_mask(0.5, 0.5, r, r)
    # cross-section of the w grid at one x: fluid points ~ pi R^2 / h^2
    h = 1.0 / 64
    sb, nb = c.SB[2], c.NB[2]
    sec = int(m[2][sb[2] + 3:nb[2] + 4, sb[1] + 3:nb[1] + 4, 10].sum())  # numpy (k, j, i)
    expect = math.pi * R ** 2 / h ** 2
    assert abs(sec - expect) < 4 * (2 * math.pi * R / h)
    # symmetric about the axis on the pressure-located u grid in y
    mu = m[0][4:4 + 65, 4:4 + 65, 10]
    assert np.array_equal(mu, mu[::-1, :]) and np.array_equal(mu, mu[:, ::-1])
