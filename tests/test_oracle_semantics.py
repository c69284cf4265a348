"""Further pins of the oracle against what the reference's code or the
mathematics fixes: ghost-fill semantics (src/mod_exchange.f90:128-152),
interpolation exactness, the upwind nonlinear term of a linear field, pressure
levelling, corner filling and the RK3 amplification factor of a time step."""
import numpy as np
import pytest

import oracle_lib as O


def _field(c, fn, grid):
    xs = [c.coords(d, 1 if grid == "uvw"[d] else 0) for d in range(3)]
    K, J, I = np.meshgrid(np.arange(c.N[2] + 7), np.arange(c.N[1] + 7), np.arange(c.N[0] + 7), indexing="ij")
    return np.ascontiguousarray(fn(xs[0][I], xs[1][J], xs[2][K]))


def _box(lo, hi):
    return tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))


def _plane(d, idx, N):
    """numpy index of Fortran plane `idx` of direction d over the transverse range 1..N."""
    sl = [slice(4, N[2] + 4), slice(4, N[1] + 4), slice(4, N[0] + 4)]
    sl[2 - d] = idx + 3
    return tuple(sl)


@pytest.mark.parametrize("vel_dir", [0, 1, 2, 3])
def test_exchange_matches_reference_semantics(vel_dir):
    # mod_exchange.f90:128-152 (and the same blocks for dir 2, 3): periodic ghosts
    # S+s <- N?p+1+s (s=-3..-1), N?p+s <- S-1+s (s=1..3) over 1..N transverse;
    # walls zero b?L..-1 and N+1..N+b?U when vel_dir == dir, else b?L..0 and N+1..N+b?U.
    N = (10, 9, 8)
    c = O.OracleCase(N, (1, 1, -1, -1, 1, 1))
    rng = np.random.default_rng(11)
    for d in range(3):
        phi = np.ascontiguousarray(rng.standard_normal(c.shape))
        ref = phi.copy()
        c.exchange(d + 1, vel_dir, phi)
        S, Np = c.Sp[d], c.Np[d]
        if c.bc[2 * d] == -1:
            for s in range(-3, 0):
                ref[_plane(d, S + s, N)] = ref[_plane(d, Np + 1 + s, N)]
            for s in range(1, 4):
                ref[_plane(d, Np + s, N)] = ref[_plane(d, S - 1 + s, N)]
        else:
            lo_end = -1 if vel_dir == d + 1 else 0
            for s in list(range(-3, lo_end + 1)) + list(range(N[d] + 1, N[d] + 4)):
                ref[_plane(d, s, N)] = 0.0
        assert np.array_equal(phi, ref), (d, vel_dir)


@pytest.mark.parametrize("deg", [0, 1, 3, 5])
def test_interpolations_exact_for_polynomials(deg):
    # 6-point interpolation stencils (cIpu -2:3, cIup -3:2) reproduce degree <= 5
    c = O.OracleCase((27, 23, 21), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0))
    f = lambda x, y, z: (x - 0.3) ** deg + 0.5 * (y - 0.6) ** deg - 0.25 * (z - 0.4) ** deg
    for m in (1, 2, 3):
        p = _field(c, f, "p")
        inter = c.zeros()
        c.interpolate2_pre_vel(m, p, inter)
        exact = _field(c, f, "uvw"[m - 1])
        lo = [c.SI[m - 1][d] + 3 for d in range(3)]
        hi = [c.NI[m - 1][d] - 3 for d in range(3)]
        np.testing.assert_allclose(inter[_box(lo, hi)], exact[_box(lo, hi)], rtol=1e-10, atol=1e-11)
        u = _field(c, f, "uvw"[m - 1])
        back = c.zeros()
        c.interpolate2_vel_pre(m, u, back)
        lo = [c.Sp[d] + 3 for d in range(3)]
        hi = [c.Np[d] - 3 for d in range(3)]
        exact_p = _field(c, f, "p")
        np.testing.assert_allclose(back[_box(lo, hi)], exact_p[_box(lo, hi)], rtol=1e-10, atol=1e-11)


def test_nonlinear_of_linear_field():
    # u = a x + b, v = w = 0: u du/dx = a (a x + b) exactly for the upwind
    # stencils of either branch and for the interpolations; nl_2 = nl_3 = 0
    c = O.OracleCase((29, 17, 15), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0))
    a, b = -0.7, 0.4   # u changes sign inside the box: both upwind branches occur
    vel = c.zeros(3)
    vel[0] = _field(c, lambda x, y, z: a * x + b, "u")
    vel = np.ascontiguousarray(vel)
    c.exchange_all_all(vel)
    work = c.zeros(3)
    c.interpolate_vel(vel, work)
    nl, pp = c.zeros(3), c.zeros()
    c.nonlinear(vel, work, nl, pp)
    exact = _field(c, lambda x, y, z: a * (a * x + b), "u")
    lo = [c.SI[0][d] + 4 for d in range(3)]
    hi = [c.NI[0][d] - 4 for d in range(3)]
    s = _box(lo, hi)
    np.testing.assert_allclose(nl[0][s], exact[s], rtol=1e-10, atol=1e-11)
    for q in (1, 2):
        s = _box([c.SI[q][d] + 4 for d in range(3)], [c.NI[q][d] - 4 for d in range(3)])
        assert np.abs(nl[q][s]).max() < 1e-11


@pytest.mark.parametrize("bc", [(1, 1, 1, 1, 1, 1), (-1, -1, 1, 1, -1, -1)])
def test_level_pressure_removes_the_mean(bc):
    c = O.OracleCase((13, 11, 9), bc)
    rng = np.random.default_rng(12)
    pre = np.ascontiguousarray(rng.standard_normal(c.shape) + 3.0)
    s = _box(c.Sp, c.Np)
    before = pre[s].copy()
    c.level_pressure(pre)
    shift = before - pre[s]
    assert np.ptp(shift) < 1e-14                      # one constant subtracted everywhere
    assert abs(pre[s].mean()) < 1e-13                 # mod_lib.f90:507-546: mean over the box


def test_fill_corners_is_mean_of_three_neighbours():
    # mod_lib.f90:549-575 on walls: edge value = mean of the two face neighbours and the diagonal one
    c = O.OracleCase((10, 9, 8), (1, 1, 1, 1, 1, 1))
    rng = np.random.default_rng(13)
    pre = np.ascontiguousarray(rng.standard_normal(c.shape))
    ref = pre.copy()
    c.fill_corners(pre)
    N1, N2, N3 = c.N
    at = lambda a, i, j, k: a[k + 3, j + 3, i + 3]
    for k in range(2, N3):   # x-y edges away from the x-z and y-z edges filled after them
        assert at(pre, 1, 1, k) == (at(ref, 2, 1, k) + at(ref, 1, 2, k) + at(ref, 2, 2, k)) / 3.0
        assert at(pre, N1, N2, k) == (at(ref, N1 - 1, N2, k) + at(ref, N1, N2 - 1, k) + at(ref, N1 - 1, N2 - 1, k)) / 3.0
    interior = _box([2, 2, 2], [N1 - 1, N2 - 1, N3 - 1])
    assert np.array_equal(pre[interior], ref[interior])


def test_rk3_amplification_of_a_shear_mode():
    """Periodic box, u = A sin(2 pi y), v = w = 0: the nonlinear term and the
    divergence vanish, so one step (3 substeps, aRK = 8/15, 5/12, 3/4,
    bRK = 0, -17/60, -5/12) multiplies u by the third-order polynomial
    1 + z + z^2/2 + z^3/6, z = dt lambda_h / Re, lambda_h the eigenvalue of
    the discrete Laplacian for this mode."""
    c = O.OracleCase((8, 32, 8), (-1, -1, -1, -1, -1, -1), (1.0, 1.0, 1.0))
    A = 0.8
    vel = c.zeros(3)
    vel[0] = _field(c, lambda x, y, z: A * np.sin(2 * np.pi * y), "u")
    vel = np.ascontiguousarray(vel)
    u0 = vel[0].copy()
    nl = c.zeros(3)
    c.helmholtz(1.0, vel.copy(), nl)      # nl = -Lap_h u
    s = _box(c.SI[0], c.NI[0])
    big = np.abs(u0[s]) > 0.5 * A
    lam = (-nl[0][s][big] / u0[s][big])
    assert np.ptp(lam) < 1e-9 * abs(lam.mean())
    lam = lam.mean()
    assert abs(lam + (2 * np.pi) ** 2) < 1e-3 * (2 * np.pi) ** 2
    Re, dt = 10.0, 5e-3
    z = dt * lam / Re
    R = 1 + z + z * z / 2 + z ** 3 / 6
    pre, nl = c.zeros(), c.zeros(3)
    p = O.StepParams(Re=Re, dtime=dt, jor_sweeps=0, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=50,
                     level_pressure=1, inflow_kind=0, inflow_amp=0.0, mask3=None)
    c.timestep(p, 0, 0.0, vel, pre, nl)
    np.testing.assert_allclose(vel[0][s], R * u0[s], rtol=0, atol=1e-12 * A)
    assert np.abs(vel[1][s]).max() < 1e-12 and np.abs(vel[2][s]).max() < 1e-12


def _random_rhs(c, seed):
    """A right-hand side in the range of product_div_grad (so the solve is consistent)."""
    rng = np.random.default_rng(seed)
    xt = np.ascontiguousarray(rng.standard_normal(c.shape))
    bb = c.zeros()
    c.product_div_grad(xt, bb, c.zeros())
    return bb


@pytest.mark.parametrize("bc", [(1, 1, 1, 1, 1, 1), (-1, -1, 1, 1, -1, -1)])
def test_mixed_bicgstab_reaches_the_fp64_solution(bc):
    """Right-preconditioned BiCGstab with FP32 JOR sweeps as M stops on the
    FP64 residual, so it reaches the same tolerance as the plain solver, in
    fewer iterations.  (The two solutions need not agree: product_div_grad has
    a null space beyond the constants, and they differ within it.)"""
    c = O.OracleCase((21, 17, 13), bc, (1.5, 1.0, 1.0), (0.0, 1.3, 0.0))
    bb = _random_rhs(c, 21)
    s = _box(c.Sp, c.Np)
    eps = 1e-10 * np.abs(bb[s]).max()
    x64, x32 = c.zeros(), c.zeros()
    st64 = c.bicgstab(eps, 3000, c.Sp, c.Np, bb, x64)
    st32 = c.bicgstab_mixed(eps, 3000, c.Sp, c.Np, bb, x32, 4, 0.8)
    assert st64.exit_reason == 0 and st32.exit_reason == 0
    assert st32.iterations < st64.iterations
    for x in (x64, x32):
        lap = c.zeros()
        c.product_div_grad(x.copy(), lap, c.zeros())
        assert np.abs(lap[s] - bb[s]).max() <= eps * 1.0001


def test_fp32_jor_correction_tracks_fp64_jor():
    """FP32 defect correction from e = 0 equals the FP64 sweeps up to binary32
    rounding: the first FP32 sweep is the FP64 sweep rounded."""
    c = O.OracleCase((19, 15, 13), (1, 1, 1, 1, 1, 1), (1.5, 1.0, 1.0), (0.0, 1.3, 1.3))
    rng = np.random.default_rng(22)
    bb = np.ascontiguousarray(rng.standard_normal(c.shape))
    s = _box(c.Sp, c.Np)
    for n in (1, 30):
        x64, x32 = c.zeros(), c.zeros()
        c.jor(n, 0.8, c.Sp, c.Np, x64, bb, c.zeros())
        c.jor_mixed(n, 0.8, c.Sp, c.Np, x32, bb)
        err = np.linalg.norm(x32[s] - x64[s]) / np.linalg.norm(x64[s])
        assert 0 < err < 5e-6 * n, (n, err)


def test_nan_residual_stops_as_breakdown():
    """The stop rule for a non-finite residual (no reference counterpart: the
    reference's MAX(ABS(phi), normInf) leaves a NaN's fate to the compiler).
    Max-norms let a NaN win, so one NaN in the right-hand side makes the residual
    norm NaN, and every BiCGstab form stops at its first stop test with exit 2
    instead of reading the NaN residual as converged."""
    c = O.OracleCase((17, 13, 11), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    bb = np.ascontiguousarray(np.random.default_rng(194).standard_normal(c.shape))
    bb[tuple(n // 2 for n in c.shape)] = np.nan
    assert np.isnan(c.residual_cdg(c.Sp, c.Np, c.zeros(), bb))
    for solve in (lambda phi: c.bicgstab(1e-8, 50, c.Sp, c.Np, bb, phi),
                  lambda phi: c.bicgstab_mixed(1e-8, 50, c.Sp, c.Np, bb, phi, 3, 0.8),
                  lambda phi: c.bicgstab_jor64(1e-8, 50, c.Sp, c.Np, bb, phi, 3, 0.8)):
        st = solve(c.zeros())
        assert st.exit_reason == 2 and st.iterations == 0, (st.exit_reason, st.iterations)
    # a finite right-hand side still converges (the rule changes nothing else)
    bb[tuple(n // 2 for n in c.shape)] = 0.0
    xt = c.zeros()
    xt[3:-3, 3:-3, 3:-3] = np.random.default_rng(195).standard_normal(xt[3:-3, 3:-3, 3:-3].shape)
    b2 = c.zeros()
    c.product_div_grad(xt, b2, c.zeros())
    st = c.bicgstab(1e-9 * np.abs(b2).max(), 500, c.Sp, c.Np, b2, c.zeros())
    assert st.exit_reason == 0
