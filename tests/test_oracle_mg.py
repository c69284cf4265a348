"""Pins of the oracle's multigrid V cycle (BiCGstab preconditioner 1).

The cycle restates multigridV and its parts (src/mod_solvers.f90:353-1141,
src/mod_laplace.f90:200-790, src/mod_coeffs.f90:1048-1248, 1683-2035).  The
reference ships no vectors for it, so each part is checked against what the
mathematics fixes:
  - the level sizes of the hierarchy rule, worked by hand;
  - level 1's cdg equals the case's cdg; on uniform grids every level's cdg is
    the closed form (1, -2, 1)/h_g^2, with the wall rows 2(0, -1, 1)/(h h_g)
    that get_stencil's use of the finest u-cell width gives;
  - interpolation reproduces linear functions of the coordinates (stretched
    grids included), away from the edges the reference averages;
  - restriction maps a field linear in the indices to its values at the
    coarse points (exact in binary arithmetic for integer data);
  - the exact solution of a level's cdg system is a fixed point of its sweep;
  - a V cycle is linear and keeps no state between calls, changes its bb only
    on the edges of wall pairs, and preconditions BiCGstab: fewer iterations
    than without it, to the same residual.
"""
import numpy as np
import pytest

import oracle_lib as O

WALLS = (1, 1, 1, 1, 1, 1)
PER_XZ = (-1, -1, 1, 1, -1, -1)
PER = (-1, -1, -1, -1, -1, -1)


def rng(seed):
    return np.random.default_rng(seed)


def prange(c):
    return tuple(slice(c.Sp[d] + 3, c.Np[d] + 4) for d in (2, 1, 0))


def level_ranges(m, g, bc):
    N = m.NN[g - 1]
    lo, hi = [], []
    for d in range(3):
        lo.append(1)
        hi.append(N[d] - 1 if bc[2 * d] == -1 else N[d])
    return lo, hi


def test_hierarchy_sizes():
    c = O.OracleCase((257, 129, 129), WALLS, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    m = O.OracleMG(c)
    assert m.NN == [(257, 129, 129), (129, 65, 65), (65, 33, 33), (33, 17, 17), (17, 9, 9), (9, 5, 5), (5, 3, 3),
                    (3, 3, 3)]
    c = O.OracleCase((70, 9, 40), (1, 1, -1, -1, 1, 1), (2.0, 1.0, 1.0), (0.0, 0.0, 1.5))
    assert O.OracleMG(c).NN == [(70, 9, 40), (70, 5, 40), (70, 3, 40)]
    c = O.OracleCase((20, 20, 20), WALLS)
    assert O.OracleMG(c).NN == [(20, 20, 20)]
    c = O.OracleCase((33, 17, 17), WALLS)
    assert O.OracleMG(c, n_grids_limit=2).NN == [(33, 17, 17), (17, 9, 9)]


@pytest.mark.parametrize("bc,st", [(WALLS, (0.0, 1.5, 1.5)), (PER_XZ, (0.0, 1.2, 0.0))])
def test_level1_cdg_is_the_case_cdg(bc, st):
    c = O.OracleCase((33, 17, 17), bc, (2.0, 1.0, 1.0), st)
    m = O.OracleMG(c)
    for d in range(3):
        assert np.array_equal(m.coeff(1, f"cdg{d + 1}"), c.coeff(f"cdg{d + 1}"))


@pytest.mark.parametrize("bc", [WALLS, PER])
def test_uniform_cdg_closed_form(bc):
    N, L = (33, 17, 9), (2.0, 1.0, 0.5)
    c = O.OracleCase(N, bc, L)
    m = O.OracleMG(c)
    assert m.NN == [(33, 17, 9), (17, 9, 5), (9, 5, 3), (5, 3, 3), (3, 3, 3)]
    for g in range(1, m.G + 1):
        for d in range(3):
            h = L[d] / (N[d] - 1)
            n = m.NN[g - 1][d]
            hg = L[d] / (n - 1)
            cdg = m.coeff(g, f"cdg{d + 1}").reshape(n, 3)
            want = np.tile(np.array([1.0, -2.0, 1.0]) / hg ** 2, (n, 1))
            if bc[2 * d] > 0:
                want[0] = 2.0 * np.array([0.0, -1.0, 1.0]) / (h * hg)
                want[-1] = 2.0 * np.array([1.0, -1.0, 0.0]) / (h * hg)
            np.testing.assert_allclose(cdg, want, rtol=1e-13, atol=0)


def test_restriction_weights_and_walls():
    c = O.OracleCase((17, 9, 9), (1, 1, -1, -1, 1, 1))
    m = O.OracleMG(c)
    for g in range(2, m.G + 1):
        n = m.NN[g - 1]
        r1 = m.coeff(g, "cR1").reshape(n[0], 3)
        r2 = m.coeff(g, "cR2").reshape(n[1], 3)
        assert np.array_equal(r2, np.tile([0.25, 0.5, 0.25], (n[1], 1)))  # periodic: full weighting everywhere
        ident = np.zeros(n[0], bool)
        ident[:2] = ident[-2:] = True
        assert np.array_equal(r1[ident], np.tile([0.0, 1.0, 0.0], (int(ident.sum()), 1)))
        assert np.array_equal(r1[~ident], np.tile([0.25, 0.5, 0.25], (int((~ident).sum()), 1)))


@pytest.mark.parametrize("bc,st", [(WALLS, (1.5, 1.5, 1.5)), (PER_XZ, (0.0, 1.5, 0.0))])
def test_interpolation_reproduces_linear_functions(bc, st):
    c = O.OracleCase((33, 17, 17), bc, (2.0, 1.0, 1.0), st)
    m = O.OracleMG(c)
    for g in range(1, m.G):
        x = [m.coeff(g, f"xp{d + 1}") for d in range(3)]
        X = [m.coeff(g + 1, f"xp{d + 1}") for d in range(3)]
        f = lambda a, b, cc: 1.0 + 2.0 * a[None, None, :] - 3.0 * b[None, :, None] + 0.5 * cc[:, None, None]
        coarse = np.ascontiguousarray(f(*X))
        fine = m.zeros(g)
        m.interpolate_add(g, coarse, fine)
        want = f(*x)
        N = m.NN[g - 1]
        ok = np.zeros(fine.shape, bool)
        ok[4:N[2] + 4, 4:N[1] + 4, 4:N[0] + 4] = True  # 1..N
        # periodic: the last interval interpolates towards the image of point 1
        for d, ax in ((0, 2), (1, 1), (2, 0)):
            if bc[2 * d] == -1:
                sl = [slice(None)] * 3
                sl[ax] = slice(N[d] - 1 + 3, N[d] + 4)
                ok[tuple(sl)] = False
        # walls: the reference averages the coarse edges of wall pairs first
        near = []
        for d, ax in ((0, 2), (1, 1), (2, 0)):
            idx = np.arange(N[d] + 7) - 3
            e = (idx <= 2) | (idx >= N[d] - 1) if bc[2 * d] > 0 else np.zeros(N[d] + 7, bool)
            shp = [1, 1, 1]
            shp[ax] = -1
            near.append(e.reshape(shp))
        two = (near[0] & near[1]) | (near[0] & near[2]) | (near[1] & near[2])
        ok &= ~two
        np.testing.assert_allclose(fine[ok], want[ok], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("bc", [WALLS, PER_XZ])
def test_restriction_maps_linear_index_fields(bc):
    c = O.OracleCase((33, 17, 17), bc, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5) if bc == WALLS else (0.0, 1.5, 0.0))
    m = O.OracleMG(c)
    checked = 0
    for g in range(1, m.G):
        N, Nc = m.NN[g - 1], m.NN[g]
        dd = [(N[d] - 1) // (Nc[d] - 1) for d in range(3)]
        i = np.arange(N[0] + 7) - 3.0
        j = np.arange(N[1] + 7) - 3.0
        k = np.arange(N[2] + 7) - 3.0
        lo, hi = level_ranges(m, g, bc)
        bb = m.zeros(g)
        full = i[None, None, :] + 10.0 * j[None, :, None] + 100.0 * k[:, None, None] + 0.0 * bb
        s = tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))
        bb[s] = full[s]
        coarse = m.zeros(g + 1)
        m.restrict_residual(g, bb, m.zeros(g), coarse)
        nmax = [Nc[d] - 1 if bc[2 * d] == -1 else Nc[d] for d in range(3)]
        for kk in range(1, nmax[2] + 1):
            for jj in range(1, nmax[1] + 1):
                for ii in range(1, nmax[0] + 1):
                    idx = (ii, jj, kk)
                    fi = [dd[d] * (idx[d] - 1) + 1 for d in range(3)]
                    # periodic: a stencil reaching across the wrap is not linear; walls: edges are averaged
                    if any(bc[2 * d] == -1 and (fi[d] - 1 < 1 or fi[d] + 1 > N[d] - 1) for d in range(3)):
                        continue
                    at_wall = [bc[2 * d] > 0 and idx[d] in (1, Nc[d]) for d in range(3)]
                    if sum(at_wall) >= 2:
                        continue
                    assert coarse[kk + 3, jj + 3, ii + 3] == fi[0] + 10.0 * fi[1] + 100.0 * fi[2]
                    checked += 1
    assert checked > 100


def _cdg_apply(m, g, bc, x):
    """The level's compact operator on its pressure range (periodic wrap, zero ghosts)."""
    N = m.NN[g - 1]
    lo, hi = level_ranges(m, g, bc)
    cd = [m.coeff(g, f"cdg{d + 1}").reshape(N[d], 3) for d in range(3)]
    xx = x.copy()
    for d, ax in ((0, 2), (1, 1), (2, 0)):
        if bc[2 * d] == -1:
            def sl(t):
                s = [slice(None)] * 3
                s[ax] = t + 3
                return tuple(s)
            xx[sl(0)] = xx[sl(N[d] - 1)]
            xx[sl(N[d])] = xx[sl(1)]
    y = np.zeros_like(x)
    I = np.arange(lo[0], hi[0] + 1)
    for k in range(lo[2], hi[2] + 1):
        for j in range(lo[1], hi[1] + 1):
            c1 = cd[0][I - 1]
            c2, c3 = cd[1][j - 1], cd[2][k - 1]
            X = lambda a, b, cc: xx[cc + 3, b + 3, a + 3]
            y[k + 3, j + 3, I + 3] = (c1[:, 0] * X(I - 1, j, k) + c1[:, 2] * X(I + 1, j, k) + c2[0] * X(I, j - 1, k)
                                      + c2[2] * X(I, j + 1, k) + c3[0] * X(I, j, k - 1) + c3[2] * X(I, j, k + 1)
                                      + (c1[:, 1] + c2[1] + c3[1]) * X(I, j, k))
    return y


@pytest.mark.parametrize("bc,st", [(WALLS, (0.0, 1.5, 1.5)), (PER_XZ, (0.0, 1.5, 0.0))])
def test_exact_solution_is_a_fixed_point_of_the_sweep(bc, st):
    c = O.OracleCase((33, 17, 17), bc, (2.0, 1.0, 1.0), st)
    m = O.OracleMG(c)
    for g in range(1, m.G + 1):
        lo, hi = level_ranges(m, g, bc)
        s = tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))
        x = m.zeros(g)
        x[s] = rng(g).standard_normal(x[s].shape)
        b = _cdg_apply(m, g, bc, x)
        y = x.copy()
        m.relax(g, 0, 1, b, y)
        np.testing.assert_allclose(y[s], x[s], rtol=0, atol=1e-11 * np.abs(x[s]).max())


@pytest.mark.parametrize("bc,st", [(WALLS, (0.0, 1.5, 1.5)), (PER_XZ, (0.0, 1.5, 0.0))])
def test_vcycle_is_linear_and_stateless(bc, st):
    c = O.OracleCase((33, 17, 17), bc, (2.0, 1.0, 1.0), st)
    m = O.OracleMG(c)
    s = prange(c)
    b1, b2 = c.zeros(), c.zeros()
    b1[s] = rng(1).standard_normal(b1[s].shape)
    b2[s] = rng(2).standard_normal(b2[s].shape)
    z1, z1b, z2, z12 = c.zeros(), c.zeros(), c.zeros(), c.zeros()
    m.vcycle(b1.copy(), z1)
    m.vcycle(b2.copy(), z2)
    m.vcycle(b1.copy(), z1b)
    assert np.array_equal(z1[s], z1b[s])
    m.vcycle(np.ascontiguousarray(2.0 * b1 + b2), z12)
    np.testing.assert_allclose(z12[s], 2.0 * z1[s] + z2[s], rtol=0, atol=1e-12 * np.abs(z12[s]).max())


def test_vcycle_changes_bb_only_on_wall_pair_edges():
    c = O.OracleCase((33, 17, 17), (1, 1, -1, -1, 1, 1), (2.0, 1.0, 1.0), (0.0, 0.0, 1.5))
    m = O.OracleMG(c)
    s = prange(c)
    b = c.zeros()
    b[s] = rng(3).standard_normal(b[s].shape)
    b0 = b.copy()
    m.vcycle(b, c.zeros())
    changed = np.argwhere(b != b0)
    assert len(changed) > 0
    for k, j, i in changed - 3:
        # x-z edges (the only wall pair here), or y's periodic ghost and image planes
        # (restrict exchanges its inputs in place)
        assert (i in (1, c.N[0]) and k in (1, c.N[2])) or j <= 0 or j >= c.N[1]


@pytest.mark.parametrize("N,bc,st", [((33, 17, 17), WALLS, (0.0, 1.5, 1.5)),
                                     ((33, 33, 17), PER_XZ, (0.0, 1.5, 0.0)),
                                     ((17, 17, 17), PER, (0.0, 0.0, 0.0))])
def test_bicgstab_mg_converges_in_fewer_iterations(N, bc, st):
    c = O.OracleCase(N, bc, (2.0, 1.0, 1.0), st)
    m = O.OracleMG(c)
    s = prange(c)
    xt = c.zeros()
    xt[s] = rng(5).standard_normal(xt[s].shape)
    b = c.zeros()
    c.product_div_grad(xt.copy(), b, c.zeros())
    bb = c.zeros()
    bb[s] = b[s]
    eps = 1e-8 * np.abs(bb).max()
    its = {}
    for name in ("plain", "mg"):
        phi = c.zeros()
        if name == "plain":
            st_ = c.bicgstab(eps, 3000, c.Sp, c.Np, bb.copy(), phi)
        else:
            st_ = c.bicgstab_mg(m, eps, 3000, c.Sp, c.Np, bb.copy(), phi)
        assert st_.exit_reason == 0
        r = c.zeros()
        c.product_div_grad(phi.copy(), r, c.zeros())
        assert np.abs((bb - r)[s]).max() < 1.01 * eps
        its[name] = st_.iterations
    assert its["mg"] < its["plain"]


@pytest.mark.parametrize("bc,st", [(WALLS, (0.0, 1.5, 1.5)), (PER_XZ, (0.0, 1.5, 0.0))])
def test_fp32_cycle_tracks_the_fp64_cycle(bc, st):
    """The mixed-precision cycle is the FP64 cycle in binary32: its result lies
    within single-precision rounding of the FP64 one (a wrong table or step
    would not), and it still preconditions BiCGstab: fewer iterations than
    without a preconditioner."""
    c = O.OracleCase((33, 17, 17), bc, (2.0, 1.0, 1.0), st)
    m = O.OracleMG(c)
    s = prange(c)
    b = c.zeros()
    b[s] = rng(7).standard_normal(b[s].shape)
    b64, b32 = b.copy(), b.copy()
    z64, z32 = c.zeros(), c.zeros()
    m.vcycle(b64, z64)
    m.vcycle_f32(b32, z32)
    assert np.array_equal(b64[s], b32[s])  # the same edge averaging of bb, in FP64
    d = np.linalg.norm(z32[s] - z64[s]) / np.linalg.norm(z64[s])
    assert 0 < d < 1e-5
    xt = c.zeros()
    xt[s] = rng(8).standard_normal(xt[s].shape)
    bb = c.zeros()
    c.product_div_grad(xt.copy(), bb, c.zeros())
    b2 = c.zeros()
    b2[s] = bb[s]
    eps = 1e-8 * np.abs(b2).max()
    p64, p32 = c.zeros(), c.zeros()
    s64 = c.bicgstab_mg(m, eps, 3000, c.Sp, c.Np, b2.copy(), p64)
    s32 = c.bicgstab_mg_f32(m, eps, 3000, c.Sp, c.Np, b2.copy(), p32)
    assert s64.exit_reason == 0 and s32.exit_reason == 0
    pp = c.zeros()
    s0 = c.bicgstab(eps, 3000, c.Sp, c.Np, b2.copy(), pp)
    assert s64.iterations < s32.iterations < s0.iterations or s32.iterations <= s64.iterations
