"""Pins of the coefficient tables and operators that the polynomial and
closed-form checks of test_oracle.py leave open.

- Every row of every mapped first-derivative table, the one-sided wall rows of
  stretched grids included, differentiates the coordinate it reads to 1.  This
  is the chain-rule mapping dxi = sum cc x (src/mod_coeffs.f90:476-480): after
  it, D(x) = 1 holds by construction, so a row built on the wrong grid's
  coordinates, or a tap offset by one, fails it.  Every second-derivative row
  (cp??, cu??) annihilates constants and its own grid's coordinate.  The
  interpolation tables are not mapped: every row reproduces the index
  coordinate of its target point (p points at i, velocity points at i + 1/2)
  and a constant, exactly.
- The explicit Helmholtz term (src/mod_diff.f90:549-762) of linear fields in
  all three components on a grid stretched and walled in every direction is
  zero: each component and direction must use the table of the grid its data
  sits on (cu11/cv22/cw33 along the component's own direction, cp?? along the
  others).  A numpy application of the other grid's table to the same data is
  checked to be far from zero, so the test separates the two.
- The upwind term (src/mod_diff.f90:784-1424) of linear fields in all three
  components advected by mixed-sign velocities: every cross-direction cNp and
  own-direction cNu table is exercised (the linear check of
  test_oracle_semantics.py has v = w = 0).
"""
import numpy as np
import pytest

import oracle_lib as O

EPS = np.finfo(np.float64).eps
STRETCHED_WALLS = ((17, 13, 11), (1,) * 6, (2.0, 1.0, 1.0), (1.3, 1.5, 1.5))

# name pattern -> (first tap, taps, grid it reads: 'p' or 'u', kind)
TABLES = {
    "cGp": (-2, 6, "p", "d1"),   # gradient, p -> velocity grid
    "cD": (-3, 6, "u", "d1"),    # divergence, velocity grid -> p
    "cNpU": (-3, 7, "p", "d1"),  # upwind derivative along a direction where the field sits on p points
    "cNpD": (-3, 7, "p", "d1"),
    "cNuU": (-3, 7, "u", "d1"),  # upwind derivative along the field's own staggered direction
    "cNuD": (-3, 7, "u", "d1"),
    "cIp": (-2, 6, "p", "i"),    # interpolation p -> velocity grid
    "cIu": (-3, 6, "u", "i"),    # interpolation velocity grid -> p
}
NAMES = {
    "cGp": ("cGp1", "cGp2", "cGp3"), "cD": ("cDu1", "cDv2", "cDw3"),
    "cNpU": ("cNp1U", "cNp2U", "cNp3U"), "cNpD": ("cNp1D", "cNp2D", "cNp3D"),
    "cNuU": ("cNu1U", "cNv2U", "cNw3U"), "cNuD": ("cNu1D", "cNv2D", "cNw3D"),
    "cIp": ("cIpu", "cIpv", "cIpw"), "cIu": ("cIup", "cIvp", "cIwp"),
}


def _rows(c, name, d, taps):
    return c.coeff(name).reshape(c.N[d] + 1, taps)


@pytest.mark.parametrize("spec", [STRETCHED_WALLS, ((21, 17, 15), (1, 1, -1, -1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 1.4))],
                         ids=["stretched-walls", "mixed"])
@pytest.mark.parametrize("table", sorted(TABLES))
def test_every_row_differentiates_or_interpolates_the_coordinate(spec, table):
    c = O.OracleCase(*spec)
    lo, taps, src, kind = TABLES[table]
    checked = 0
    for d in range(3):
        name = NAMES[table][d]
        w = _rows(c, name, d, taps)
        x_src = c.coords(d, 1 if src == "u" else 0)     # (-3:N+3) at [i+3]
        x_dst = c.coords(d, 0 if src == "u" else 1)
        for i in range(c.N[d] + 1):
            row = w[i]
            if not row.any():
                continue  # a row the reference never fills (outside the range that uses it)
            idx = np.arange(lo, lo + taps) + i + 3
            ok = (idx >= 0) & (idx < c.N[d] + 7)
            assert not row[~ok].any(), (name, i)
            xs = x_src[idx[ok]]
            r = row[ok]
            bound = 16 * EPS * np.abs(r * xs).sum() + 1e-300
            if kind == "d1":
                assert abs((r * xs).sum() - 1.0) <= bound + 16 * EPS, (name, i, (r * xs).sum())
                assert abs(r.sum()) <= 16 * EPS * np.abs(r).sum(), (name, i, r.sum())
            else:
                # index coordinates: p points at i, velocity points at i + 1/2
                xi = idx[ok] - 3 + (0.5 if src == "u" else 0.0)
                target = i + (0.0 if src == "u" else 0.5)
                assert (r * xi).sum() == target, (name, i, (r * xi).sum(), target)
                assert r.sum() == 1.0, (name, i, r.sum())
            checked += 1
    assert checked > 0


def _field(c, fn, grid):
    xs = [c.coords(d, 1 if grid == "uvw"[d] else 0) for d in range(3)]
    K, J, I = np.meshgrid(np.arange(c.N[2] + 7), np.arange(c.N[1] + 7), np.arange(c.N[0] + 7), indexing="ij")
    return np.ascontiguousarray(fn(xs[0][I], xs[1][J], xs[2][K]))


def _interior(c, m, pad=3):
    lo = [c.SI[m][d] + pad for d in range(3)]
    hi = [c.NI[m][d] - pad for d in range(3)]
    return tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))


SECOND = {0: ("cp11", "cu11"), 1: ("cp22", "cv22"), 2: ("cp33", "cw33")}  # (pressure grid, velocity grid)


@pytest.mark.parametrize("spec", [STRETCHED_WALLS, ((21, 17, 15), (1, 1, -1, -1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 1.4))],
                         ids=["stretched-walls", "mixed"])
def test_second_derivative_rows_annihilate_their_coordinate(spec):
    c = O.OracleCase(*spec)
    for d in range(3):
        for which, name in enumerate(SECOND[d]):
            w = _rows(c, name, d, 7)
            x = c.coords(d, which)
            for i in range(c.N[d] + 1):
                r = w[i]
                if not r.any():
                    continue
                idx = np.arange(-3, 4) + i + 3
                ok = (idx >= 0) & (idx < c.N[d] + 7)
                assert not r[~ok].any(), (name, i)
                terms = r[ok] * x[idx[ok]]
                assert abs(terms.sum()) <= 16 * EPS * np.abs(terms).sum(), (name, i, terms.sum())
                assert abs(r.sum()) <= 16 * EPS * np.abs(r).sum(), (name, i, r.sum())


def test_helmholtz_of_linear_fields_in_all_components_is_zero():
    c = O.OracleCase((25, 23, 21), (1,) * 6, (1.0, 1.0, 1.0), (1.2, 1.5, 1.3))
    abg = [(1.0, -0.5, 0.25), (-0.75, 2.0, 0.5), (0.3, 0.6, -1.5)]
    vel = c.zeros(3)
    for m in range(3):
        a, b, g = abg[m]
        vel[m] = _field(c, lambda x, y, z, a=a, b=b, g=g: a * x + b * y + g * z + 0.5, "uvw"[m])
    nl = c.zeros(3)
    c.helmholtz(0.5, vel, nl)
    for m in range(3):
        s = _interior(c, m)
        assert nl[m][s].size > 0
        np.testing.assert_allclose(nl[m][s], 0.0, atol=1e-9)
    # negative control: component 2 differentiated along y with the pressure
    # grid's table (cp22) instead of its own grid's (cv22) is not zero
    w = _rows(c, "cp22", 1, 7)
    v = vel[1]
    k, i = (c.SI[1][2] + c.NI[1][2]) // 2 + 3, (c.SI[1][0] + c.NI[1][0]) // 2 + 3
    wrong = [sum(w[j, t] * v[k, j + t - 3 + 3, i] for t in range(7)) for j in range(c.SI[1][1] + 3, c.NI[1][1] - 2)]
    assert max(abs(x) for x in wrong) > 1e-3


def test_upwind_of_linear_fields_in_all_components_with_mixed_signs():
    c = O.OracleCase((25, 23, 21), (1,) * 6, (1.0, 1.0, 1.0), (1.2, 1.5, 1.3))
    slopes = [(1.0, -2.0, 0.5), (0.25, 1.5, -1.0), (-0.5, 0.75, 2.0)]  # d u_m / d x_d
    vel = c.zeros(3)
    for m in range(3):
        s = slopes[m]
        vel[m] = _field(c, lambda x, y, z, s=s: s[0] * x + s[1] * y + s[2] * z + 0.125, "uvw"[m])
    adv = (0.75, -1.25, 0.5)  # constant advecting velocities on the pressure grid, mixed signs
    work = c.zeros(3)
    for d in range(3):
        work[d][:] = adv[d]
    nl = c.zeros(3)
    c.nonlinear(vel, work, nl, c.zeros())
    for m in range(3):
        want = sum(adv[d] * slopes[m][d] for d in range(3))
        got = nl[m][_interior(c, m)]
        assert got.size > 0
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)
