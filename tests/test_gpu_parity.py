"""GPU parity: every operator of the HIP library against the CPU oracle.

The library is called through its C ABI (ctypes binding) with fields resident
on the device; the oracle runs on the host on the same seeded inputs.  All FP64
work here is compared bit for bit: both sides evaluate the reference loops in
the same operation order without fused multiply-adds, and the dot products
use the same fixed reduction tree.
"""
import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O

pytestmark = pytest.mark.gpu

GRIDS = [
    pytest.param(((17, 13, 11), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5)), id="walls-stretched"),
    pytest.param(((16, 12, 10), (-1, -1, 1, 1, -1, -1), (1.0, 1.0, 1.0), (0.0, 1.2, 0.0)), id="periodic-xz"),
]
# several operator tiles in x and y and two z chunks of the fused kernels
MULTI = [
    pytest.param(((70, 9, 40), (1, 1, -1, -1, 1, 1), (2.0, 1.0, 1.0), (0.0, 0.0, 1.5)), id="multi-walls-x"),
    pytest.param(((67, 10, 37), (-1, -1, 1, 1, -1, -1), (1.0, 1.0, 1.0), (0.0, 1.3, 0.0)), id="multi-periodic-xz"),
]


def pair(spec):
    N, bc, L, st = spec
    oc = O.OracleCase(N, bc, L, st)
    gc = ih.Case(N, bc, L, st)
    return oc, gc


def rnd(shape, seed):
    return np.ascontiguousarray(np.random.default_rng(seed).standard_normal(shape))


def same(a, b, what=""):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: {int((a != b).sum())} values differ, max |diff| {np.nanmax(d):.3e}")


def box(lo, hi):
    return tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))


def readable(oc):
    """Cells a one-dimensional stencil from the pressure box can read: at most
    one index outside the box.  Edge and corner ghosts are filled from other
    ghosts by the exchange order and are never read."""
    out = np.zeros(oc.shape, dtype=int)
    for d, ax in ((0, 2), (1, 1), (2, 0)):
        idx = np.arange(oc.N[d] + 7) - 3
        outside = (idx < oc.Sp[d]) | (idx > oc.Np[d])
        shp = [1, 1, 1]
        shp[ax] = -1
        out = out + outside.reshape(shp)
    return out <= 1


@pytest.mark.parametrize("spec", GRIDS)
def test_exchange(spec):
    oc, gc = pair(spec)
    for d in (1, 2, 3):
        for vd in (0, 1, 2, 3):
            f = rnd(oc.shape, 10 * d + vd)
            gc.upload(ih.TMP1, f)
            oc.exchange(d, vd, f)
            gc.exchange(d, vd, ih.TMP1)
            same(gc.download(ih.TMP1), f, f"exchange dir {d} vel_dir {vd}")


@pytest.mark.parametrize("spec", GRIDS)
def test_divergence2(spec):
    oc, gc = pair(spec)
    vel = np.ascontiguousarray(np.stack([rnd(oc.shape, s) for s in (1, 2, 3)]))
    div = rnd(oc.shape, 4)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
    gc.upload(ih.RES, div)
    oc.divergence2(vel, div)
    gc.divergence2()
    same(gc.download(ih.RES), div, "div")
    for q in range(3):
        same(gc.download(ih.VEL1 + q), vel[q], f"vel{q + 1} ghosts")


@pytest.mark.parametrize("spec", GRIDS)
@pytest.mark.parametrize("m", [1, 2, 3])
def test_gradient(spec, m):
    oc, gc = pair(spec)
    p = rnd(oc.shape, 5)
    g = rnd(oc.shape, 6)
    gc.upload(ih.PRE, p)
    gc.upload(ih.GPRE, g)
    oc.gradient(m, p, g)
    gc.gradient(m, ih.PRE, ih.GPRE)
    same(gc.download(ih.GPRE), g, "grad")
    same(gc.download(ih.PRE), p, "pre ghosts")


@pytest.mark.parametrize("spec", GRIDS)
def test_interpolate_vel(spec):
    oc, gc = pair(spec)
    vel = np.ascontiguousarray(np.stack([rnd(oc.shape, s) for s in (7, 8, 9)]))
    work = np.ascontiguousarray(np.stack([rnd(oc.shape, s) for s in (10, 11, 12)]))
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.WORK1 + q, work[q])
    oc.interpolate_vel(vel, work)
    gc.interpolate_vel()
    for q in range(3):
        same(gc.download(ih.WORK1 + q), work[q], f"work{q + 1}")


@pytest.mark.parametrize("spec", GRIDS)
@pytest.mark.parametrize("m", [1, 2, 3])
def test_interpolate_pre_vel(spec, m):
    oc, gc = pair(spec)
    phi = rnd(oc.shape, 13)
    out = rnd(oc.shape, 14)
    gc.upload(ih.TMP1, phi)
    gc.upload(ih.TMP2, out)
    oc.interpolate2_pre_vel(m, phi, out)
    gc.interpolate2_pre_vel(m, ih.TMP1, ih.TMP2)
    same(gc.download(ih.TMP2), out, "pre->vel")


@pytest.mark.parametrize("spec", GRIDS)
@pytest.mark.parametrize("cid", range(1, 10))
def test_nonlinear_upwind_component(spec, cid):
    """device_Nonlinear_upwind(component_id): pp of both signs."""
    oc, gc = pair(spec)
    q, d = (cid - 1) // 3, (cid - 1) % 3
    pp = rnd(oc.shape, 20 + cid)
    vel = rnd(oc.shape, 30 + cid)
    nl = rnd(oc.shape, 40 + cid)
    gc.upload(ih.TMP1, pp)
    gc.upload(ih.TMP2, vel)
    gc.upload(ih.TMP3, nl)
    gc.nonlinear_upwind(cid, ih.TMP1, ih.TMP2, ih.TMP3)
    got = gc.download(ih.TMP3)
    # values are checked against the oracle in the next test; here: only the
    # component's interior range is written
    s = box(oc.SI[q], oc.NI[q])
    mask = np.ones(oc.shape, bool)
    mask[s] = False
    same(got[mask], nl[mask], "outside the interior range")
    assert np.isfinite(got[s]).all()


@pytest.mark.parametrize("spec", GRIDS)
def test_nonlinear_and_helmholtz_against_oracle(spec):
    """nonlinear() + Helmholtz_explicit() through the per-component calls of the
    reference interface: pp = I_m(work_d), then device_Nonlinear_upwind, then
    device_helmholtz, compared with the oracle's routines."""
    oc, gc = pair(spec)
    vel = np.ascontiguousarray(np.stack([rnd(oc.shape, s) for s in (50, 51, 52)]))
    work = np.ascontiguousarray(np.stack([rnd(oc.shape, s) for s in (53, 54, 55)]))
    nl = np.ascontiguousarray(np.stack([rnd(oc.shape, s) for s in (56, 57, 58)]))
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.WORK1 + q, work[q])
        gc.upload(ih.NL1 + q, nl[q])
    pp = oc.zeros()
    oc.nonlinear(vel, work, nl, pp)
    multL = 1.0 / 137.0
    oc.helmholtz(multL, vel, nl)
    for q in range(3):
        for d in range(3):
            gc.interpolate2_pre_vel(q + 1, ih.WORK1 + d, ih.PP)
            gc.nonlinear_upwind(3 * q + d + 1, ih.PP, ih.VEL1 + q, ih.NL1 + q)
        gc.helmholtz(q + 1, multL, ih.VEL1 + q, ih.NL1 + q)
    for q in range(3):
        same(gc.download(ih.NL1 + q), nl[q], f"nl{q + 1}")


@pytest.mark.parametrize("spec", GRIDS + MULTI)
def test_product_div_grad(spec):
    oc, gc = pair(spec)
    p = rnd(oc.shape, 60)
    lap = rnd(oc.shape, 61)
    gc.upload(ih.TMP1, p)
    gc.upload(ih.TMP2, lap)
    oc.product_div_grad(p, lap, oc.zeros())
    gc.product_div_grad(ih.TMP1, ih.TMP2)
    same(gc.download(ih.TMP2), lap, "div grad")


@pytest.mark.parametrize("spec", GRIDS + MULTI)
@pytest.mark.parametrize("sweeps", [1, 2, 7])
def test_jor_and_residual(spec, sweeps):
    oc, gc = pair(spec)
    rel = rnd(oc.shape, 70)
    bb = rnd(oc.shape, 71)
    gc.upload(ih.PRE, rel)
    gc.upload(ih.RES, bb)
    oc.jor(sweeps, 0.8, oc.Sp, oc.Np, rel, bb, oc.zeros())
    gc.jor(sweeps, 0.8, oc.Sp, oc.Np)
    r = readable(oc)
    same(gc.download(ih.PRE)[r], rel[r], "jor")
    r_o = oc.residual_cdg(oc.Sp, oc.Np, rel, bb)
    r_g = gc.residual_cdg(oc.Sp, oc.Np)
    assert r_o == r_g


@pytest.mark.parametrize("spec", GRIDS + MULTI)
@pytest.mark.parametrize("lo,hi", [((2, 2, 2), (-1, -1, -1)), ((3, 1, 2), (0, -2, -1)), ((2, 3, 1), (-2, 0, 0)),
                                   ((4, 2, 3), (-3, -1, -2))], ids=["even-start", "odd-start", "even-short", "inner"])
def test_jor_sub_box(spec, lo, hi):
    """JOR over boxes that start and end anywhere in a row: the vector sweep's
    first and last vectors then reach past the box on either side (a start at
    an even i is not aligned to a 2-point vector)."""
    oc, gc = pair(spec)
    ss = list(lo)
    nn = [oc.N[d] + hi[d] for d in range(3)]
    rel = rnd(oc.shape, 80)
    bb = rnd(oc.shape, 81)
    gc.upload(ih.PRE, rel)
    gc.upload(ih.RES, bb)
    oc.jor(3, 0.8, ss, nn, rel, bb, oc.zeros())
    gc.jor(3, 0.8, ss, nn)
    same(gc.download(ih.PRE)[readable(oc)], rel[readable(oc)], f"jor on {ss}..{nn}")


@pytest.mark.parametrize("spec", GRIDS + MULTI)
def test_dot_fixed_tree(spec):
    oc, gc = pair(spec)
    a = rnd(oc.shape, 80)
    b = rnd(oc.shape, 81)
    gc.upload(ih.TMP1, a)
    gc.upload(ih.TMP2, b)
    oc.set_reduction_mode(1)
    assert oc.dot(oc.Sp, oc.Np, a, b) == gc.dot(oc.Sp, oc.Np, ih.TMP1, ih.TMP2)


@pytest.mark.parametrize("spec", GRIDS + MULTI)
def test_bicgstab(spec):
    oc, gc = pair(spec)
    xt = rnd(oc.shape, 90)
    bb = oc.zeros()
    oc.product_div_grad(xt.copy(), bb, oc.zeros())
    phi = np.ascontiguousarray(0.1 * rnd(oc.shape, 91))
    gc.upload(ih.RES, bb)
    gc.upload(ih.PRE, phi)
    eps = 1e-9 * np.abs(bb).max()
    so = oc.bicgstab(eps, 500, oc.Sp, oc.Np, bb, phi)
    sg = gc.bicgstab(eps, 500, oc.Sp, oc.Np)
    assert so.iterations == sg.iterations and so.exit_reason == sg.exit_reason
    assert so.norm_inf_final == sg.norm_inf_final
    same(gc.download(ih.PRE), phi, "bicgstab phi")


@pytest.mark.parametrize("spec", GRIDS)
def test_fill_corners_and_level_pressure(spec):
    oc, gc = pair(spec)
    p = rnd(oc.shape, 95)
    gc.upload(ih.PRE, p)
    oc.fill_corners(p)
    oc.level_pressure(p)
    gc.fill_corners(ih.PRE)
    gc.level_pressure(ih.PRE)
    same(gc.download(ih.PRE), p, "corners + level")


def tube_radius(x, R=0.3, x0=1.0, ell=0.4, delta=1.0 - 1.0 / np.sqrt(2.0)):
    r = np.full_like(x, R)
    inside = np.abs(x - x0) <= ell
    r[inside] = R * (1.0 - delta / 2.0 * (1.0 + np.cos(np.pi * (x[inside] - x0) / ell)))
    return r


@pytest.mark.parametrize("spec", GRIDS)
def test_mask_bit_exact(spec):
    oc, gc = pair(spec)
    r_p = tube_radius(oc.coords(0, 0), x0=0.5 * spec[2][0])
    r_u = tube_radius(oc.coords(0, 1), x0=0.5 * spec[2][0])
    m_o = oc.tube_mask(0.5, 0.5, r_u, r_p)
    gc.set_tube_mask(0.5, 0.5, r_u, r_p)
    m_g = gc.get_mask()
    assert np.array_equal(m_o, m_g)
    assert 0 < int((m_o[0] == 0).sum())


def step_params_pair(kind, jor=20, tol=1e-9, use_mask=0, mask3=None, Re=100.0, dt=1e-3):
    po = O.StepParams(Re=Re, dtime=dt, jor_sweeps=jor, jor_omega=0.8, bicg_rel_tol=tol, bicg_max_it=3000,
                      level_pressure=1, inflow_kind=kind, inflow_amp=1.0, mask3=mask3)
    pg = ih.StepParams(Re=Re, dtime=dt, jor_sweeps=jor, jor_omega=0.8, bicg_rel_tol=tol, bicg_max_it=3000,
                       level_pressure=1, inflow_kind=kind, inflow_amp=1.0, use_mask=use_mask, jor_fp32=0)
    return po, pg


def compare_state(oc, gc, vel, pre, nl, tag):
    for q in range(3):
        s = box(oc.SB[q], oc.NB[q])
        same(gc.download(ih.VEL1 + q)[s], vel[q][s], f"{tag} vel{q + 1}")
        s = box(oc.SI[q], oc.NI[q])
        same(gc.download(ih.NL1 + q)[s], nl[q][s], f"{tag} nl{q + 1}")
    s = box(oc.Sp, oc.Np)
    same(gc.download(ih.PRE)[s], pre[s], f"{tag} pre")


def test_timestep_parabolic_channel():
    """Config-1 shape (walls, parabolic inflow/outflow, seeded noise) on 17^3."""
    spec = ((17, 17, 17), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    oc, gc = pair(spec)
    rng = np.random.default_rng(7)
    yp, zp = oc.coords(1, 0), oc.coords(2, 0)
    vel = oc.zeros(3)
    vel[0] = (16 * yp[None, :, None] * (1 - yp[None, :, None]) * zp[:, None, None] * (1 - zp[:, None, None])
              * np.ones(oc.shape))
    for q in range(3):
        vel[q] += rng.normal(0.0, 0.01, size=oc.shape)
    vel = np.ascontiguousarray(vel)
    pre = oc.zeros()
    nl = oc.zeros(3)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.NL1 + q, nl[q])
    gc.upload(ih.PRE, pre)
    po, pg = step_params_pair(1)
    t = 0.0
    for step in range(2):
        so = oc.timestep(po, step, t, vel, pre, nl)
        sg = gc.timestep(pg, t)
        t += 1e-3
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        assert list(so.bicg_final) == list(sg.bicg_final)
        assert so.div_max_after == sg.div_max_after
        compare_state(oc, gc, vel, pre, nl, f"step {step}")


def test_timestep_stenosed_tube():
    """Config-2 shape (tanh-stretched y,z, stenosed tube mask, pulsatile inflow)."""
    spec = ((33, 17, 17), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    oc, gc = pair(spec)
    r_u = tube_radius(oc.coords(0, 1))
    r_p = tube_radius(oc.coords(0, 0))
    mask = oc.tube_mask(0.5, 0.5, r_u, r_p)
    gc.set_tube_mask(0.5, 0.5, r_u, r_p)
    vel = oc.zeros(3)
    vel[0] = 1.0 * mask[0]
    vel = np.ascontiguousarray(vel)
    pre = oc.zeros()
    nl = oc.zeros(3)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.NL1 + q, nl[q])
    gc.upload(ih.PRE, pre)
    import ctypes
    po, pg = step_params_pair(2, use_mask=1, mask3=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                              Re=500.0, dt=5e-4)
    t = 0.0
    for step in range(2):
        so = oc.timestep(po, step, t, vel, pre, nl)
        sg = gc.timestep(pg, t)
        t += 5e-4
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        compare_state(oc, gc, vel, pre, nl, f"step {step}")


def test_reference_signature_entry_points():
    """Host-array calls with the reference's argument lists."""
    spec = ((17, 13, 11), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    oc, gc = pair(spec)
    p = rnd(oc.shape, 100)
    g = rnd(oc.shape, 101)
    p2, g2 = p.copy(), g.copy()
    oc.gradient(2, p, g)
    gc.device_gradient(2, p2, g2)
    same(g2, g, "device_gradient")
    vel = np.ascontiguousarray(np.stack([rnd(oc.shape, s) for s in (102, 103, 104)]))
    div = rnd(oc.shape, 105)
    vel2, div2 = vel.copy(), div.copy()
    oc.divergence2(vel, div)
    gc.device_divergence2(vel2, div2)
    same(div2, div, "device_divergence2")
    x, y = rnd(oc.shape, 106), rnd(oc.shape, 107)
    z = np.zeros_like(x)
    gc.device_axpby(1.0, -1.0, x, y, z)
    same(z, x - y, "device_axpby")
    phi = rnd(oc.shape, 108)
    inter = rnd(oc.shape, 109)
    inter2 = inter.copy()
    oc.interpolate2_vel_pre(3, phi, inter)
    gc.device_interpolate2(3, phi.copy(), inter2)
    same(inter2, inter, "device_interpolate2")
