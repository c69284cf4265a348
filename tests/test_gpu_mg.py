"""GPU parity of the multigrid V cycle (BiCGstab preconditioner 1).

Level tables, one V cycle (its result and the edge averaging it applies to its
right-hand side), BiCGstab preconditioned by it, and a projection step with it,
each against the oracle's restatement, bit for bit: both sides evaluate the
reference loops in the same operation order without fused multiply-adds.
Values outside the pressure range (ghosts, periodic image planes) are not
compared: the reference leaves them to exchange_relax, which both sides apply
before any read."""
import ctypes

import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O
from test_gpu_parity import box, compare_state, pair, rnd, same

pytestmark = pytest.mark.gpu

SPECS = [
    pytest.param(((33, 17, 17), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5)), id="walls-stretched"),
    pytest.param(((33, 17, 33), (-1, -1, 1, 1, -1, -1), (1.0, 1.0, 1.0), (0.0, 1.2, 0.0)), id="periodic-xz"),
    pytest.param(((17, 17, 17), (-1, -1, -1, -1, -1, -1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)), id="periodic"),
    pytest.param(((70, 9, 40), (1, 1, -1, -1, 1, 1), (2.0, 1.0, 1.0), (0.0, 0.0, 1.5)), id="ragged-walls-x"),
    pytest.param(((67, 10, 37), (-1, -1, 1, 1, -1, -1), (1.0, 1.0, 1.0), (0.0, 1.3, 0.0)), id="ragged-periodic-xz"),
    pytest.param(((20, 20, 20), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)), id="one-level"),
    # level 2 (33x33x17) above the fused-kernel size: launched level by level, levels 3.. fused
    pytest.param(((65, 65, 33), (1, 1, -1, -1, 1, 1), (1.0, 1.0, 0.5), (0.0, 0.0, 1.5)), id="launched-levels"),
    # edge lines of 4 x 257 points on level 1: more than one workgroup's threads, so the
    # launched edge averaging takes a second pass over a line set.  x is the stretched
    # direction: tanh stretching 1.5 over 9 points puts the wall velocity coordinates
    # outside atanh's domain (NaN, in the reference's coord_tanh as here)
    pytest.param(((257, 9, 9), (1, 1, 1, 1, 1, 1), (4.0, 1.0, 1.0), (1.5, 0.0, 0.0)), id="long-lines"),
]


def prange(oc):
    return box(oc.Sp, oc.Np)


@pytest.mark.parametrize("spec", SPECS)
def test_levels_and_tables(spec):
    oc, gc = pair(spec)
    om = O.OracleMG(oc)
    assert gc.mg_levels() == om.NN
    for g in range(1, om.G + 1):
        names = [f"cdg{d}" for d in (1, 2, 3)]
        if g >= 2:
            names += [f"cR{d}" for d in (1, 2, 3)]
        if g < om.G:
            names += [f"cI{d}" for d in (1, 2, 3)]
        for nm in names:
            same(gc.mg_coeff(g, nm), om.coeff(g, nm), f"level {g} {nm}")


def _vcycle_pair(spec, cfg=None, seed=300, fp32=False):
    oc, gc = pair(spec)
    om = O.OracleMG(oc, *cfg) if cfg else O.OracleMG(oc)
    if cfg:
        gc.mg_configure(*cfg)
    s = prange(oc)
    bb = oc.zeros()
    bb[s] = rnd(oc.shape, seed)[s]
    z0 = rnd(oc.shape, seed + 1)  # the result must not depend on what phi held
    gc.upload(ih.RES, bb)
    gc.upload(ih.PRE, z0)
    z = z0.copy()
    if fp32:
        gc.multigrid_v_f32(ih.RES, ih.PRE)
        om.vcycle_f32(bb, z)
    else:
        gc.multigrid_v(ih.RES, ih.PRE)
        om.vcycle(bb, z)
    return oc, gc, om, bb, z, s


@pytest.mark.parametrize("spec", SPECS)
def test_vcycle(spec):
    oc, gc, om, bb, z, s = _vcycle_pair(spec)
    same(gc.download(ih.PRE)[s], z[s], "V cycle result")
    same(gc.download(ih.RES)[s], bb[s], "V cycle right-hand side after the edge averaging")
    # a second cycle on the averaged right-hand side: no state kept between calls
    z2 = oc.zeros()
    om.vcycle(bb, z2)
    gc.multigrid_v(ih.RES, ih.PRE)
    same(gc.download(ih.PRE)[s], z2[s], "second V cycle")


@pytest.mark.parametrize("cfg", [(2, 2, 3, 1), (15, 1, 1, 2), (3, 3, 2, 5)])
def test_vcycle_configurations(cfg):
    spec = ((33, 17, 17), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    oc, gc, om, bb, z, s = _vcycle_pair(spec, cfg, seed=310)
    same(gc.download(ih.PRE)[s], z[s], f"V cycle {cfg}")


@pytest.mark.parametrize("spec", SPECS[:5] + SPECS[6:])
def test_bicgstab_mg(spec):
    oc, gc = pair(spec)
    om = O.OracleMG(oc)
    xt = rnd(oc.shape, 320)
    bb = oc.zeros()
    oc.product_div_grad(xt.copy(), bb, oc.zeros())
    b2 = oc.zeros()
    s = prange(oc)
    b2[s] = bb[s]
    bb = b2
    phi = np.ascontiguousarray(0.1 * rnd(oc.shape, 321))
    gc.upload(ih.RES, bb)
    gc.upload(ih.PRE, phi)
    eps = 1e-9 * np.abs(bb).max()
    so = oc.bicgstab_mg(om, eps, 1000, oc.Sp, oc.Np, bb, phi)
    sg = gc.bicgstab_precond(eps, 1000, oc.Sp, oc.Np, 1)
    assert so.exit_reason == 0
    assert so.iterations == sg.iterations and so.exit_reason == sg.exit_reason
    assert so.norm_inf_final == sg.norm_inf_final
    same(gc.download(ih.PRE)[s], phi[s], "multigrid BiCGstab phi")


def test_timestep_with_multigrid():
    """The stenosed-tube step (mask, pulsatile inflow) with preconditioner 1."""
    from test_gpu_parity import tube_radius
    spec = ((33, 17, 17), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    oc, gc = pair(spec)
    r_u = tube_radius(oc.coords(0, 1))
    r_p = tube_radius(oc.coords(0, 0))
    mask = oc.tube_mask(0.5, 0.5, r_u, r_p)
    gc.set_tube_mask(0.5, 0.5, r_u, r_p)
    vel = oc.zeros(3)
    vel[0] = 1.0 * mask[0]
    vel = np.ascontiguousarray(vel)
    pre, nl = oc.zeros(), oc.zeros(3)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.NL1 + q, nl[q])
    gc.upload(ih.PRE, pre)
    kw = dict(Re=500.0, dtime=5e-4, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=3000,
              level_pressure=1, inflow_kind=2, inflow_amp=1.0, jor_fp32=0, precond_sweeps=0, precond_poisson=1)
    po = O.StepParams(mask3=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), **kw)
    pg = ih.StepParams(use_mask=1, **kw)
    t = 0.0
    for step in range(2):
        so = oc.timestep(po, step, t, vel, pre, nl)
        sg = gc.timestep(pg, t)
        t += 5e-4
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        assert list(so.bicg_final) == list(sg.bicg_final)
        compare_state(oc, gc, vel, pre, nl, f"multigrid step {step}")


# ---- the FP32 cycle of the mixed-precision solve -----------------------------

@pytest.mark.parametrize("spec", SPECS)
def test_vcycle_f32(spec):
    oc, gc, om, bb, z, s = _vcycle_pair(spec, seed=330, fp32=True)
    same(gc.download(ih.PRE)[s], z[s], "FP32 V cycle result")
    same(gc.download(ih.RES)[s], bb[s], "FP32 V cycle: FP64 edge averaging of the right-hand side")


@pytest.mark.parametrize("spec", SPECS[:2] + SPECS[3:5] + SPECS[6:])
def test_bicgstab_mg_f32(spec):
    oc, gc = pair(spec)
    om = O.OracleMG(oc)
    xt = rnd(oc.shape, 340)
    bb = oc.zeros()
    oc.product_div_grad(xt.copy(), bb, oc.zeros())
    s = prange(oc)
    b2 = oc.zeros()
    b2[s] = bb[s]
    phi = np.ascontiguousarray(0.1 * rnd(oc.shape, 341))
    gc.upload(ih.RES, b2)
    gc.upload(ih.PRE, phi)
    # parity, converged or not: an FP32 M can stall short of tight tolerances
    eps = 1e-8 * np.abs(b2).max()
    so = oc.bicgstab_mg_f32(om, eps, 400, oc.Sp, oc.Np, b2, phi)
    sg = gc.bicgstab_mg_f32(eps, 400, oc.Sp, oc.Np)
    assert so.iterations == sg.iterations and so.exit_reason == sg.exit_reason
    assert so.norm_inf_final == sg.norm_inf_final
    same(gc.download(ih.PRE)[s], phi[s], "FP32-multigrid BiCGstab phi")


def test_timestep_mixed_multigrid():
    """The tube step in the mixed form with the FP32 cycle: FP32 JOR defect
    correction, BiCGstab preconditioned by the FP32 V cycle, bit-exact with the
    oracle's binary32 restatement."""
    from test_gpu_parity import tube_radius
    spec = ((33, 17, 17), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    oc, gc = pair(spec)
    r_u = tube_radius(oc.coords(0, 1))
    r_p = tube_radius(oc.coords(0, 0))
    mask = oc.tube_mask(0.5, 0.5, r_u, r_p)
    gc.set_tube_mask(0.5, 0.5, r_u, r_p)
    vel = oc.zeros(3)
    vel[0] = 1.0 * mask[0]
    vel = np.ascontiguousarray(vel)
    pre, nl = oc.zeros(), oc.zeros(3)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.NL1 + q, nl[q])
    gc.upload(ih.PRE, pre)
    kw = dict(Re=500.0, dtime=5e-4, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=3000,
              level_pressure=1, inflow_kind=2, inflow_amp=1.0, jor_fp32=1, precond_sweeps=0, precond_poisson=1,
              mg_fp32=1)
    po = O.StepParams(mask3=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), **kw)
    pg = ih.StepParams(use_mask=1, **kw)
    t = 0.0
    for step in range(2):
        so = oc.timestep(po, step, t, vel, pre, nl)
        sg = gc.timestep(pg, t)
        t += 5e-4
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        compare_state(oc, gc, vel, pre, nl, f"mixed multigrid step {step}")
