"""GPU parity of the mixed-precision Poisson path.

FP32 JOR smoothing (as defect correction, and as the right preconditioner of
BiCGstab) with residuals, inner products and updates in FP64.  The FP32
arithmetic is IEEE binary32 on both sides in the same operation order, so the
device is compared bit for bit with the oracle's restatement; its distance
from the plain FP64 result is checked separately."""
import ctypes

import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O
from test_gpu_parity import GRIDS, MULTI, box, compare_state, pair, rnd, same, tube_radius

pytestmark = pytest.mark.gpu

# walls on every side at ragged sizes: rows of 70, 131 and 261 points, so the
# sweeps' 16-byte vectors, halo vectors and wavefront overlaps fall at every
# offset, in FP32 and FP64
WALLED = [
    pytest.param(((70, 9, 40), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.3, 1.5)), id="walls-70x9x40"),
    pytest.param(((131, 23, 15), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 1.2)), id="walls-131x23x15"),
    pytest.param(((261, 9, 13), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.1, 0.0)), id="walls-261x9x13"),
]


@pytest.mark.parametrize("spec", GRIDS + MULTI + WALLED)
def test_bicgstab_mixed(spec):
    oc, gc = pair(spec)
    xt = rnd(oc.shape, 190)
    bb = oc.zeros()
    oc.product_div_grad(xt.copy(), bb, oc.zeros())
    phi = np.ascontiguousarray(0.1 * rnd(oc.shape, 191))
    gc.upload(ih.RES, bb)
    gc.upload(ih.PRE, phi)
    eps = 1e-9 * np.abs(bb).max()
    so = oc.bicgstab_mixed(eps, 500, oc.Sp, oc.Np, bb, phi, 3, 0.8)
    sg = gc.bicgstab_mixed(eps, 500, oc.Sp, oc.Np, 3, 0.8)
    assert so.iterations == sg.iterations and so.exit_reason == sg.exit_reason == 0
    assert so.norm_inf_final == sg.norm_inf_final
    same(gc.download(ih.PRE), phi, "mixed bicgstab phi")


@pytest.mark.parametrize("sweeps", [1, 2, 5])
def test_bicgstab_mixed_sweep_counts(sweeps):
    """M = 1, 2 and 5 FP32 sweeps: one sweep is its own kernel; from two on the
    p and s updates form the first sweep, with no middle sweep at two and an
    odd number of them at five.  Bit for bit against the oracle."""
    oc, gc = pair(WALLED[0].values[0])
    xt = rnd(oc.shape, 195)
    bb = oc.zeros()
    oc.product_div_grad(xt.copy(), bb, oc.zeros())
    phi = np.ascontiguousarray(0.1 * rnd(oc.shape, 196))
    gc.upload(ih.RES, bb)
    gc.upload(ih.PRE, phi)
    eps = 1e-9 * np.abs(bb).max()
    so = oc.bicgstab_mixed(eps, 3000, oc.Sp, oc.Np, bb, phi, sweeps, 0.8)
    sg = gc.bicgstab_mixed(eps, 3000, oc.Sp, oc.Np, sweeps, 0.8)
    assert so.iterations == sg.iterations and so.exit_reason == sg.exit_reason == 0
    assert so.norm_inf_final == sg.norm_inf_final
    same(gc.download(ih.PRE), phi, f"mixed bicgstab phi, {sweeps} sweeps")


def _tube_case(mixed):
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
              level_pressure=1, inflow_kind=2, inflow_amp=1.0, jor_fp32=1 if mixed else 0,
              precond_sweeps=4 if mixed else 0)
    po = O.StepParams(mask3=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), **kw)
    pg = ih.StepParams(use_mask=1, **kw)
    return oc, gc, vel, pre, nl, po, pg, mask


def test_timestep_mixed_bit_exact():
    oc, gc, vel, pre, nl, po, pg, mask = _tube_case(True)
    t = 0.0
    for step in range(2):
        so = oc.timestep(po, step, t, vel, pre, nl)
        sg = gc.timestep(pg, t)
        t += 5e-4
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        compare_state(oc, gc, vel, pre, nl, f"mixed step {step}")


def test_mixed_against_fp64():
    """Two steps of the stenosed tube, FP64 and mixed, from the same state:
    velocity within 1e-8 (relative L2), pressure within 1e-5; BiCGstab needs
    fewer than half the iterations with the FP32 preconditioner."""
    res = {}
    for mixed in (False, True):
        oc, gc, vel, pre, nl, po, pg, mask = _tube_case(mixed)
        its = []
        t = 0.0
        for step in range(2):
            its += list(gc.timestep(pg, t).bicg_iterations)
            t += 5e-4
        res[mixed] = ([gc.download(ih.VEL1 + q) for q in range(3)], gc.download(ih.PRE), sum(its))
    (v64, p64, n64), (v32, p32, n32) = res[False], res[True]
    ref = np.linalg.norm(v64[0][box(oc.SB[0], oc.NB[0])])
    for q in range(3):
        s = box(oc.SB[q], oc.NB[q])
        assert np.linalg.norm(v32[q][s] - v64[q][s]) <= 1e-8 * ref, q
    s = box(oc.Sp, oc.Np)
    assert np.linalg.norm(p32[s] - p64[s]) <= 1e-5 * np.linalg.norm(p64[s])
    assert 2 * n32 < n64, (n32, n64)


@pytest.mark.parametrize("spec", GRIDS + MULTI + WALLED)
def test_bicgstab_jor64(spec):
    """The like-for-like FP64 baseline: the same right-preconditioned BiCGstab
    with M = FP64 JOR sweeps, bit for bit against the oracle."""
    oc, gc = pair(spec)
    xt = rnd(oc.shape, 192)
    bb = oc.zeros()
    oc.product_div_grad(xt.copy(), bb, oc.zeros())
    phi = np.ascontiguousarray(0.1 * rnd(oc.shape, 193))
    gc.upload(ih.RES, bb)
    gc.upload(ih.PRE, phi)
    eps = 1e-9 * np.abs(bb).max()
    so = oc.bicgstab_jor64(eps, 500, oc.Sp, oc.Np, bb, phi, 3, 0.8)
    sg = gc.bicgstab_jor64(eps, 500, oc.Sp, oc.Np, 3, 0.8)
    assert so.iterations == sg.iterations and so.exit_reason == sg.exit_reason == 0
    assert so.norm_inf_final == sg.norm_inf_final
    same(gc.download(ih.PRE), phi, "FP64-preconditioned bicgstab phi")


def test_timestep_fp64_preconditioned_bit_exact():
    oc, gc, vel, pre, nl, po, pg, mask = _tube_case(False)
    po.precond_sweeps = pg.precond_sweeps = 4
    po.precond_fp64 = pg.precond_fp64 = 1
    t = 0.0
    for step in range(2):
        so = oc.timestep(po, step, t, vel, pre, nl)
        sg = gc.timestep(pg, t)
        t += 5e-4
        assert list(so.bicg_iterations) == list(sg.bicg_iterations)
        compare_state(oc, gc, vel, pre, nl, f"FP64-preconditioned step {step}")


@pytest.mark.parametrize("spec", GRIDS)
def test_nan_residual_stops_as_breakdown(spec):
    """One NaN in the right-hand side: the max-norm of the residual is NaN, and
    BiCGstab (plain, FP32-preconditioned, FP64-preconditioned) stops at its first
    stop test with a breakdown (exit 2) instead of reading the NaN residual as a
    small one and reporting convergence.  A mixed time step relies on that exit
    to redo the solve in FP64."""
    oc, gc = pair(spec)
    bb = rnd(oc.shape, 194)
    bb[tuple(n // 2 for n in oc.shape)] = np.nan
    phi = oc.zeros()
    gc.upload(ih.RES, bb)
    gc.upload(ih.PRE, phi)
    assert np.isnan(gc.residual_cdg(oc.Sp, oc.Np))
    for solve in (lambda: gc.bicgstab(1e-8, 50, oc.Sp, oc.Np),
                  lambda: gc.bicgstab_mixed(1e-8, 50, oc.Sp, oc.Np, 3, 0.8),
                  lambda: gc.bicgstab_jor64(1e-8, 50, oc.Sp, oc.Np, 3, 0.8)):
        gc.upload(ih.PRE, phi)
        st = solve()
        assert st.exit_reason == 2 and st.iterations == 0, (st.exit_reason, st.iterations)
