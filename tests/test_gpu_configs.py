"""Parity at the sizes the benchmark configurations state.

- 65^3 channel (walls, parabolic inflow 16 y(1-y) z(1-z), seeded N(0, 0.01)
  noise, Re 100, dt 1e-3, 200 JOR sweeps + BiCGstab to 1e-8): three full steps
  bit for bit against the oracle, then twenty steps on the device with the
  projection's own invariant: after the correction, div u = b - A p, so the
  largest |div u| equals the final BiCGstab residual (src/mod_test.f90:304-352
  checks the same quantity).
- 257x129x129 stenosed tube (tanh 1.5 in y, z; radius 0.3, 50 % area cosine
  stenosis; pulsatile plug inflow; Re 500, dt 5e-4): the IBM mask bytes and one
  full step (200 JOR sweeps, BiCGstab bounded to a few iterations a substep so
  the oracle finishes in seconds) bit for bit, in FP64 and in the mixed form
  (FP32 sweeps and FP32-swept preconditioner, FP64 residuals and updates).
- 385x193x193 tube through the serving loop: the served result equals the same
  run on a case of its own.

The oracle runs its OpenMP build (identical results) on the host's cores.
"""
import math

import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O
from test_gpu_parity import box, compare_state, same

pytestmark = pytest.mark.gpu


def _channel_state(oc, seed=2026):
    rng = np.random.default_rng(seed)
    yp, zp = oc.coords(1, 0), oc.coords(2, 0)
    vel = oc.zeros(3)
    vel[0] = (16 * yp[None, :, None] * (1 - yp[None, :, None]) * zp[:, None, None] * (1 - zp[:, None, None])
              * np.ones(oc.shape))
    for q in range(3):
        vel[q] += rng.normal(0.0, 0.01, size=oc.shape)
    return np.ascontiguousarray(vel)


def test_channel_65_cubed_three_steps_exact_then_twenty():
    spec = ((65, 65, 65), (1,) * 6, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    print("config: 65^3 channel, 3 steps against the oracle", flush=True)
    oc = O.OracleCase(*spec)
    gc = ih.Case(*spec)
    vel = _channel_state(oc)
    pre, nl = oc.zeros(), oc.zeros(3)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.NL1 + q, nl[q])
    gc.upload(ih.PRE, pre)
    kw = dict(Re=100.0, dtime=1e-3, jor_sweeps=200, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=20000,
              level_pressure=1, inflow_kind=1, inflow_amp=1.0)
    po = O.StepParams(**kw)
    pg = ih.StepParams(use_mask=0, **kw)
    t = 0.0
    for step in range(20):
        sg = gc.timestep(pg, t)
        if step < 3:
            so = oc.timestep(po, step, t, vel, pre, nl)
            assert list(sg.bicg_iterations) == list(so.bicg_iterations)
            assert list(sg.bicg_final) == list(so.bicg_final)
            assert sg.div_max_after == so.div_max_after
            compare_state(oc, gc, vel, pre, nl, f"channel step {step}")
        t += 1e-3
        # every solve converged, within a bounded number of iterations
        assert all(0 < n < 2000 for n in sg.bicg_iterations), list(sg.bicg_iterations)
        # div u after the last substep's correction is that solve's residual
        assert sg.div_max_after == pytest.approx(sg.bicg_final[2], rel=1e-4)
        assert sg.div_max_after < 1e-8
    print(f"config: 20 steps, last iterations {list(sg.bicg_iterations)}, max|div u| {sg.div_max_after:.3e}",
          flush=True)


def _tube_radius(x, R=0.3, x0=1.0, ell=0.4, delta=1.0 - 1.0 / math.sqrt(2.0)):
    r = np.full_like(x, R)
    s = np.abs(x - x0) <= ell
    r[s] = R * (1.0 - delta / 2.0 * (1.0 + np.cos(np.pi * (x[s] - x0) / ell)))
    return r


TUBE = ((257, 129, 129), (1,) * 6, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_stenosed_tube_full_size_step(mixed):
    print(f"config: 257x129x129 stenosed tube, one step, mixed={mixed}", flush=True)
    import ctypes
    oc = O.OracleCase(*TUBE)
    gc = ih.Case(*TUBE)
    r_u, r_p = _tube_radius(oc.coords(0, 1)), _tube_radius(oc.coords(0, 0))
    mask = oc.tube_mask(0.5, 0.5, r_u, r_p)
    gc.set_tube_mask(0.5, 0.5, r_u, r_p)
    same(gc.get_mask(), mask, "IBM mask bytes at 257x129x129")
    # the tube's cross-section holds about pi R^2 / (dy dz) points at every x
    vel = oc.zeros(3)
    vel[0] = mask[0].astype(np.float64)
    vel = np.ascontiguousarray(vel)
    pre, nl = oc.zeros(), oc.zeros(3)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.NL1 + q, nl[q])
    gc.upload(ih.PRE, pre)
    kw = dict(Re=500.0, dtime=5e-4, jor_sweeps=200, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=3,
              level_pressure=1, inflow_kind=2, inflow_amp=1.0, jor_fp32=int(mixed), precond_sweeps=8 if mixed else 0)
    po = O.StepParams(mask3=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), **kw)
    pg = ih.StepParams(use_mask=1, **kw)
    so = oc.timestep(po, 0, 0.0, vel, pre, nl)
    sg = gc.timestep(pg, 0.0)
    assert list(sg.bicg_iterations) == list(so.bicg_iterations)
    assert list(sg.bicg_final) == list(so.bicg_final)
    compare_state(oc, gc, vel, pre, nl, f"tube 257x129x129 mixed={mixed}")
    print(f"config: iterations {list(sg.bicg_iterations)}", flush=True)


def test_served_385_request_equals_its_own_run():
    """One 384x192x192-cell request through ih_serve (a pool slot, reset and
    reused) gives the checksums of the same run on a case of its own."""
    print("config: 385x193x193 served request against a direct run", flush=True)
    N = (385, 193, 193)
    step = ih.StepParams(Re=500.0, dtime=5e-4, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=40,
                         level_pressure=1, inflow_kind=2, inflow_amp=1.0, use_mask=1)
    req = [dict(arrival_s=0.0, N=N, steps=2, radius=0.3, stenosis=0.5),
           dict(arrival_s=0.0, N=N, steps=2, radius=0.3, stenosis=0.5)]
    res, summ, rc = ih.serve(req, step, hbm_budget=1 << 36, max_active=1)
    assert rc == 0 and all(r.status == 0 for r in res)
    assert summ.pool_slots == 1     # the second request reused the first one's slot
    c = ih.Case(N, (1,) * 6, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    delta = 1.0 - math.sqrt(1.0 - 0.5)

    def radius(xs):  # the serving loop's profile, with the C library's cosine (math.cos)
        return np.array([0.3 if abs(x - 1.0) > 0.4 else
                         0.3 * (1.0 - delta / 2.0 * (1.0 + math.cos(math.pi * (x - 1.0) / 0.4))) for x in xs])

    c.set_tube_mask(0.5, 0.5, radius(c.coords(0, 1)), radius(c.coords(0, 0)))
    c.fill(ih.VEL1, 1.0)
    c.apply_mask()
    its = 0
    for s in range(2):
        st = c.timestep(step, s * 5e-4)
        its += sum(st.bicg_iterations)
    pb = (c.Sp, c.Np)
    assert (res[0].checksum_vel, res[0].checksum_pre, res[0].bicg_iterations) == (
        c.dot(*pb, ih.VEL1, ih.VEL1), c.dot(*pb, ih.PRE, ih.PRE), its)
    assert (res[1].checksum_vel, res[1].checksum_pre) == (res[0].checksum_vel, res[0].checksum_pre)
    c.close()
