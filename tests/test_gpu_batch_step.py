"""Batched time steps (ih_batch_timestep): many stenosed tubes of uneven grid
size, each with its own radius, stenosis and IBM mask, advanced together; the
Poisson solves of all cases share their launches.  Every case must end each
step bit-identical to the same case advanced by its own ih_timestep
(src/mod_timeint.f90:275-377 per instance), in the mixed-precision form (FP32
JOR sweeps, BiCGstab right-preconditioned by FP32 sweeps) and in plain FP64;
the smallest case is also checked against the oracle."""
import ctypes
import math

import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O
from test_gpu_parity import compare_state

pytestmark = pytest.mark.gpu


def radius_profile(x, R, stenosis, x0=1.0, ell=0.4):
    """cosine constriction removing `stenosis` of the area at x0"""
    delta = 1.0 - math.sqrt(1.0 - stenosis)
    r = np.full_like(x, R)
    s = np.abs(x - x0) <= ell
    r[s] = R * (1.0 - delta / 2.0 * (1.0 + np.cos(np.pi * (x[s] - x0) / ell)))
    return r


def tube(N, R, sten, bc=(1,) * 6):
    c = ih.Case(N, bc, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    c.set_tube_mask(0.5, 0.5, radius_profile(c.coords(0, 1), R, sten), radius_profile(c.coords(0, 0), R, sten))
    c.fill(ih.VEL1, 1.0)
    c.apply_mask()
    return c


# config 4's sizes (cells 192x96x96, 128x64x64, 96^3, 64^3) and a z-periodic case
CASES = [((193, 97, 97), 0.3, 0.5, (1,) * 6),
         ((129, 65, 65), 0.22, 0.7, (1,) * 6),
         ((97, 97, 97), 0.35, 0.3, (1,) * 6),
         ((65, 65, 65), 0.25, 0.45, (1,) * 6),
         ((49, 33, 34), 0.3, 0.5, (1, 1, 1, 1, -1, -1))]


def _state(c):
    return [c.download(f) for f in (ih.VEL1, ih.VEL2, ih.VEL3, ih.PRE, ih.NL1, ih.NL2, ih.NL3)]


@pytest.mark.parametrize("mixed", [True, False], ids=["mixed", "fp64"])
def test_batched_steps_equal_single_case_steps(mixed):
    kw = dict(Re=500.0, dtime=5e-4, jor_sweeps=40, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=4000,
              level_pressure=1, inflow_kind=2, inflow_amp=1.0, use_mask=1, jor_fp32=int(mixed),
              precond_sweeps=8 if mixed else 0)
    p = ih.StepParams(**kw)
    print(f"batch step: {len(CASES)} cases, mixed={mixed}", flush=True)
    batched = [tube(N, R, s, bc) for N, R, s, bc in CASES]
    single = [tube(N, R, s, bc) for N, R, s, bc in CASES]
    bt = ih.Batch(batched)
    t = 0.0
    for step in range(2):
        sb = bt.timestep(p, t)
        for q, c in enumerate(single):
            ss = c.timestep(p, t)
            assert list(sb[q].bicg_iterations) == list(ss.bicg_iterations), (q, step)
            assert list(sb[q].bicg_final) == list(ss.bicg_final), (q, step)
            assert sb[q].div_max_after == ss.div_max_after, (q, step)
            for a, b in zip(_state(batched[q]), _state(c)):
                assert np.array_equal(a, b), (q, step)
        t += 5e-4
    print("batch step: iterations", [list(s.bicg_iterations) for s in sb], flush=True)
    bt.close()
    for c in batched + single:
        c.close()


@pytest.mark.parametrize("sweeps", [1, 2])
def test_batched_preconditioner_sweep_counts(sweeps):
    """M = 1 FP32 sweep (its own batched kernel) and 2 (the batched p and s
    updates form the first sweep, no middle sweep): each case of the batch
    bit-identical to its own step."""
    kw = dict(Re=500.0, dtime=5e-4, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=4000,
              level_pressure=1, inflow_kind=2, inflow_amp=1.0, use_mask=1, jor_fp32=1, precond_sweeps=sweeps)
    p = ih.StepParams(**kw)
    spec = CASES[3:]
    print(f"batch step: {len(spec)} cases, {sweeps} preconditioner sweeps", flush=True)
    batched = [tube(N, R, s, bc) for N, R, s, bc in spec]
    single = [tube(N, R, s, bc) for N, R, s, bc in spec]
    bt = ih.Batch(batched)
    sb = bt.timestep(p, 0.0)
    for q, c in enumerate(single):
        ss = c.timestep(p, 0.0)
        assert list(sb[q].bicg_iterations) == list(ss.bicg_iterations), q
        assert list(sb[q].bicg_final) == list(ss.bicg_final), q
        for a, b in zip(_state(batched[q]), _state(c)):
            assert np.array_equal(a, b), q
    bt.close()
    for c in batched + single:
        c.close()


def test_batched_step_matches_the_oracle():
    """the smallest tube of a batch against the oracle's step"""
    N, R, sten, bc = (33, 17, 17), 0.3, 0.5, (1,) * 6
    spec = (N, bc, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    oc = O.OracleCase(*spec)
    small = tube(N, R, sten, bc)
    other = tube((65, 33, 33), 0.25, 0.6)
    mask = oc.tube_mask(0.5, 0.5, radius_profile(oc.coords(0, 1), R, sten), radius_profile(oc.coords(0, 0), R, sten))
    assert np.array_equal(small.get_mask(), mask)
    vel = oc.zeros(3)
    vel[0] = mask[0].astype(np.float64)
    vel = np.ascontiguousarray(vel)
    pre, nl = oc.zeros(), oc.zeros(3)
    for q in range(3):
        small.upload(ih.VEL1 + q, vel[q])
        small.upload(ih.NL1 + q, nl[q])
    small.upload(ih.PRE, pre)
    kw = dict(Re=500.0, dtime=5e-4, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=2000,
              level_pressure=1, inflow_kind=2, inflow_amp=1.0)
    po = O.StepParams(mask3=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), **kw)
    pg = ih.StepParams(use_mask=1, jor_fp32=0, **kw)
    bt = ih.Batch([other, small])
    sb = bt.timestep(pg, 0.0)
    so = oc.timestep(po, 0, 0.0, vel, pre, nl)
    assert list(sb[1].bicg_iterations) == list(so.bicg_iterations)
    compare_state(oc, small, vel, pre, nl, "batched tube 33x17x17")
    bt.close()
    small.close()
    other.close()
