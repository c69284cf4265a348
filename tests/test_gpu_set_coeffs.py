"""ih_case_set_coeffs, the path by which IMPACT's host uploads the coefficient
tables it computed itself (GPU_memcopy, src/mod_mem_GPU.f90:348-434), fed
the oracle's tables in the reference layout c(lo:hi, 0:N): every table reads
back unchanged and a full stenosed-tube step equals the oracle's step bit for
bit."""
import ctypes
import math

import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O
from test_gpu_parity import compare_state

pytestmark = pytest.mark.gpu

SETS = {
    "cGp": ("cGp1", "cGp2", "cGp3"), "cD": ("cDu1", "cDv2", "cDw3"), "cIp": ("cIpu", "cIpv", "cIpw"),
    "cIu": ("cIup", "cIvp", "cIwp"), "cNpD": ("cNp1D", "cNp2D", "cNp3D"), "cNpU": ("cNp1U", "cNp2U", "cNp3U"),
    "cNuD": ("cNu1D", "cNv2D", "cNw3D"), "cNuU": ("cNu1U", "cNv2U", "cNw3U"), "cp2": ("cp11", "cp22", "cp33"),
    "cu2": ("cu11", "cv22", "cw33"), "cdg": ("cdg1", "cdg2", "cdg3"),
}


def test_set_coeffs_with_the_oracle_tables_then_a_step():
    spec = ((33, 17, 17), (1,) * 6, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    oc = O.OracleCase(*spec)
    gc = ih.Case(*spec, generate=False)
    keep = {}
    cs = ih.CoeffSet()
    for field, names in SETS.items():
        arr = getattr(cs, field)
        for d, name in enumerate(names):
            keep[name] = np.ascontiguousarray(oc.coeff(name))
            arr[d] = keep[name].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ih.check(gc.lib.ih_case_set_coeffs(gc.h, ctypes.byref(cs)), "set_coeffs")
    for name, want in keep.items():
        assert np.array_equal(gc.coeff(name), want), name

    def radius(x, R=0.3, x0=1.0, ell=0.4, delta=1.0 - 1.0 / math.sqrt(2.0)):
        r = np.full_like(x, R)
        s = np.abs(x - x0) <= ell
        r[s] = R * (1.0 - delta / 2.0 * (1.0 + np.cos(np.pi * (x[s] - x0) / ell)))
        return r

    r_u, r_p = radius(oc.coords(0, 1)), radius(oc.coords(0, 0))
    mask = oc.tube_mask(0.5, 0.5, r_u, r_p)
    gc.set_tube_mask(0.5, 0.5, r_u, r_p)
    assert np.array_equal(gc.get_mask(), mask)
    vel = oc.zeros(3)
    vel[0] = mask[0].astype(np.float64)
    vel = np.ascontiguousarray(vel)
    pre, nl = oc.zeros(), oc.zeros(3)
    for q in range(3):
        gc.upload(ih.VEL1 + q, vel[q])
        gc.upload(ih.NL1 + q, nl[q])
    gc.upload(ih.PRE, pre)
    kw = dict(Re=500.0, dtime=5e-4, jor_sweeps=20, jor_omega=0.8, bicg_rel_tol=1e-8, bicg_max_it=2000,
              level_pressure=1, inflow_kind=2, inflow_amp=1.0)
    so = oc.timestep(O.StepParams(mask3=mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), **kw), 0, 0.0, vel,
                     pre, nl)
    sg = gc.timestep(ih.StepParams(use_mask=1, jor_fp32=0, **kw), 0.0)
    assert list(sg.bicg_iterations) == list(so.bicg_iterations)
    compare_state(oc, gc, vel, pre, nl, "step on uploaded tables")
    gc.close()
