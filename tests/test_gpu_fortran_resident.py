"""Residency registry through the reference's names: tests/fortran/resident_driver.f90
runs one projection substep with the drop-in device_* subroutines twice from
the same inputs, first as the reference's wrappers work (every array over PCIe
on every call), then with the module arrays bound to resident fields
(ih_bind_field) and one ih_sync_all at the end.  Both passes must equal the
oracle's substep bit for bit, and the bound pass must move each array at most
once each way."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_fortran import pattern

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "fortran" / "_build" / "resident_driver"
N = (17, 13, 11)
BOUND_ARRAYS = 16  # vel, work, nl, rhs (3 each), pp, pre, div, gpre


def test_substep_through_reference_names_resident(tmp_path):
    if not DRIVER.exists():
        pytest.skip("Fortran driver not built (amdflang missing at build time)")
    print(f"running {DRIVER.name}", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    def got(name):
        return np.fromfile(tmp_path / f"{name}.bin", dtype=np.float64).reshape(N[2] + 7, N[1] + 7, N[0] + 7)

    # the oracle's substep on the same inputs
    oc = O.OracleCase(N, (1,) * 6, (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
    vel = np.ascontiguousarray(np.stack([pattern(oc, 60 + q) for q in (1, 2, 3)]))
    work = np.ascontiguousarray(np.stack([pattern(oc, 63 + q) for q in (1, 2, 3)]))
    nl = np.ascontiguousarray(np.stack([pattern(oc, 66 + q) for q in (1, 2, 3)]))
    rhs = np.ascontiguousarray(np.stack([pattern(oc, 69 + q) for q in (1, 2, 3)]))
    pp, div, gpre = pattern(oc, 73), pattern(oc, 74), pattern(oc, 75)
    pre = oc.zeros()
    dta = -1.0 / 64.0
    for d in range(3):
        oc.interpolate2_vel_pre(d + 1, vel[d], work[d])
    oc.nonlinear(vel, work, nl, pp)
    oc.helmholtz(1.0 / 137.0, vel, nl)
    for q in range(3):
        rhs[q] = 1.0 * vel[q] + dta * nl[q]
    oc.divergence2(rhs, div)
    rc, sweeps = O.device_jacobi_solver(oc, 0.8, oc.Sp, oc.Np, pre, div)
    assert rc == 0 and sweeps > 0
    for q in range(3):
        oc.gradient(q + 1, pre, gpre)
        vel[q] = 1.0 * rhs[q] + (-1.0) * gpre

    for tag in ("plain", "bound"):
        for q in range(3):
            assert np.array_equal(got(f"{tag}_vel{q + 1}"), vel[q]), f"{tag} vel{q + 1}"
            assert np.array_equal(got(f"{tag}_nl{q + 1}"), nl[q]), f"{tag} nl{q + 1}"
            assert np.array_equal(got(f"{tag}_rhs{q + 1}"), rhs[q]), f"{tag} rhs{q + 1}"
        assert np.array_equal(got(f"{tag}_pre"), pre), f"{tag} pre"
        assert np.array_equal(got(f"{tag}_div"), div), f"{tag} div"

    lines = (tmp_path / "transfers.txt").read_text().split("\n")
    (up_plain, down_plain), (up_bound, down_bound) = [tuple(int(v) for v in ln.split()) for ln in lines[:2]]
    print(f"transfers: plain {up_plain} up / {down_plain} down, bound {up_bound} up / {down_bound} down", flush=True)
    assert up_bound <= BOUND_ARRAYS and down_bound <= BOUND_ARRAYS
    assert up_plain > 3 * up_bound and down_plain > 2 * down_bound
