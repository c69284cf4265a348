"""A Fortran host on the HDF5 layer: tests/fortran/io_driver.f90, built with
amdflang against include/impact_io_c.f90, calls ih_io_write_xdmf,
ih_io_write_fields, ih_io_write_restart and ih_io_read_restart in the order of
IMPACT's time loop.  The files it writes are read back with tests/h5read.py;
the fields it reads back and the run state it ends with are compared with
what the reference's routines give for the same inputs."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import h5read
import oracle_lib as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "fortran" / "_build" / "io_driver"
N = (17, 13, 11)


def pattern(s):
    """mod(7(i+3) + 13(j+3) + 29(k+3) + s, 97)/32 - 1.5 on (-3:N+3)^3, numpy order (k, j, i)."""
    k, j, i = np.meshgrid(np.arange(N[2] + 7), np.arange(N[1] + 7), np.arange(N[0] + 7), indexing="ij")
    return np.ascontiguousarray(((7 * i + 13 * j + 29 * k + s) % 97) / 32.0 - 1.5)


def box(lo, hi):
    return tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))


def test_fortran_host_hdf5_io(tmp_path):
    if not DRIVER.exists():
        pytest.skip("Fortran I/O driver not built (amdflang or libimpact_io missing at build time)")
    print(f"running {DRIVER.name}", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    oc = O.OracleCase(N, (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))

    def back(name):
        return np.fromfile(tmp_path / f"{name}.bin", dtype=np.float64).reshape(oc.shape)

    # read_restart restored every field over its range and nothing else
    for m in range(3):
        got, want = back(f"back_vel{m + 1}"), pattern(61 + m)
        s = box(oc.SB[m], oc.NB[m])
        assert np.array_equal(got[s], want[s]), f"velocity {m + 1}"
        rest = np.ones(oc.shape, dtype=bool)
        rest[s] = False
        assert not got[rest].any(), f"velocity {m + 1} outside its range"
    sp = box(oc.Sp, oc.Np)
    assert np.array_equal(back("back_pre")[sp], pattern(64)[sp])

    # write_fields: velocity m after exchange(m, m, .) interpolated to the pressure
    # grid, the medium size its stride-2 subsample; the pressure as it is
    for m, vn in enumerate(("velX", "velY", "velZ")):
        v, inter = pattern(61 + m), oc.zeros()
        oc.exchange(m + 1, m + 1, v)
        oc.interpolate2_vel_pre(m + 1, v, inter)
        with h5read.File(tmp_path / f"{vn}_large_00002.h5") as f:
            big = f.dataset(vn)[0]
        with h5read.File(tmp_path / f"{vn}_med_00002.h5") as f:
            med = f.dataset(vn)[0]
        assert np.array_equal(big, inter[sp]), vn
        assert np.array_equal(med, big[::2, ::2, ::2]), vn
    with h5read.File(tmp_path / "pre_large_00002.h5") as f:
        assert np.array_equal(f.dataset("pre")[0], pattern(64)[sp])
    for tag in ("large", "med"):
        assert (tmp_path / f"IMPACT_xdmf3d_{tag}_00002.xmf").exists()
    assert sorted(p.name for p in tmp_path.glob("*_restart005.h5")) == [
        "pre_restart005.h5", "velX_restart005.h5", "velY_restart005.h5", "velZ_restart005.h5"]

    # the run state: write_fields advanced the count and the output time;
    # read_restart restored the state and re-based the output times on the
    # configured intervals (src/mod_inout.f90:452-487)
    si = np.fromfile(tmp_path / "state_i.bin", dtype=np.int32)
    assert list(si) == [3, 0, 5, 3, 100]
    sr = np.fromfile(tmp_path / "state_r.bin", dtype=np.float64)
    tov = 0.6 + 0.1
    assert sr[0] == tov
    assert sr[1] == 0.5 and sr[2] == 1e-3
    assert sr[3] == (tov - 0.1) + 0.05
    assert sr[4] == (0.7 - 0.2) + 0.4
    assert sr[5] == 2.0
