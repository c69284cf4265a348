"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
LIB_PATH = ORACLE_DIR / "_build" / "liboracle.so"
OMP_LIB_PATH = ORACLE_DIR / "_build" / "liboracle_omp.so"  # same code, point loops on OpenMP threads

_lib = None
# The OpenMP build runs the point loops on the host's threads; every point is
# written by one thread and sums are not split, so its results are the serial
# build's bit for bit.  It is used when built (the full-size parity tests need it).
_path = OMP_LIB_PATH if OMP_LIB_PATH.exists() else LIB_PATH


def use_parallel_build():
    """Load the OpenMP build instead (identical results, for timing the CPU
    baseline); must be called before the first OracleCase."""
    global _path
    if _lib is not None and _path != OMP_LIB_PATH:
        raise RuntimeError("the serial oracle is already loaded")
    _path = OMP_LIB_PATH


def build(force: bool = False) -> Path:
    if force or not _path.exists():
        subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return _path


def lib():
    global _lib
    if _lib is None:
        if not _path.exists():
            build()
        L = C.CDLL(str(_path))
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        vp = C.c_void_p
        L.orc_case_create.restype = vp
        L.orc_case_create.argtypes = [ip, ip, dp, dp]
        L.orc_case_destroy.argtypes = [vp]
        L.orc_field_size.restype = C.c_int64
        L.orc_field_size.argtypes = [vp]
        L.orc_ranges.argtypes = [vp, ip]
        L.orc_coords.argtypes = [vp, C.c_int, C.c_int, dp]
        L.orc_coeff.restype = C.c_int
        L.orc_coeff.argtypes = [vp, C.c_char_p, dp]
        L.orc_exact_table.restype = C.c_int
        L.orc_exact_table.argtypes = [C.c_int, C.c_int, C.c_double, dp]
        L.orc_exchange.argtypes = [vp, C.c_int, C.c_int, dp]
        L.orc_exchange_all_all.argtypes = [vp, dp]
        L.orc_exchange_relax.argtypes = [vp, dp]
        L.orc_divergence2.argtypes = [vp, dp, dp]
        L.orc_gradient.argtypes = [vp, C.c_int, dp, dp]
        L.orc_interpolate2_vel_pre.argtypes = [vp, C.c_int, dp, dp]
        L.orc_interpolate2_pre_vel.argtypes = [vp, C.c_int, dp, dp]
        L.orc_interpolate_vel.argtypes = [vp, dp, dp]
        L.orc_nonlinear.argtypes = [vp, dp, dp, dp, dp]
        L.orc_helmholtz_explicit.argtypes = [vp, C.c_double, dp, dp]
        L.orc_product_div_grad.argtypes = [vp, dp, dp, dp]
        L.orc_set_reduction_mode.argtypes = [vp, C.c_int]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [vp, ip, ip, dp, dp]
        L.orc_jor.argtypes = [vp, C.c_int, C.c_double, ip, ip, dp, dp, dp]
        L.orc_residual_cdg.restype = C.c_double
        L.orc_residual_cdg.argtypes = [vp, ip, ip, dp, dp]
        L.orc_bicgstab.argtypes = [vp, C.c_double, C.c_int, ip, ip, dp, dp, dp, C.POINTER(BicgStats)]
        L.orc_bicgstab_mixed.argtypes = [vp, C.c_double, C.c_int, ip, ip, dp, dp, C.c_int, C.c_double,
                                         C.POINTER(BicgStats)]
        L.orc_jor_mixed.argtypes = [vp, C.c_int, C.c_double, ip, ip, dp, dp]
        L.orc_timestep.argtypes = [vp, C.POINTER(StepParams), C.c_int, C.c_double, dp, dp, dp, C.POINTER(StepStats)]
        L.orc_tube_mask.argtypes = [vp, C.c_double, C.c_double, dp, dp, C.POINTER(C.c_uint8)]
        L.orc_apply_mask.argtypes = [vp, C.POINTER(C.c_uint8), dp]
        L.orc_level_pressure.argtypes = [vp, dp]
        L.orc_fill_corners.argtypes = [vp, dp]
        L.orc_inflow_amplitude.restype = C.c_double
        L.orc_inflow_amplitude.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_mg_create.restype = vp
        L.orc_mg_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_mg_destroy.argtypes = [vp]
        L.orc_mg_grids.restype = C.c_int
        L.orc_mg_grids.argtypes = [vp]
        L.orc_mg_levels.restype = C.c_int
        L.orc_mg_levels.argtypes = [vp, ip]
        L.orc_mg_coeff.restype = C.c_int
        L.orc_mg_coeff.argtypes = [vp, C.c_int, C.c_char_p, dp]
        L.orc_mg_field_size.restype = C.c_int64
        L.orc_mg_field_size.argtypes = [vp, C.c_int]
        L.orc_mg_vcycle.argtypes = [vp, dp, dp]
        L.orc_mg_restrict_residual.argtypes = [vp, C.c_int, dp, dp, dp]
        L.orc_mg_interpolate_add.argtypes = [vp, C.c_int, dp, dp]
        L.orc_mg_relax.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_bicgstab_mg.argtypes = [vp, vp, C.c_double, C.c_int, ip, ip, dp, dp, C.POINTER(BicgStats)]
        L.orc_bicgstab_mg_f32.argtypes = [vp, vp, C.c_double, C.c_int, ip, ip, dp, dp, C.POINTER(BicgStats)]
        L.orc_mg_vcycle_f32.argtypes = [vp, dp, dp]
        L.orc_vtk_read.restype = C.c_int
        L.orc_vtk_read.argtypes = [C.c_char_p, ip, C.POINTER(dp)]
        L.orc_waveform_convert.argtypes = [C.c_int, dp, dp, C.c_int, dp, dp, dp]
        L.orc_ct_build.restype = C.c_int
        L.orc_ct_build.argtypes = [vp, ip, dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.POINTER(C.c_uint8), ip, dp]
        L.orc_ct_inflow.restype = C.c_double
        L.orc_ct_inflow.argtypes = [C.c_int, dp, dp, C.c_double, ip]
        L.orc_ct_forcing.argtypes = [vp, C.POINTER(C.c_uint8), C.c_int, dp, dp, C.c_double, dp]
        L.orc_waveform_convert_branch.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, dp]
        L.orc_wk_pressure_flow.argtypes = [C.c_int, C.c_double, dp, dp, dp, dp]
        L.orc_wk_echo.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, dp, dp]
        L.orc_ct_apply_forcing.argtypes = [vp, C.POINTER(C.c_uint8), C.POINTER(CtState), C.c_double, dp]
        L.orc_set_pressure_bc.argtypes = [vp, C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_double, dp]
        L.orc_bicgstab_jor64.argtypes = [vp, C.c_double, C.c_int, ip, ip, dp, dp, C.c_int, C.c_double,
                                         C.POINTER(BicgStats)]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_device_jacobi_solver.restype = C.c_int
        L.orc_device_jacobi_solver.argtypes = [vp, C.c_double, ip, ip, dp, dp, C.POINTER(C.c_uint8), C.c_int,
                                               C.c_int, C.c_double, dp, C.c_int, ip]
        _lib = L
    return _lib


class BicgStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("norm_inf_initial", C.c_double),
                ("norm_inf_final", C.c_double), ("exit_reason", C.c_int)]


class CtState(C.Structure):
    """orc_ct_state: the module state of both apply_CT_forcing branches."""
    _fields_ = [("ct_based", C.c_int), ("idealized", C.c_int), ("npts", C.c_int),
                ("t_ms", C.POINTER(C.c_double)), ("q_raw", C.POINTER(C.c_double)),
                ("t_conv", C.POINTER(C.c_double)), ("q_conv", C.POINTER(C.c_double)),
                ("ref", C.c_double * 3), ("inlet_voxels", C.c_int), ("k_start", C.c_int),
                ("wk_state", C.c_double * 3), ("pres", C.c_double * 3)]


def ct_state(ct_based, idealized, t_ms, q_raw, t_conv, q_conv, ref, inlet_voxels, k_start):
    """a CtState over the given arrays (kept alive on the object)"""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (t_ms, q_raw, t_conv, q_conv)]
    st = CtState(ct_based=int(ct_based), idealized=int(idealized), npts=len(arrs[0]),
                 t_ms=_dp(arrs[0]), q_raw=_dp(arrs[1]), t_conv=_dp(arrs[2]), q_conv=_dp(arrs[3]),
                 inlet_voxels=int(inlet_voxels), k_start=int(k_start))
    for d in range(3):
        st.ref[d] = float(ref[d])
    st._keep = arrs
    return st


class StepParams(C.Structure):
    _fields_ = [("Re", C.c_double), ("dtime", C.c_double), ("jor_sweeps", C.c_int),
                ("jor_omega", C.c_double), ("bicg_rel_tol", C.c_double), ("bicg_max_it", C.c_int),
                ("level_pressure", C.c_int), ("inflow_kind", C.c_int), ("inflow_amp", C.c_double),
                ("mask3", C.POINTER(C.c_uint8)), ("jor_fp32", C.c_int), ("precond_sweeps", C.c_int),
                ("precond_poisson", C.c_int), ("mg_fp32", C.c_int), ("ct_lab", C.POINTER(C.c_uint8)),
                ("ct_npts", C.c_int), ("ct_t", C.POINTER(C.c_double)), ("ct_q", C.POINTER(C.c_double)),
                ("ct_state", C.POINTER(CtState)), ("pressure_bc", C.c_int), ("precond_fp64", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("bicg_iterations", C.c_int * 3), ("bicg_final", C.c_double * 3),
                ("div_max_after", C.c_double)]


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(seq):
    arr = (C.c_int * len(seq))(*seq)
    return arr


class OracleCase:
    """A single-block grid of the oracle with its coefficients."""

    def __init__(self, N, bc, L=(1.0, 1.0, 1.0), stretch=(0.0, 0.0, 0.0)):
        self.L = lib()
        self.N = tuple(int(n) for n in N)
        self.bc = tuple(int(b) for b in bc)
        Ld = (C.c_double * 3)(*L)
        Sd = (C.c_double * 3)(*stretch)
        self.h = self.L.orc_case_create(_ip(self.N), _ip(self.bc), Ld, Sd)
        if not self.h:
            raise ValueError("oracle rejected the case")
        self.shape = (self.N[2] + 7, self.N[1] + 7, self.N[0] + 7)  # numpy (k, j, i)
        self.size = int(self.L.orc_field_size(self.h))
        r = (C.c_int * 48)()
        self.L.orc_ranges(self.h, r)
        r = list(r)
        self.Sp, self.Np = r[0:3], r[3:6]
        self.SB = [r[6 + 3 * m:9 + 3 * m] for m in range(3)]
        self.NB = [r[15 + 3 * m:18 + 3 * m] for m in range(3)]
        self.SI = [r[24 + 3 * m:27 + 3 * m] for m in range(3)]
        self.NI = [r[33 + 3 * m:36 + 3 * m] for m in range(3)]
        self.SNF = (r[42:45], r[45:48])

    def __del__(self):
        try:
            if self.h:
                self.L.orc_case_destroy(self.h)
        except Exception:
            pass

    # --- fields -------------------------------------------------------------
    def zeros(self, n=1):
        if n == 1:
            return np.zeros(self.shape, dtype=np.float64)
        return np.zeros((n,) + self.shape, dtype=np.float64)

    @staticmethod
    def at(f, i, j, k):
        """Fortran-index view helper: f[k+3, j+3, i+3]."""
        return f[k + 3, j + 3, i + 3]

    def coords(self, d, which):
        out = np.zeros(self.N[d] + 7)
        self.L.orc_coords(self.h, d, which, _dp(out))
        return out

    def coeff(self, name):
        n = max(self.N) + 1
        buf = np.zeros(7 * n + 16)
        cnt = self.L.orc_coeff(self.h, name.encode(), _dp(buf))
        if cnt < 0:
            raise KeyError(name)
        return buf[:cnt].copy()

    # --- operators ----------------------------------------------------------
    def exchange(self, d, vel_dir, f):
        self.L.orc_exchange(self.h, d, vel_dir, _dp(f))

    def exchange_all_all(self, vel3):
        self.L.orc_exchange_all_all(self.h, _dp(vel3))

    def exchange_relax(self, f):
        self.L.orc_exchange_relax(self.h, _dp(f))

    def divergence2(self, vel3, div):
        self.L.orc_divergence2(self.h, _dp(vel3), _dp(div))

    def gradient(self, m, phi, grad):
        self.L.orc_gradient(self.h, m, _dp(phi), _dp(grad))

    def interpolate_vel(self, vel3, work3):
        self.L.orc_interpolate_vel(self.h, _dp(vel3), _dp(work3))

    def interpolate2_vel_pre(self, m, phi, inter):
        self.L.orc_interpolate2_vel_pre(self.h, m, _dp(phi), _dp(inter))

    def interpolate2_pre_vel(self, m, phi, inter):
        self.L.orc_interpolate2_pre_vel(self.h, m, _dp(phi), _dp(inter))

    def nonlinear(self, vel3, work3, nl3, pp):
        self.L.orc_nonlinear(self.h, _dp(vel3), _dp(work3), _dp(nl3), _dp(pp))

    def helmholtz(self, multL, vel3, nl3):
        self.L.orc_helmholtz_explicit(self.h, multL, _dp(vel3), _dp(nl3))

    def product_div_grad(self, phi, lap, dig):
        self.L.orc_product_div_grad(self.h, _dp(phi), _dp(lap), _dp(dig))

    def set_reduction_mode(self, mode):
        self.L.orc_set_reduction_mode(self.h, int(mode))

    def dot(self, ss, nn, a, b):
        return self.L.orc_dot(self.h, _ip(ss), _ip(nn), _dp(a), _dp(b))

    def jor(self, n, omega, ss, nn, rel, bb, comp):
        self.L.orc_jor(self.h, n, omega, _ip(ss), _ip(nn), _dp(rel), _dp(bb), _dp(comp))

    def residual_cdg(self, ss, nn, rel, bb):
        return self.L.orc_residual_cdg(self.h, _ip(ss), _ip(nn), _dp(rel), _dp(bb))

    def bicgstab(self, eps, nmax, ss, nn, bb, phi):
        st = BicgStats()
        self.L.orc_bicgstab(self.h, eps, nmax, _ip(ss), _ip(nn), _dp(bb), _dp(phi), None, C.byref(st))
        return st

    def bicgstab_mixed(self, eps, nmax, ss, nn, bb, phi, sweeps, omega=0.8):
        st = BicgStats()
        self.L.orc_bicgstab_mixed(self.h, eps, nmax, _ip(ss), _ip(nn), _dp(bb), _dp(phi), int(sweeps),
                                  float(omega), C.byref(st))
        return st

    def bicgstab_mg(self, mg, eps, nmax, ss, nn, bb, phi):
        st = BicgStats()
        self.L.orc_bicgstab_mg(self.h, mg.h, eps, nmax, _ip(ss), _ip(nn), _dp(bb), _dp(phi), C.byref(st))
        return st

    def bicgstab_mg_f32(self, mg, eps, nmax, ss, nn, bb, phi):
        st = BicgStats()
        self.L.orc_bicgstab_mg_f32(self.h, mg.h, eps, nmax, _ip(ss), _ip(nn), _dp(bb), _dp(phi), C.byref(st))
        return st

    def bicgstab_jor64(self, eps, nmax, ss, nn, bb, phi, sweeps, omega=0.8):
        st = BicgStats()
        self.L.orc_bicgstab_jor64(self.h, eps, nmax, _ip(ss), _ip(nn), _dp(bb), _dp(phi), int(sweeps), float(omega),
                                  C.byref(st))
        return st

    def jor_mixed(self, n, omega, ss, nn, rel, bb):
        self.L.orc_jor_mixed(self.h, n, omega, _ip(ss), _ip(nn), _dp(rel), _dp(bb))

    def timestep(self, params: StepParams, step, time, vel3, pre, nl3):
        st = StepStats()
        self.L.orc_timestep(self.h, C.byref(params), step, time, _dp(vel3), _dp(pre), _dp(nl3), C.byref(st))
        return st

    def tube_mask(self, yc, zc, r_u, r_p):
        m = np.zeros((3,) + self.shape, dtype=np.uint8)
        self.L.orc_tube_mask(self.h, yc, zc, _dp(np.ascontiguousarray(r_u)), _dp(np.ascontiguousarray(r_p)),
                             m.ctypes.data_as(C.POINTER(C.c_uint8)))
        return m

    def apply_mask(self, mask3, vel3):
        self.L.orc_apply_mask(self.h, mask3.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(vel3))

    def level_pressure(self, pre):
        self.L.orc_level_pressure(self.h, _dp(pre))

    def fill_corners(self, pre):
        self.L.orc_fill_corners(self.h, _dp(pre))


class OracleMG:
    """The multigrid V cycle of a case (preconditioner 1): levels 1..G."""

    def __init__(self, case: OracleCase, n_grids_limit=15, n_down=4, n_up=4, n_bottom=4):
        self.L = lib()
        self.case = case
        self.h = self.L.orc_mg_create(case.h, n_grids_limit, n_down, n_up, n_bottom)
        if not self.h:
            raise ValueError("oracle rejected the multigrid configuration")
        self.G = int(self.L.orc_mg_grids(self.h))
        nn = (C.c_int * (3 * self.G))()
        self.L.orc_mg_levels(self.h, nn)
        self.NN = [tuple(nn[3 * g:3 * g + 3]) for g in range(self.G)]  # level g+1

    def __del__(self):
        try:
            if self.h:
                self.L.orc_mg_destroy(self.h)
        except Exception:
            pass

    def shape(self, g):
        N = self.NN[g - 1]
        return (N[2] + 7, N[1] + 7, N[0] + 7)

    def zeros(self, g):
        return np.zeros(self.shape(g), dtype=np.float64)

    def coeff(self, g, name):
        buf = np.zeros(3 * max(self.NN[g - 1]) + 16)
        cnt = self.L.orc_mg_coeff(self.h, g, name.encode(), _dp(buf))
        if cnt < 0:
            raise KeyError(name)
        return buf[:cnt].copy()

    def vcycle(self, bb, phi):
        self.L.orc_mg_vcycle(self.h, _dp(bb), _dp(phi))

    def vcycle_f32(self, bb, phi):
        self.L.orc_mg_vcycle_f32(self.h, _dp(bb), _dp(phi))

    def restrict_residual(self, g, bb, x, coarse):
        self.L.orc_mg_restrict_residual(self.h, g, _dp(bb), _dp(x), _dp(coarse))

    def interpolate_add(self, g, coarse, fine):
        self.L.orc_mg_interpolate_add(self.h, g, _dp(coarse), _dp(fine))

    def relax(self, g, init, n, bb, rel):
        self.L.orc_mg_relax(self.h, g, int(init), int(n), _dp(bb), _dp(rel))


def vtk_read(path):
    """read_vtk_3D's reader: (dims, values) or None."""
    dims = (C.c_int * 3)()
    ptr = C.POINTER(C.c_double)()
    n = lib().orc_vtk_read(str(path).encode(), dims, C.byref(ptr))
    if n < 0:
        return None
    vals = np.ctypeslib.as_array(ptr, shape=(n,)).copy()
    C.CDLL(None).free(ptr)
    return tuple(dims), vals


def waveform_convert(t_ms, q, M1):
    t_ms = np.ascontiguousarray(t_ms, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    to, qo, ref = np.zeros_like(t_ms), np.zeros_like(q), np.zeros(3)
    lib().orc_waveform_convert(len(t_ms), _dp(t_ms), _dp(q), int(M1), _dp(to), _dp(qo), _dp(ref))
    return to, qo, ref


def ct_build(case, dims, values, offsets, ref_velocity, ref_length):
    lab = np.zeros(case.shape, dtype=np.uint8)
    info = (C.c_int * 6)()
    centre = np.zeros(4)
    rc = lib().orc_ct_build(case.h, _ip(dims), _dp(np.ascontiguousarray(values, dtype=np.float64)), *offsets,
                            float(ref_velocity), float(ref_length), lab.ctypes.data_as(C.POINTER(C.c_uint8)), info,
                            _dp(centre))
    if rc != 0:
        raise ValueError("oracle rejected the CT volume")
    return lab, list(info), centre


def ct_inflow(t_conv, q_conv, time):
    found = (C.c_int * 1)()
    v = lib().orc_ct_inflow(len(t_conv), _dp(t_conv), _dp(q_conv), float(time), found)
    return v, bool(found[0])


def ct_forcing(case, lab, t_conv, q_conv, time, vel3):
    lib().orc_ct_forcing(case.h, lab.ctypes.data_as(C.POINTER(C.c_uint8)), len(t_conv), _dp(t_conv), _dp(q_conv),
                         float(time), _dp(vel3))


def waveform_convert_branch(t_ms, q, M1, idealized):
    t_ms = np.ascontiguousarray(t_ms, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    to, qo, ref = np.zeros_like(t_ms), np.zeros_like(q), np.zeros(3)
    lib().orc_waveform_convert_branch(len(t_ms), _dp(t_ms), _dp(q), int(M1), int(idealized), _dp(to), _dp(qo),
                                      _dp(ref))
    return to, qo, ref


def wk_pressure_flow(branch, time, ref, state):
    """get_pressure_and_flow_rate_WKM: (p, q) nondimensional; state (3 doubles, p_end_systole_*) is updated"""
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    p, q = np.zeros(1), np.zeros(1)
    lib().orc_wk_pressure_flow(int(branch), float(time), _dp(ref), _dp(state), _dp(p), _dp(q))
    return p[0], q[0]


def wk_echo(t_ms, q_raw, t_conv, time, ref_velocity):
    pi, po = np.zeros(1), np.zeros(1)
    lib().orc_wk_echo(len(t_ms), _dp(np.ascontiguousarray(t_ms, dtype=np.float64)),
                      _dp(np.ascontiguousarray(q_raw, dtype=np.float64)),
                      _dp(np.ascontiguousarray(t_conv, dtype=np.float64)), float(time), float(ref_velocity),
                      _dp(pi), _dp(po))
    return pi[0], po[0]


def ct_apply_forcing(case, lab, st: CtState, time, vel3):
    lib().orc_ct_apply_forcing(case.h, lab.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(st), float(time), _dp(vel3))


def set_pressure_bc(case, lab, types, k_start, p_in, p_out, p_out2, rel):
    lib().orc_set_pressure_bc(case.h, lab.ctypes.data_as(C.POINTER(C.c_uint8)), int(types[0]), int(types[1]),
                              int(types[2]), int(k_start), float(p_in), float(p_out), float(p_out2), _dp(rel))


def device_jacobi_solver(case, omega, ss, nn, rel, bb, lab=None, k_start=0, substep=0, dtime=0.0, pres=(0, 0, 0),
                         max_iter=200000):
    """returns (status, sweeps)"""
    sweeps = (C.c_int * 1)()
    pres = np.ascontiguousarray(pres, dtype=np.float64)
    labp = lab.ctypes.data_as(C.POINTER(C.c_uint8)) if lab is not None else None
    rc = lib().orc_device_jacobi_solver(case.h, float(omega), _ip(ss), _ip(nn), _dp(rel), _dp(bb), labp,
                                        int(k_start), int(substep), float(dtime), _dp(pres), int(max_iter), sweeps)
    return rc, sweeps[0]


def set_threads(n: int) -> int:
    """threads of the OpenMP build's point loops (n <= 0 only queries)"""
    return int(lib().orc_set_threads(int(n)))


def exact_table(abl, n, dx):
    cc = np.zeros(7)
    ok = lib().orc_exact_table(abl, n, dx, _dp(cc))
    return cc[:n].copy() if ok else None
