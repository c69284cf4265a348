"""Python binding of libimpact_hip (the MI355X projection-step library).

The library is the product; this module only loads it with ctypes and mirrors
its C ABI (include/impact_hip.h) with numpy arrays in the reference's Fortran
layout phi(-3:N1+3, -3:N2+3, -3:N3+3), viewed from numpy as shape
(N3+7, N2+7, N1+7) so that element (i, j, k) is ``a[k+3, j+3, i+3]``.

There is no CPU fallback: if the built library is missing, importing the
binding's loader raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "lib" / "libimpact_hip.so"
IO_LIB_PATH = PKG_ROOT / "lib" / "libimpact_io.so"
IH_IO_ABI_VERSION = 1

IH_ABI_VERSION = 4

# field ids (include/impact_hip.h)
(VEL1, VEL2, VEL3, NL1, NL2, NL3, RHS1, RHS2, RHS3, WORK1, WORK2, WORK3, PRE, RES, GPRE, PP, COMP, DIG,
 RR, RH, PK, AP, AR, TMP1, TMP2, TMP3, TMP4, Z1, Z2, NFIELDS) = range(30)


IH_ENOTCONV = -1002


class IHError(RuntimeError):
    pass


class GridDesc(C.Structure):
    _fields_ = [("version", C.c_int), ("N", C.c_int * 3), ("bc", C.c_int * 6), ("L", C.c_double * 3),
                ("stretch", C.c_double * 3)]


class CoeffSet(C.Structure):
    _fields_ = [(name, C.POINTER(C.c_double) * 3) for name in
                ("cGp", "cD", "cIp", "cIu", "cNpD", "cNpU", "cNuD", "cNuU", "cp2", "cu2", "cdg")]


class BicgStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("norm_inf_initial", C.c_double), ("norm_inf_final", C.c_double),
                ("exit_reason", C.c_int)]


class CtInfo(C.Structure):
    _fields_ = [("inlet_cells", C.c_int), ("outlet_cells", C.c_int), ("outlet2_cells", C.c_int),
                ("inlet_jet_voxels", C.c_int), ("k_start", C.c_int), ("k_start2", C.c_int),
                ("inlet_centre", C.c_double * 2), ("outlet_centre", C.c_double * 2),
                ("reference_velocity", C.c_double), ("reference_length", C.c_double),
                ("reference_time", C.c_double)]


class StepParams(C.Structure):
    _fields_ = [("Re", C.c_double), ("dtime", C.c_double), ("jor_sweeps", C.c_int), ("jor_omega", C.c_double),
                ("bicg_rel_tol", C.c_double), ("bicg_max_it", C.c_int), ("level_pressure", C.c_int),
                ("inflow_kind", C.c_int), ("inflow_amp", C.c_double), ("use_mask", C.c_int), ("jor_fp32", C.c_int),
                ("precond_sweeps", C.c_int), ("precond_poisson", C.c_int),
                ("mg_fp32", C.c_int), ("use_ct", C.c_int), ("pressure_bc", C.c_int), ("precond_fp64", C.c_int)]


class IoState(C.Structure):
    """ih_io_state (include/impact_io.h): the run state of the HDF5 files."""
    _fields_ = ([("version", C.c_int)] +
                [(n, C.c_int) for n in ("restart", "write_count", "timestep", "write_out_vect", "write_out_scal",
                                         "new_dtime", "Euler_yes", "twostep_yes", "mapping_yes", "upwind_yes")] +
                [(n, C.c_double) for n in ("time", "dtime", "time_out_vect", "time_out_scal", "dtime_out_vect",
                                            "dtime_out_scal", "Re", "CFL", "thetaL", "epsU", "p_end_systole_in",
                                            "p_end_systole_out", "p_end_systole_out_2", "U_ref", "L_ref")])

    def __init__(self, **kw):
        super().__init__()
        self.version = IH_IO_ABI_VERSION
        for k, v in kw.items():
            setattr(self, k, v)


class Request(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("N", C.c_int * 3), ("steps", C.c_int), ("radius", C.c_double),
                ("stenosis", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("start_s", C.c_double), ("finish_s", C.c_double), ("latency_s", C.c_double),
                ("bicg_iterations", C.c_int), ("checksum_vel", C.c_double), ("checksum_pre", C.c_double),
                ("status", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("bicg_iterations", C.c_int * 3), ("bicg_final", C.c_double * 3), ("div_max_after", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_int64), ("max_active", C.c_int), ("time_scale", C.c_double),
                ("step", StepParams), ("progress_every", C.c_int)]


class ServeSummary(C.Structure):
    _fields_ = [("makespan_s", C.c_double), ("peak_bytes", C.c_int64), ("peak_active", C.c_int),
                ("setup_s", C.c_double), ("pool_bytes", C.c_int64), ("pool_slots", C.c_int),
                ("admit_wait_max_s", C.c_double), ("admit_wait_mean_s", C.c_double)]


_lib = None


def load(path: os.PathLike | None = None):
    """Load libimpact_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise IHError(f"{p} is missing: build it with `make -C {PKG_ROOT}` (or __graft_entry__.build())")
    L = C.CDLL(str(p))
    vp, dp, ip, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64)
    sig = {
        "ih_abi_version": (C.c_int, []),
        "ih_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
        "ih_device_count": (C.c_int, [ip]),
        "ih_case_create": (C.c_int, [C.POINTER(GridDesc), C.c_int, C.POINTER(vp)]),
        "ih_case_destroy": (None, [vp]),
        "ih_case_set_coeffs": (C.c_int, [vp, C.POINTER(CoeffSet)]),
        "ih_case_generate_coeffs": (C.c_int, [vp]),
        "ih_case_get_coeff": (C.c_int, [vp, C.c_char_p, dp, i64p]),
        "ih_case_get_coords": (C.c_int, [vp, C.c_int, C.c_int, dp]),
        "ih_grid_coeff": (C.c_int, [C.POINTER(GridDesc), C.c_char_p, dp, i64p]),
        "ih_grid_coords": (C.c_int, [C.POINTER(GridDesc), C.c_int, C.c_int, dp]),
        "ih_case_get_ranges": (C.c_int, [vp, ip]),
        "ih_field_elems": (C.c_int64, [vp]),
        "ih_case_bytes": (C.c_int64, [vp]),
        "ih_field_upload": (C.c_int, [vp, C.c_int, dp]),
        "ih_field_download": (C.c_int, [vp, C.c_int, dp]),
        "ih_field_fill": (C.c_int, [vp, C.c_int, C.c_double]),
        "ih_field_copy": (C.c_int, [vp, C.c_int, C.c_int]),
        "ih_field_get_box": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), dp]),
        "ih_field_set_box": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), dp]),
        "ih_case_get_grid": (C.c_int, [vp, C.POINTER(GridDesc)]),
        "ih_synchronize": (C.c_int, [vp]),
        "ih_exchange": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "ih_exchange_relax": (C.c_int, [vp, C.c_int]),
        "ih_divergence2": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ih_gradient": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "ih_interpolate2_vel_pre": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "ih_interpolate2_pre_vel": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "ih_interpolate_vel": (C.c_int, [vp]),
        "ih_nonlinear_upwind": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ih_helmholtz": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.c_int]),
        "ih_axpby": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]),
        "ih_product_div_grad": (C.c_int, [vp, C.c_int, C.c_int]),
        "ih_dot": (C.c_int, [vp, ip, ip, C.c_int, C.c_int, dp]),
        "ih_dot_noalloc": (C.c_int, [vp, ip, ip, C.c_int, C.c_int, dp]),
        "ih_jor": (C.c_int, [vp, C.c_int, C.c_double, ip, ip, C.c_int, C.c_int]),
        "ih_residual_cdg": (C.c_int, [vp, ip, ip, C.c_int, C.c_int, dp]),
        "ih_bicgstab": (C.c_int, [vp, C.c_double, C.c_int, ip, ip, C.c_int, C.c_int, C.POINTER(BicgStats)]),
        "ih_bicgstab_mixed": (C.c_int, [vp, C.c_double, C.c_int, ip, ip, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.POINTER(BicgStats)]),
        "ih_bicgstab_jor64": (C.c_int, [vp, C.c_double, C.c_int, ip, ip, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.POINTER(BicgStats)]),
        "ih_bicgstab_precond": (C.c_int, [vp, C.c_double, C.c_int, ip, ip, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(BicgStats)]),
        "ih_mg_configure": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ih_mg_levels": (C.c_int, [vp, ip, ip]),
        "ih_mg_get_coeff": (C.c_int, [vp, C.c_int, C.c_char_p, dp, C.POINTER(C.c_int64)]),
        "ih_multigrid_v": (C.c_int, [vp, C.c_int, C.c_int]),
        "ih_multigrid_v_f32": (C.c_int, [vp, C.c_int, C.c_int]),
        "ih_bicgstab_mg_f32": (C.c_int, [vp, C.c_double, C.c_int, ip, ip, C.c_int, C.c_int, C.POINTER(BicgStats)]),
        "ih_ct_load_vtk": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, C.c_int,
                                     C.POINTER(CtInfo)]),
        "ih_ct_get_labels": (C.c_int, [vp, C.POINTER(C.c_uint8)]),
        "ih_ct_inflow": (C.c_int, [vp, C.c_double, dp, ip]),
        "ih_ct_forcing": (C.c_int, [vp, C.c_double]),
        "ih_ct_configure": (C.c_int, [vp, C.c_int, C.c_int]),
        "ih_ct_get_state": (C.c_int, [vp, dp, dp]),
        "ih_ct_set_state": (C.c_int, [vp, dp, dp]),
        "ih_set_pressure_bc": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double]),
        "ih_case_set_substep": (C.c_int, [vp, C.c_int, C.c_double]),
        "ih_fill_corners": (C.c_int, [vp, C.c_int]),
        "ih_level_pressure": (C.c_int, [vp, C.c_int]),
        "ih_set_tube_mask": (C.c_int, [vp, C.c_double, C.c_double, dp, dp]),
        "ih_get_mask": (C.c_int, [vp, C.POINTER(C.c_uint8)]),
        "ih_apply_mask": (C.c_int, [vp]),
        "ih_timestep": (C.c_int, [vp, C.POINTER(StepParams), C.c_double, C.POINTER(StepStats)]),
        "ih_device_jacobi_solver": (C.c_int, [vp, C.c_int, C.c_double] + [C.c_int] * 6 + [dp, dp]),
        "ih_device_gradient": (C.c_int, [vp, C.c_int, dp, dp]),
        "ih_device_divergence2": (C.c_int, [vp, dp, dp]),
        "ih_device_nonlinear_upwind": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "ih_device_helmholtz": (C.c_int, [vp, C.c_int, C.c_double, dp, dp]),
        "ih_device_axpby": (C.c_int, [vp, C.c_double, C.c_double, dp, dp, dp]),
        "ih_device_interpolate": (C.c_int, [vp, C.c_int, dp, dp]),
        "ih_device_interpolate2": (C.c_int, [vp, C.c_int, dp, dp]),
        "ih_batch_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
        "ih_batch_destroy": (None, [vp]),
        "ih_batch_autotune": (C.c_int, [vp, C.c_int, ip, ip, ip, dp]),
        "ih_batch_set_shape": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "ih_batch_set_groups": (C.c_int, [vp, C.c_int64]),
        "ih_batch_get_groups": (C.c_int, [vp, C.POINTER(C.c_int64), ip]),
        "ih_batch_jor": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.c_int]),
        "ih_batch_timestep": (C.c_int, [vp, C.POINTER(StepParams), C.c_double, C.POINTER(StepStats)]),
        "ih_grid_bytes": (C.c_int64, [C.POINTER(GridDesc)]),
        "ih_grid_bytes_step": (C.c_int64, [C.POINTER(GridDesc), C.POINTER(StepParams)]),
        "ih_serve": (C.c_int, [C.POINTER(ServeConfig), C.POINTER(Request), C.c_int, C.POINTER(RequestResult),
                               C.POINTER(ServeSummary)]),
        "ih_timer_start": (C.c_int, [vp]),
        "ih_timer_stop": (C.c_int, [vp, dp]),
        "ih_kernel_timing_enable": (C.c_int, [vp, C.c_int]),
        "ih_kernel_timing_get": (C.c_int, [vp, C.c_char_p, dp, i64p]),
        "ih_bench_kernel": (C.c_int, [vp, C.c_char_p, C.c_int, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.ih_abi_version() != IH_ABI_VERSION:
        raise IHError("ABI version mismatch")
    _lib = L
    return L


_io = None


def load_io(path: os.PathLike | None = None):
    """libimpact_io.so (HDF5 output and restart); raises if it is not built."""
    global _io
    if _io is not None:
        return _io
    load()
    p = Path(path) if path else IO_LIB_PATH
    if not p.exists():
        raise IHError(f"{p} is not built (make -C gpu-impact-ct_amd)")
    L = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    vp, ip, cs = C.c_void_p, C.POINTER(C.c_int), C.c_char_p
    sp = C.POINTER(IoState)
    sig = {
        "ih_io_abi_version": (C.c_int, []),
        "ih_io_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
        "ih_io_write_hdf": (C.c_int, [vp, cs, cs, C.c_int, C.c_int, ip, sp]),
        "ih_io_read_hdf": (C.c_int, [vp, cs, cs, C.c_int, C.c_int, sp]),
        "ih_io_write_restart": (C.c_int, [vp, cs, sp]),
        "ih_io_read_restart": (C.c_int, [vp, cs, C.c_int, sp]),
        "ih_io_write_fields": (C.c_int, [vp, cs, ip, ip, sp]),
        "ih_io_write_xdmf": (C.c_int, [vp, cs, ip, ip, sp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.ih_io_abi_version() != IH_IO_ABI_VERSION:
        raise IHError("I/O ABI version mismatch")
    _io = L
    return L


def io_check(rc: int, what: str = ""):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        load_io().ih_io_last_error(buf, 1024)
        raise IHError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")


def exported_symbols():
    """Names the binding expects the library to export."""
    load()
    return [n for n in dir(_lib) if n.startswith("ih_")]


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    load().ih_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str = ""):
    if rc != 0:
        raise IHError(f"{what} failed ({rc}): {last_error()}")


def _dp(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a contiguous float64 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ints(seq):
    return (C.c_int * len(seq))(*[int(x) for x in seq])


def grid_desc(N, bc, L=(1.0, 1.0, 1.0), stretch=(0.0, 0.0, 0.0)) -> GridDesc:
    g = GridDesc()
    g.version = IH_ABI_VERSION
    for d in range(3):
        g.N[d] = int(N[d])
        g.L[d] = float(L[d])
        g.stretch[d] = float(stretch[d])
    for s in range(6):
        g.bc[s] = int(bc[s])
    return g


def grid_coeff(N, bc, L, stretch, name: str) -> np.ndarray:
    """Coefficient array the library generates for a grid (host only, no device)."""
    L_ = load()
    g = grid_desc(N, bc, L, stretch)
    n = C.c_int64(0)
    check(L_.ih_grid_coeff(C.byref(g), name.encode(), None, C.byref(n)), "ih_grid_coeff")
    out = np.zeros(n.value)
    check(L_.ih_grid_coeff(C.byref(g), name.encode(), _dp(out), C.byref(n)), "ih_grid_coeff")
    return out


def grid_coords(N, bc, L, stretch, d: int, which: int) -> np.ndarray:
    L_ = load()
    g = grid_desc(N, bc, L, stretch)
    out = np.zeros(int(N[d]) + 7)
    check(L_.ih_grid_coords(C.byref(g), d, which, _dp(out)), "ih_grid_coords")
    return out


def device_count() -> int:
    n = C.c_int(0)
    check(load().ih_device_count(C.byref(n)), "ih_device_count")
    return n.value


class Case:
    """One grid block resident on a GPU (ih_case)."""

    def __init__(self, N, bc, L=(1.0, 1.0, 1.0), stretch=(0.0, 0.0, 0.0), device: int = 0, generate: bool = True):
        self.lib = load()
        self.N = tuple(int(n) for n in N)
        self.bc = tuple(int(b) for b in bc)
        self.L = tuple(float(x) for x in L)
        self.stretch = tuple(float(x) for x in stretch)
        g = grid_desc(self.N, self.bc, self.L, self.stretch)
        h = C.c_void_p()
        check(self.lib.ih_case_create(C.byref(g), device, C.byref(h)), "ih_case_create")
        self.h = h
        self.shape = (self.N[2] + 7, self.N[1] + 7, self.N[0] + 7)
        r = (C.c_int * 48)()
        check(self.lib.ih_case_get_ranges(self.h, r), "ranges")
        r = list(r)
        self.Sp, self.Np = r[0:3], r[3:6]
        self.SB = [r[6 + 3 * m:9 + 3 * m] for m in range(3)]
        self.NB = [r[15 + 3 * m:18 + 3 * m] for m in range(3)]
        self.SI = [r[24 + 3 * m:27 + 3 * m] for m in range(3)]
        self.NI = [r[33 + 3 * m:36 + 3 * m] for m in range(3)]
        if generate:
            check(self.lib.ih_case_generate_coeffs(self.h), "ih_case_generate_coeffs")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ih_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- fields -------------------------------------------------------------
    def upload(self, f: int, a: np.ndarray):
        assert a.shape == self.shape
        check(self.lib.ih_field_upload(self.h, f, _dp(np.ascontiguousarray(a))), "upload")

    def download(self, f: int) -> np.ndarray:
        a = np.zeros(self.shape)
        check(self.lib.ih_field_download(self.h, f, _dp(a)), "download")
        return a

    def fill(self, f: int, v: float):
        check(self.lib.ih_field_fill(self.h, f, float(v)), "fill")

    def sync(self):
        check(self.lib.ih_synchronize(self.h), "sync")

    def bytes(self) -> int:
        return int(self.lib.ih_case_bytes(self.h))

    def coeff(self, name: str) -> np.ndarray:
        n = C.c_int64(0)
        check(self.lib.ih_case_get_coeff(self.h, name.encode(), None, C.byref(n)), "coeff")
        out = np.zeros(n.value)
        check(self.lib.ih_case_get_coeff(self.h, name.encode(), _dp(out), C.byref(n)), "coeff")
        return out

    @staticmethod
    def _box_shape(lo, hi, stride):
        """(n3, n2, n1) of a valid box; the library rejects an invalid one
        before it touches the buffer, which then only needs to exist."""
        n = [max((int(hi[d]) - int(lo[d])) // max(int(stride[d]), 1) + 1, 1) for d in range(3)]
        return (n[2], n[1], n[0])

    def get_box(self, f: int, lo, hi, stride=(1, 1, 1)) -> np.ndarray:
        """The strided box of a field, shape (n3, n2, n1) (x fastest)."""
        out = np.zeros(self._box_shape(lo, hi, stride))
        check(self.lib.ih_field_get_box(self.h, f, _ints(lo), _ints(hi), _ints(stride), _dp(out)), "get_box")
        return out

    def set_box(self, f: int, lo, hi, stride, data: np.ndarray):
        data = np.ascontiguousarray(data, dtype=np.float64)
        if data.size != int(np.prod(self._box_shape(lo, hi, stride))):
            raise ValueError("data does not have the box's size")
        check(self.lib.ih_field_set_box(self.h, f, _ints(lo), _ints(hi), _ints(stride), _dp(data)), "set_box")

    def grid(self) -> GridDesc:
        g = GridDesc()
        check(self.lib.ih_case_get_grid(self.h, C.byref(g)), "grid")
        return g

    # --- HDF5 output and restart (libimpact_io) ---------------------------
    def write_hdf(self, filename: str, dsetname: str, f: int, vel_dir: int, stride, st: IoState):
        io_check(load_io().ih_io_write_hdf(self.h, str(filename).encode(), dsetname.encode(), f, vel_dir,
                                           _ints(stride), C.byref(st)), "write_hdf")

    def read_hdf(self, filename: str, dsetname: str, f: int, vel_dir: int, st: IoState | None = None):
        io_check(load_io().ih_io_read_hdf(self.h, str(filename).encode(), dsetname.encode(), f, vel_dir,
                                          C.byref(st) if st is not None else None), "read_hdf")

    def write_restart(self, prefix: str, st: IoState):
        io_check(load_io().ih_io_write_restart(self.h, str(prefix).encode(), C.byref(st)), "write_restart")

    def read_restart(self, prefix: str, restart: int, st: IoState):
        io_check(load_io().ih_io_read_restart(self.h, str(prefix).encode(), restart, C.byref(st)), "read_restart")

    def write_fields(self, prefix: str, sizes, strides, st: IoState):
        io_check(load_io().ih_io_write_fields(self.h, str(prefix).encode(), _ints(sizes),
                                              _ints([s for t in strides for s in t]), C.byref(st)), "write_fields")

    def write_xdmf(self, prefix: str, sizes, strides, st: IoState):
        io_check(load_io().ih_io_write_xdmf(self.h, str(prefix).encode(), _ints(sizes),
                                            _ints([s for t in strides for s in t]), C.byref(st)), "write_xdmf")

    def coords(self, d: int, which: int) -> np.ndarray:
        out = np.zeros(self.N[d] + 7)
        check(self.lib.ih_case_get_coords(self.h, d, which, _dp(out)), "coords")
        return out

    # --- operators (field ids) ---------------------------------------------
    def exchange(self, d, vel_dir, f):
        check(self.lib.ih_exchange(self.h, d, vel_dir, f), "exchange")

    def divergence2(self, u=VEL1, v=VEL2, w=VEL3, div=RES):
        check(self.lib.ih_divergence2(self.h, u, v, w, div), "divergence2")

    def gradient(self, m, phi, grad):
        check(self.lib.ih_gradient(self.h, m, phi, grad), "gradient")

    def interpolate_vel(self):
        check(self.lib.ih_interpolate_vel(self.h), "interpolate_vel")

    def interpolate2_pre_vel(self, m, phi, inter):
        check(self.lib.ih_interpolate2_pre_vel(self.h, m, phi, inter), "interpolate2_pre_vel")

    def interpolate2_vel_pre(self, m, phi, inter):
        check(self.lib.ih_interpolate2_vel_pre(self.h, m, phi, inter), "interpolate2_vel_pre")

    def nonlinear_upwind(self, cid, pp, vel_c, nl_c):
        check(self.lib.ih_nonlinear_upwind(self.h, cid, pp, vel_c, nl_c), "nonlinear_upwind")

    def helmholtz(self, comp, multL, vel_c, nl_c):
        check(self.lib.ih_helmholtz(self.h, comp, float(multL), vel_c, nl_c), "helmholtz")

    def product_div_grad(self, phi, lap):
        check(self.lib.ih_product_div_grad(self.h, phi, lap), "product_div_grad")

    def dot(self, ss, nn, a, b) -> float:
        out = C.c_double(0.0)
        check(self.lib.ih_dot(self.h, _ints(ss), _ints(nn), a, b, C.byref(out)), "dot")
        return out.value

    def dot_noalloc(self, ss, nn, a, b) -> float:
        """ih_dot_noalloc: the same product through the library's fixed scratch array"""
        out = C.c_double(0.0)
        check(self.lib.ih_dot_noalloc(self.h, _ints(ss), _ints(nn), a, b, C.byref(out)), "dot_noalloc")
        return out.value

    def jor(self, n, omega, ss, nn, rel=PRE, bb=RES):
        check(self.lib.ih_jor(self.h, int(n), float(omega), _ints(ss), _ints(nn), rel, bb), "jor")

    def residual_cdg(self, ss, nn, rel=PRE, bb=RES) -> float:
        out = C.c_double(0.0)
        check(self.lib.ih_residual_cdg(self.h, _ints(ss), _ints(nn), rel, bb, C.byref(out)), "residual_cdg")
        return out.value

    def bicgstab(self, eps, nmax, ss, nn, bb=RES, phi=PRE) -> BicgStats:
        st = BicgStats()
        check(self.lib.ih_bicgstab(self.h, float(eps), int(nmax), _ints(ss), _ints(nn), bb, phi, C.byref(st)),
              "bicgstab")
        return st

    def bicgstab_mixed(self, eps, nmax, ss, nn, sweeps, omega=0.8, bb=RES, phi=PRE) -> BicgStats:
        st = BicgStats()
        check(self.lib.ih_bicgstab_mixed(self.h, float(eps), int(nmax), _ints(ss), _ints(nn), bb, phi, int(sweeps),
                                         float(omega), C.byref(st)), "bicgstab_mixed")
        return st

    def bicgstab_jor64(self, eps, nmax, ss, nn, sweeps, omega=0.8, bb=RES, phi=PRE) -> BicgStats:
        """the FP64-swept form of bicgstab_mixed (the like-for-like FP64 baseline)"""
        st = BicgStats()
        check(self.lib.ih_bicgstab_jor64(self.h, float(eps), int(nmax), _ints(ss), _ints(nn), bb, phi, int(sweeps),
                                         float(omega), C.byref(st)), "bicgstab_jor64")
        return st

    def bicgstab_precond(self, eps, nmax, ss, nn, preconditioner, bb=RES, phi=PRE) -> BicgStats:
        """BiCGstab_orig with the reference's preconditioner argument (0 none, 1 multigrid V cycle)."""
        st = BicgStats()
        check(self.lib.ih_bicgstab_precond(self.h, float(eps), int(nmax), _ints(ss), _ints(nn), bb, phi,
                                           int(preconditioner), C.byref(st)), "bicgstab_precond")
        return st

    def mg_configure(self, n_grids_limit=15, n_relax_down=4, n_relax_up=4, n_relax_bottom=4):
        check(self.lib.ih_mg_configure(self.h, int(n_grids_limit), int(n_relax_down), int(n_relax_up),
                                       int(n_relax_bottom)), "mg_configure")

    def mg_levels(self):
        G = C.c_int(0)
        nn = (C.c_int * 45)()
        check(self.lib.ih_mg_levels(self.h, C.byref(G), nn), "mg_levels")
        return [tuple(nn[3 * g:3 * g + 3]) for g in range(G.value)]

    def mg_coeff(self, level, name) -> np.ndarray:
        n = C.c_int64(0)
        check(self.lib.ih_mg_get_coeff(self.h, int(level), name.encode(), None, C.byref(n)), "mg_get_coeff")
        out = np.zeros(n.value)
        check(self.lib.ih_mg_get_coeff(self.h, int(level), name.encode(), _dp(out), C.byref(n)), "mg_get_coeff")
        return out

    def multigrid_v(self, bb, phi):
        """phi = M bb (one V cycle); bb's edges of wall pairs are averaged in place."""
        check(self.lib.ih_multigrid_v(self.h, bb, phi), "multigrid_v")

    def multigrid_v_f32(self, bb, phi):
        """The FP32 cycle: phi = (double) M32 (float) bb; bb's edges averaged in FP64."""
        check(self.lib.ih_multigrid_v_f32(self.h, bb, phi), "multigrid_v_f32")

    def bicgstab_mg_f32(self, eps, nmax, ss, nn, bb=RES, phi=PRE) -> BicgStats:
        st = BicgStats()
        check(self.lib.ih_bicgstab_mg_f32(self.h, float(eps), int(nmax), _ints(ss), _ints(nn), bb, phi,
                                          C.byref(st)), "bicgstab_mg_f32")
        return st

    def ct_load_vtk(self, path, offsets, wave_t_ms, wave_q) -> CtInfo:
        """read_vtk_3D: the CT label image at offsets (x, y, z) and the echo waveform."""
        t = np.ascontiguousarray(wave_t_ms, dtype=np.float64)
        q = np.ascontiguousarray(wave_q, dtype=np.float64)
        info = CtInfo()
        check(self.lib.ih_ct_load_vtk(self.h, str(path).encode(), int(offsets[0]), int(offsets[1]), int(offsets[2]),
                                      _dp(t), _dp(q), len(t), C.byref(info)), "ct_load_vtk")
        return info

    def ct_labels(self) -> np.ndarray:
        out = np.zeros(self.shape, dtype=np.uint8)
        check(self.lib.ih_ct_get_labels(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8))), "ct_get_labels")
        return out

    def ct_inflow(self, time):
        q, found = C.c_double(0.0), C.c_int(0)
        check(self.lib.ih_ct_inflow(self.h, float(time), C.byref(q), C.byref(found)), "ct_inflow")
        return q.value, bool(found.value)

    def ct_forcing(self, time):
        check(self.lib.ih_ct_forcing(self.h, float(time)), "ct_forcing")

    def ct_configure(self, ct_based: int, idealized: int):
        """apply_CT_forcing's branches (CT_based_aorta, idealized_aorta); before ct_load_vtk"""
        check(self.lib.ih_ct_configure(self.h, int(ct_based), int(idealized)), "ct_configure")

    def ct_state(self):
        """(pre_inlet, pre_outlet, pre_outlet_2), (p_end_systole_in, _out, _out_2)"""
        pres, wk = np.zeros(3), np.zeros(3)
        check(self.lib.ih_ct_get_state(self.h, _dp(pres), _dp(wk)), "ct_get_state")
        return pres, wk

    def ct_set_state(self, pres, wk):
        check(self.lib.ih_ct_set_state(self.h, _dp(np.ascontiguousarray(pres, dtype=np.float64)),
                                       _dp(np.ascontiguousarray(wk, dtype=np.float64))), "ct_set_state")

    def set_pressure_bc(self, rel, types, p_in, p_out, p_out2):
        check(self.lib.ih_set_pressure_bc(self.h, rel, int(types[0]), int(types[1]), int(types[2]), float(p_in),
                                          float(p_out), float(p_out2)), "set_pressure_bc")

    def set_substep(self, substep: int, dtime: float):
        check(self.lib.ih_case_set_substep(self.h, int(substep), float(dtime)), "set_substep")

    def fill_corners(self, f=PRE):
        check(self.lib.ih_fill_corners(self.h, f), "fill_corners")

    def level_pressure(self, f=PRE):
        check(self.lib.ih_level_pressure(self.h, f), "level_pressure")

    def set_tube_mask(self, yc, zc, r_u, r_p):
        check(self.lib.ih_set_tube_mask(self.h, float(yc), float(zc), _dp(np.ascontiguousarray(r_u, dtype=np.float64)),
                                        _dp(np.ascontiguousarray(r_p, dtype=np.float64))), "set_tube_mask")

    def get_mask(self) -> np.ndarray:
        m = np.zeros((3,) + self.shape, dtype=np.uint8)
        check(self.lib.ih_get_mask(self.h, m.ctypes.data_as(C.POINTER(C.c_uint8))), "get_mask")
        return m

    def apply_mask(self):
        check(self.lib.ih_apply_mask(self.h), "apply_mask")

    def timestep(self, params: StepParams, time: float, stats: bool = True) -> StepStats | None:
        st = StepStats()
        check(self.lib.ih_timestep(self.h, C.byref(params), float(time), C.byref(st) if stats else None), "timestep")
        return st if stats else None

    # --- reference-signature calls (host arrays) ---------------------------
    def device_gradient(self, m, phi, grad):
        check(self.lib.ih_device_gradient(self.h, m, _dp(phi), _dp(grad)), "device_gradient")

    def device_divergence2(self, phi3, div):
        check(self.lib.ih_device_divergence2(self.h, _dp(phi3), _dp(div)), "device_divergence2")

    def device_nonlinear_upwind(self, cid, pp, vel_c, nl_c):
        check(self.lib.ih_device_nonlinear_upwind(self.h, cid, _dp(pp), _dp(vel_c), _dp(nl_c)), "device_nonlinear")

    def device_helmholtz(self, cid, multL, vel_c, nl_c):
        check(self.lib.ih_device_helmholtz(self.h, cid, float(multL), _dp(vel_c), _dp(nl_c)), "device_helmholtz")

    def device_axpby(self, a, b, x, y, z):
        check(self.lib.ih_device_axpby(self.h, float(a), float(b), _dp(x), _dp(y), _dp(z)), "device_axpby")

    def device_interpolate(self, m, phi, inter):
        check(self.lib.ih_device_interpolate(self.h, m, _dp(phi), _dp(inter)), "device_interpolate")

    def device_interpolate2(self, m, phi, inter):
        check(self.lib.ih_device_interpolate2(self.h, m, _dp(phi), _dp(inter)), "device_interpolate2")

    def device_jacobi_solver(self, omega, ss, nn, rel, bb) -> int:
        """returns the status: 0, or IH_ENOTCONV when the sweep limit stopped it (rel is still written)"""
        rc = self.lib.ih_device_jacobi_solver(self.h, 1, float(omega), *[int(x) for x in ss], *[int(x) for x in nn],
                                              _dp(rel), _dp(bb))
        if rc not in (0, IH_ENOTCONV):
            check(rc, "device_jacobi_solver")
        return rc

    # --- timing ------------------------------------------------------------
    def bench_kernel(self, name: str, launches: int) -> float:
        """ms per launch of one kernel, back to back on the case's stream"""
        ms = C.c_double(0.0)
        check(self.lib.ih_bench_kernel(self.h, name.encode(), int(launches), C.byref(ms)), "bench_kernel")
        return ms.value

    def timer_start(self):
        check(self.lib.ih_timer_start(self.h), "timer_start")

    def timer_stop(self) -> float:
        ms = C.c_double(0.0)
        check(self.lib.ih_timer_stop(self.h, C.byref(ms)), "timer_stop")
        return ms.value


class Batch:
    """Many cases relaxed together (ih_batch_*); the cases must outlive the batch."""

    def __init__(self, cases):
        self.lib = load()
        self.cases = list(cases)
        arr = (C.c_void_p * len(self.cases))(*[c.h for c in self.cases])
        h = C.c_void_p()
        check(self.lib.ih_batch_create(arr, len(self.cases), C.byref(h)), "batch_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.ih_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, sweeps=10):
        """returns (threads per tile, z chunk, layout, ms per sweep) of the fastest launch shape"""
        tpb, zc, lay, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(self.lib.ih_batch_autotune(self.h, int(sweeps), C.byref(tpb), C.byref(zc), C.byref(lay), C.byref(ms)),
              "autotune")
        return tpb.value, zc.value, lay.value, ms.value

    def set_shape(self, tpb, zchunk, layout):
        check(self.lib.ih_batch_set_shape(self.h, int(tpb), int(zchunk), int(layout)), "batch_set_shape")

    def set_groups(self, nbytes):
        """relax the cases in groups whose fields fit nbytes (0: one group)"""
        check(self.lib.ih_batch_set_groups(self.h, int(nbytes)), "batch_set_groups")

    def groups(self):
        """(group budget in bytes, number of groups)"""
        nb, cnt = C.c_int64(), C.c_int()
        check(self.lib.ih_batch_get_groups(self.h, C.byref(nb), C.byref(cnt)), "batch_get_groups")
        return nb.value, cnt.value

    def jor(self, n, omega=0.8, fp32=False, graph=False):
        check(self.lib.ih_batch_jor(self.h, int(n), float(omega), 1 if fp32 else 0, 1 if graph else 0), "batch_jor")

    def timestep(self, params: StepParams, time: float, stats: bool = True):
        """one projection time step of every case (ih_batch_timestep); a list of
        StepStats, one per case, or None"""
        st = (StepStats * len(self.cases))() if stats else None
        check(self.lib.ih_batch_timestep(self.h, C.byref(params), float(time), st), "batch_timestep")
        return list(st) if stats else None


def serve(requests, step: StepParams, hbm_budget: int, max_active: int, time_scale: float = 1.0, device: int = 0,
          progress_every: int = 0):
    """Run a request trace (list of dicts with arrival_s, N, steps, radius,
    stenosis) through ih_serve; returns (results, summary, status).
    progress_every > 0 prints a line on stderr after every that many finished
    requests."""
    L = load()
    n = len(requests)
    req = (Request * max(n, 1))()
    for q, r in enumerate(requests):
        req[q].arrival_s = float(r["arrival_s"])
        req[q].N = (C.c_int * 3)(*r["N"])
        req[q].steps = int(r["steps"])
        req[q].radius = float(r["radius"])
        req[q].stenosis = float(r["stenosis"])
    out = (RequestResult * max(n, 1))()
    cfg = ServeConfig(device=device, hbm_budget=int(hbm_budget), max_active=int(max_active),
                      time_scale=float(time_scale), step=step, progress_every=int(progress_every))
    summ = ServeSummary()
    rc = L.ih_serve(C.byref(cfg), req, n, out, C.byref(summ))
    return list(out[:n]), summ, rc

