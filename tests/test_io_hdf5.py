"""HDF5 field output, restart and XDMF (libimpact_io, src/mod_inout.f90).

The files are read back with tests/h5read.py, a reader written against the
HDF5 C library that shares no code with the writer.  What is checked:
- layout: one 3-D dataset, x fastest, extent S?w..M?w of filespace_props;
- values: the field's box with the stride, bit for bit; the velocity of
  write_fields equals the oracle's exchange + interpolate2_vel_pre;
- the run state: every attribute of write_hdf (src/mod_inout.f90:596-656)
  with its name (trailing blanks included), type class, size and length; the
  coordinates y?p/y?u and VectorX/Y/Z equal the oracle's coordinates;
- restart: write_restart then read_restart gives back every field over its
  range and the run state, with read_restart's output-time update;
- the two input checks of write_hdf, and the XDMF text of write_3D_xdmf."""
import ctypes as C
import re
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import pytest

import h5read
import impact_hip as ih

ROOT = Path(__file__).resolve().parent.parent

# write_hdf's attributes (src/mod_inout.f90:596-656): name -> (kind, length);
# "M" is M?, "M+1" is M?+1 of the attribute's direction
ATTRS = {
    "restart": ("i", 1), "write_count": ("i", 1), "time_out_vect": ("r", 1), "time_out_scal": ("r", 1),
    "dtime_out_vect": ("r", 1), "dtime_out_scal": ("r", 1), "write_out_vect": ("i", 1),
    "write_out_scal": ("i", 1), "new_dtime": ("i", 1), "M1  M2  M3 ": ("i", 3), "S1w S2w S3w": ("i", 3),
    "M1w M2w M3w": ("i", 3), "NB1 NB2 NB3": ("i", 3), "ls1 ls2 ls3": ("i", 3), "time": ("r", 1),
    "dtime": ("r", 1), "timestep": ("i", 1), "Re": ("r", 1), "CFL   ": ("r", 1), "thetaL": ("r", 1),
    "epsU  ": ("r", 1), "p_end_systole_in": ("r", 1), "p_end_systole_out": ("r", 1),
    "p_end_systole_out_2": ("r", 1), "Euler_yes": ("i", 1), "twostep_yes": ("i", 1), "mapping_yes": ("i", 1),
    "upwind_yes": ("i", 1), "BC_iL": ("i", 3), "BC_iU": ("i", 3),
    **{f"ncb{d}{t}": ("i", 4) for d in (1, 2, 3) for t in "cgd"},
    "y1p": ("r", "M1"), "y2p": ("r", "M2"), "y3p": ("r", "M3"),
    "y1u": ("r", "M1+1"), "y2v": ("r", "M2+1"), "y3w": ("r", "M3+1"),
    "U_ref": ("r", 1), "L_ref": ("r", 1),
}

WALLS = ((33, 17, 17), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5))
PERIODIC_XZ = ((33, 17, 33), (-1, -1, 1, 1, -1, -1), (1.0, 1.0, 1.0), (0.0, 1.2, 0.0))


def state(**kw):
    base = dict(restart=3, write_count=12, timestep=345, write_out_vect=1, write_out_scal=0, new_dtime=1,
                Euler_yes=0, twostep_yes=1, mapping_yes=1, upwind_yes=1, time=0.75, dtime=5e-4,
                time_out_vect=0.8, time_out_scal=0.9, dtime_out_vect=0.05, dtime_out_scal=0.1, Re=500.0, CFL=0.6,
                thetaL=0.25, epsU=1e-6, p_end_systole_in=1.5, p_end_systole_out=1.25, p_end_systole_out_2=1.125,
                U_ref=0.3, L_ref=0.02)
    base.update(kw)
    return ih.IoState(**base)


# ---- CPU: the library, the reader and the expected layout ------------------

def test_io_library_exports_its_header():
    L = ih.load_io()
    decl = re.findall(r"\bint\s+(ih_io_\w+)\s*\(", (ROOT / "include" / "impact_io.h").read_text())
    assert len(decl) == 8
    for name in decl:
        assert hasattr(L, name), name


def test_reader_on_a_file_written_with_the_c_library(tmp_path):
    """h5read reads back what the HDF5 C library writes directly."""
    L = h5read.lib()
    hid, hs = h5read.hid_t, h5read.hsize_t
    L.H5Fcreate.restype = hid
    L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid, hid]
    L.H5Screate_simple.restype = hid
    L.H5Screate_simple.argtypes = [C.c_int, C.POINTER(hs), C.POINTER(hs)]
    L.H5Dcreate2.restype = hid
    L.H5Dcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid, hid]
    L.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    L.H5Acreate2.restype = hid
    L.H5Acreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid]
    L.H5Awrite.argtypes = [hid, hid, C.c_void_p]
    dbl, i32 = h5read._native("H5T_NATIVE_DOUBLE_g"), h5read._native("H5T_NATIVE_INT_g")
    f = L.H5Fcreate(str(tmp_path / "t.h5").encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    data = np.arange(24, dtype=np.float64).reshape(2, 3, 4) * 0.5
    dims = (hs * 3)(2, 3, 4)
    sp = L.H5Screate_simple(3, dims, None)
    ds = L.H5Dcreate2(f, b"field", dbl, sp, 0, 0, 0)
    L.H5Dwrite(ds, dbl, 0, 0, 0, data.ctypes.data)
    one = (hs * 1)(3)
    asp = L.H5Screate_simple(1, one, None)
    a = L.H5Acreate2(ds, b"M1  M2  M3 ", i32, asp, 0, 0)
    v = np.array([7, 8, 9], dtype=np.int32)
    L.H5Awrite(a, i32, v.ctypes.data)
    for h, close in ((a, L.H5Aclose), (asp, L.H5Sclose), (ds, L.H5Dclose), (sp, L.H5Sclose), (f, L.H5Fclose)):
        close.argtypes = [hid]
        close(h)
    with h5read.File(tmp_path / "t.h5") as r:
        assert r.names() == ["field"]
        got, cls, size = r.dataset("field")
        assert cls == h5read.H5T_FLOAT and size == 8 and np.array_equal(got, data)
        at = r.attributes("field")
        assert list(at) == ["M1  M2  M3 "]
        assert np.array_equal(at["M1  M2  M3 "][0], v) and at["M1  M2  M3 "][1:] == (h5read.H5T_INTEGER, 4)


# ---- GPU: files written from device fields --------------------------------

def _pair(spec):
    from test_gpu_parity import pair
    return pair(spec)


def _file_extent(oc, vel_dir, stride):
    """filespace_props for one block (src/mod_inout.f90:2713-2815)."""
    Sw, Mw = [], []
    for d in range(3):
        M, periodic = oc.N[d], oc.bc[2 * d] == -1
        if periodic:
            Sw.append(1)
            Mw.append((M - 1) // stride[d])
        elif vel_dir == d + 1:
            Sw.append(0)
            Mw.append(M)
        else:
            Sw.append(1)
            Mw.append((M - 1) // stride[d] + 1)
    return Sw, Mw


def _check_attrs(at, oc, st, Sw, Mw):
    assert set(at) == set(ATTRS), set(at) ^ set(ATTRS)
    for name, (kind, n) in ATTRS.items():
        v, cls, size = at[name]
        if isinstance(n, str):
            d = int(n[1]) - 1
            n = oc.N[d] + (1 if n.endswith("+1") else 0)
        assert v.shape == (n,), name
        assert (cls, size) == ((h5read.H5T_FLOAT, 8) if kind == "r" else (h5read.H5T_INTEGER, 4)), name
    val = {k: v[0] for k, v in at.items()}
    for k in ("restart", "write_count", "timestep", "write_out_vect", "new_dtime", "twostep_yes"):
        assert val[k][0] == getattr(st, k), k
    for k in ("time", "dtime", "time_out_vect", "time_out_scal", "dtime_out_vect", "Re", "thetaL", "U_ref",
              "L_ref", "p_end_systole_out_2"):
        assert val[k][0] == getattr(st, k), k
    assert val["CFL   "][0] == st.CFL and val["epsU  "][0] == st.epsU
    assert list(val["M1  M2  M3 "]) == list(oc.N)
    assert list(val["S1w S2w S3w"]) == Sw and list(val["M1w M2w M3w"]) == Mw
    assert list(val["NB1 NB2 NB3"]) == [1, 1, 1] and list(val["ls1 ls2 ls3"]) == [-1, -1, -1]
    assert list(val["BC_iL"]) == list(oc.bc[0::2]) and list(val["BC_iU"]) == list(oc.bc[1::2])
    # boundary stencil widths of src/mod_dims.f90:65-81
    for d in (1, 2, 3):
        assert list(val[f"ncb{d}c"]) == [4, 5, 5, 7]
        assert list(val[f"ncb{d}g"]) == [3, 4, 4, 6]
        assert list(val[f"ncb{d}d"]) == [4, 4, 6, 6]
    for d, (pn, un) in enumerate((("y1p", "y1u"), ("y2p", "y2v"), ("y3p", "y3w"))):
        xp, xu = oc.coords(d, 0), oc.coords(d, 1)
        assert np.array_equal(val[pn], xp[4:oc.N[d] + 4]), pn
        assert np.array_equal(val[un], xu[3:oc.N[d] + 4]), un


@pytest.mark.gpu
@pytest.mark.parametrize("spec", [pytest.param(WALLS, id="walls"), pytest.param(PERIODIC_XZ, id="periodic-xz")])
@pytest.mark.parametrize("stride", [(1, 1, 1), (2, 2, 2), (4, 2, 1)])
def test_write_hdf_pressure(tmp_path, spec, stride):
    from test_gpu_parity import rnd
    oc, gc = _pair(spec)
    pre = rnd(oc.shape, 400)
    gc.upload(ih.PRE, pre)
    st = state()
    gc.write_hdf(str(tmp_path / "pre_large_00012"), "pre", ih.PRE, 0, stride, st)
    Sw, Mw = _file_extent(oc, 0, stride)
    with h5read.File(tmp_path / "pre_large_00012.h5") as r:
        assert sorted(r.names()) == ["VectorX", "VectorY", "VectorZ", "pre"]
        data, cls, size = r.dataset("pre")
        assert (cls, size) == (h5read.H5T_FLOAT, 8)
        dims = [Mw[d] - Sw[d] + 1 for d in range(3)]
        assert data.shape == (dims[2], dims[1], dims[0])
        # the field's values at S?p, S?p + s, ... (x fastest)
        sl = tuple(slice(oc.Sp[d] + 3, oc.Sp[d] + 3 + stride[d] * dims[d], stride[d]) for d in (2, 1, 0))
        assert np.array_equal(data, pre[sl])
        _check_attrs(r.attributes("pre"), oc, st, Sw, Mw)
        for d, vn in enumerate(("VectorX", "VectorY", "VectorZ")):
            v, _, _ = r.dataset(vn)
            assert np.array_equal(v, oc.coords(d, 0)[Sw[d] + 3::stride[d]][:dims[d]]), vn


@pytest.mark.gpu
@pytest.mark.parametrize("spec", [pytest.param(WALLS, id="walls"), pytest.param(PERIODIC_XZ, id="periodic-xz")])
def test_restart_round_trip(tmp_path, spec):
    from test_gpu_parity import rnd
    oc, gc = _pair(spec)
    fields = {f: rnd(oc.shape, 410 + f) for f in (ih.VEL1, ih.VEL2, ih.VEL3, ih.PRE)}
    for f, a in fields.items():
        gc.upload(f, a)
    st = state(restart=7, time=1.25, dtime=0.01, time_out_vect=2.0, dtime_out_vect=0.25, time_out_scal=1.3,
               dtime_out_scal=0.5)
    prefix = str(tmp_path) + "/"
    gc.write_restart(prefix, st)
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["pre_restart007.h5", "velX_restart007.h5", "velY_restart007.h5", "velZ_restart007.h5"]
    # the velocity files: component m over SmB..NmB, y?u along m
    for m, nm in enumerate(("velX_restart", "velY_restart", "velZ_restart")):
        with h5read.File(tmp_path / f"{nm}007.h5") as r:
            data, _, _ = r.dataset(nm)
            lo, hi = oc.SB[m], oc.NB[m]
            assert data.shape == tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
            assert np.array_equal(data, fields[ih.VEL1 + m][tuple(slice(lo[d] + 3, hi[d] + 4) for d in (2, 1, 0))])
            Sw, Mw = _file_extent(oc, m + 1, (1, 1, 1))
            assert data.shape == tuple(Mw[d] - Sw[d] + 1 for d in (2, 1, 0))
            v, _, _ = r.dataset(("VectorX", "VectorY", "VectorZ")[m])
            assert np.array_equal(v, oc.coords(m, 1)[Sw[m] + 3:Mw[m] + 4])
    for f in fields:
        gc.upload(f, oc.zeros())
    # configured intervals 0.5 (vect) and 0.05 (scal): (2.0 - 0.25) + 0.5 = 2.25 stays;
    # (1.3 - 0.5) + 0.05 = 0.85 is not after time 1.25, so it moves to time + dtime
    st2 = state(dtime_out_vect=0.5, dtime_out_scal=0.05, time=0.0, timestep=0, write_count=0)
    gc.read_restart(prefix, 7, st2)
    assert st2.restart == 7 and st2.time == 1.25 and st2.dtime == 0.01 and st2.timestep == st.timestep
    assert st2.write_count == st.write_count and st2.new_dtime == st.new_dtime
    assert st2.time_out_vect == (2.0 - 0.25) + 0.5
    assert st2.time_out_scal == 1.25 + 0.01
    assert st2.p_end_systole_out == st.p_end_systole_out
    for m in range(3):
        s = tuple(slice(oc.SB[m][d] + 3, oc.NB[m][d] + 4) for d in (2, 1, 0))
        np.testing.assert_array_equal(gc.download(ih.VEL1 + m)[s], fields[ih.VEL1 + m][s])
    s = tuple(slice(oc.Sp[d] + 3, oc.Np[d] + 4) for d in (2, 1, 0))
    np.testing.assert_array_equal(gc.download(ih.PRE)[s], fields[ih.PRE][s])


@pytest.mark.gpu
def test_write_fields_and_xdmf(tmp_path):
    """velX/Y/Z of write_fields are the oracle's exchange(m, m, vel_m) then
    interpolate2_vel_pre(m); the medium size is its stride-2 subsample."""
    from test_gpu_parity import rnd
    oc, gc = _pair(WALLS)
    vel = np.ascontiguousarray(np.stack([rnd(oc.shape, 420 + m) for m in range(3)]))
    pre = rnd(oc.shape, 425)
    for m in range(3):
        gc.upload(ih.VEL1 + m, vel[m])
    gc.upload(ih.PRE, pre)
    st = state(write_count=4, time=0.123456789, time_out_vect=0.5, dtime_out_vect=0.05)
    prefix = str(tmp_path) + "/"
    sizes, strides = (1, 1, 0), ((1, 1, 1), (2, 2, 2), (1, 1, 1))
    gc.write_xdmf(prefix, sizes, strides, st)
    gc.write_fields(prefix, sizes, strides, st)
    assert st.write_count == 5 and st.time_out_vect == 0.5 + 0.05 and st.write_out_vect == 0
    s = tuple(slice(oc.Sp[d] + 3, oc.Np[d] + 4) for d in (2, 1, 0))
    for m, vn in enumerate(("velX", "velY", "velZ")):
        v = vel[m].copy()
        inter = oc.zeros()
        oc.exchange(m + 1, m + 1, v)
        oc.interpolate2_vel_pre(m + 1, v, inter)
        with h5read.File(tmp_path / f"{vn}_large_00004.h5") as r:
            big, _, _ = r.dataset(vn)
        assert np.array_equal(big, inter[s]), vn
        with h5read.File(tmp_path / f"{vn}_med_00004.h5") as r:
            med, _, _ = r.dataset(vn)
        assert np.array_equal(med, big[::2, ::2, ::2]), vn
    with h5read.File(tmp_path / "pre_large_00004.h5") as r:
        assert np.array_equal(r.dataset("pre")[0], pre[s])
    assert not list(tmp_path.glob("*_small_*"))
    # the XDMF of each size, pointing at the files written with the same count
    for tag, stride in (("large", (1, 1, 1)), ("med", (2, 2, 2))):
        path = tmp_path / f"IMPACT_xdmf3d_{tag}_00004.xmf"
        text = path.read_text().splitlines()
        assert text[0] == '<?xml version="1.0" ?>'
        assert text[1] == '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>'
        dims = [(oc.N[d] - 1) // stride[d] + 1 for d in range(3)]
        assert f'      <Topology TopologyType="3DRectMesh" Dimensions="{dims[2]} {dims[1]} {dims[0]}"/>' in text
        # Fortran f15.9: width 15, nine decimals
        assert '        <Time TimeType="Single" Value="    0.123456789"/>' in text
        root = ET.parse(path).getroot()
        items = [e.text.strip() for e in root.iter("DataItem")]
        assert items[:3] == [f"pre_{tag}_00004.h5:/Vector{a}" for a in "XYZ"]
        assert items[3:] == [f"{n}_{tag}_00004.h5:/{n}" for n in ("pre", "velX", "velY", "velZ")]
        for it in items:
            assert (tmp_path / it.split(":")[0]).exists()
        with h5read.File(tmp_path / f"pre_{tag}_00004.h5") as r:
            assert r.dataset("pre")[0].shape == (dims[2], dims[1], dims[0])


@pytest.mark.gpu
def test_write_hdf_rejects_what_the_reference_rejects(tmp_path):
    oc, gc = _pair(WALLS)
    st = state()
    with pytest.raises(ih.IHError, match="stride /= 1 in the corresponding direction"):
        gc.write_hdf(str(tmp_path / "u"), "velX", ih.VEL1, 1, (2, 1, 1), st)
    with pytest.raises(ih.IHError, match="Cannot write field with this stride"):
        gc.write_hdf(str(tmp_path / "p"), "pre", ih.PRE, 0, (3, 1, 1), st)
    with pytest.raises(ih.IHError, match="Cannot open file"):
        gc.read_hdf(str(tmp_path / "missing"), "pre", ih.PRE, 0)
    assert not list(tmp_path.iterdir())


@pytest.mark.gpu
def test_field_box_transfer():
    """ih_field_get_box / ih_field_set_box: the strided box equals the numpy
    slice of the field; boxes outside -3..N+3 and strides below 1 are rejected
    before anything runs on the device."""
    from test_gpu_parity import rnd
    oc, gc = _pair(PERIODIC_XZ)
    a = rnd(oc.shape, 430)
    gc.upload(ih.TMP2, a)
    for lo, hi, st in (((-3, -3, -3), (oc.N[0] + 3, oc.N[1] + 3, oc.N[2] + 3), (1, 1, 1)),
                       ((1, 0, 2), (31, 17, 30), (3, 2, 4)), ((5, 5, 5), (5, 5, 5), (1, 1, 1))):
        got = gc.get_box(ih.TMP2, lo, hi, st)
        want = a[tuple(slice(lo[d] + 3, hi[d] + 4, st[d]) for d in (2, 1, 0))]
        assert np.array_equal(got, want), (lo, hi, st)
    # scatter a box back: only its points change
    lo, hi, st = (2, 1, 3), (30, 16, 27), (2, 3, 4)
    box_vals = np.full(gc.get_box(ih.TMP2, lo, hi, st).shape, 7.5)
    gc.set_box(ih.TMP2, lo, hi, st, box_vals)
    b = a.copy()
    b[tuple(slice(lo[d] + 3, hi[d] + 4, st[d]) for d in (2, 1, 0))] = 7.5
    assert np.array_equal(gc.download(ih.TMP2), b)
    for lo, hi, st in (((-4, 0, 0), (1, 1, 1), (1, 1, 1)), ((0, 0, 0), (1, 1, oc.N[2] + 4), (1, 1, 1)),
                       ((0, 0, 0), (1, 1, 1), (1, 0, 1)), ((2, 0, 0), (1, 1, 1), (1, 1, 1))):
        with pytest.raises(ih.IHError, match="outside the field's index range"):
            gc.get_box(ih.TMP2, lo, hi, st)
    assert np.array_equal(gc.download(ih.TMP2), b)


@pytest.mark.gpu
def test_read_hdf_places_the_file_as_read2_hdf_does(tmp_path):
    """read2_hdf reads the file's S?w..M?w (its attributes) into the overlap with
    the target range, from file offset 0 (src/mod_inout.f90:1604-1690), after
    setting the whole field to 0 (:1675): a stride-2 output lands in the first
    points of the range, a pressure file read as velocity 1 fills the points both
    ranges share, and every other point (ghosts included) reads 0."""
    from test_gpu_parity import rnd
    oc, gc = _pair(WALLS)
    pre = rnd(oc.shape, 440)
    gc.upload(ih.PRE, pre)
    st = state()
    gc.write_hdf(str(tmp_path / "p2"), "pre", ih.PRE, 0, (2, 2, 2), st)
    base = rnd(oc.shape, 441)
    gc.upload(ih.TMP3, base)
    st2 = state(time=0.0, timestep=0)
    gc.read_hdf(str(tmp_path / "p2"), "pre", ih.TMP3, 0, st2)
    assert st2.time == st.time and st2.timestep == st.timestep and st2.write_count == st.write_count
    dims = [(oc.N[d] - 1) // 2 + 1 for d in range(3)]
    want = np.zeros_like(base)  # read2_hdf sets phi = 0. first (src/mod_inout.f90:1675)
    strided = pre[tuple(slice(4, 4 + 2 * dims[d], 2) for d in (2, 1, 0))]
    want[tuple(slice(4, 4 + dims[d]) for d in (2, 1, 0))] = strided
    assert np.array_equal(gc.download(ih.TMP3), want)
    # pressure file (S?w = 1 .. M?w = N) into velocity 1 (SB = 0 .. NB = N in x)
    gc.write_hdf(str(tmp_path / "p1"), "pre", ih.PRE, 0, (1, 1, 1), st)
    gc.upload(ih.TMP3, base)
    gc.read_hdf(str(tmp_path / "p1"), "pre", ih.TMP3, 1)
    want = np.zeros_like(base)
    s = tuple(slice(4, oc.N[d] + 4) for d in (2, 1, 0))
    want[s] = pre[s]
    assert np.array_equal(gc.download(ih.TMP3), want)
