// HDF5 field output, restart and XDMF (src/mod_inout.f90) over the C ABI of
// libimpact_hip: fields leave and enter the device as strided boxes
// (ih_field_get_box / ih_field_set_box), so only the written points cross the
// host link.  One block: M = N, iShift = 0, iB = 1, so every file offset is 0.
#include "impact_io.h"

#include <hdf5.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_io_err = buf;
    return code;
}

// a failed call of the case library: its message becomes ours
int lib_fail(int rc, const char *what) {
    char buf[400] = {0};
    ih_last_error(buf, sizeof buf);
    return io_fail(rc, "%s: %s", what, buf);
}

// ls1 = ls2 = ls3 = -1 (src/mod_vars.f90:689); ncb?c/g/d as the coefficient
// generator uses them (src/mod_dims.f90:64-84)
constexpr int LS = -1;
const int NCBC[4] = {4, 5, 5, 7};
const int NCBG[4] = {3, 4, 4, 6};
const int NCBD[4] = {4, 4, 6, 6};

struct Grid {
    int N[3];
    int bcL[3], bcU[3];
    int Sp[3], Np[3], SB[3][3], NB[3][3];  // SB[m][d]
    std::vector<double> xp[3], xu[3];      // (-3:N+3) at [i+3]
};

int load_grid(ih_case *c, Grid &G, const ih_grid_desc *gd) {
    int r[48];
    int rc = ih_case_get_ranges(c, r);
    if (rc) return lib_fail(rc, "ranges");
    for (int d = 0; d < 3; ++d) {
        G.N[d] = gd->N[d];
        G.bcL[d] = gd->bc[2 * d];
        G.bcU[d] = gd->bc[2 * d + 1];
        G.Sp[d] = r[d];
        G.Np[d] = r[3 + d];
        for (int m = 0; m < 3; ++m) {
            G.SB[m][d] = r[6 + 3 * m + d];
            G.NB[m][d] = r[15 + 3 * m + d];
        }
        G.xp[d].resize((size_t)G.N[d] + 7);
        G.xu[d].resize((size_t)G.N[d] + 7);
        if ((rc = ih_case_get_coords(c, d, 0, G.xp[d].data()))) return lib_fail(rc, "coordinates");
        if ((rc = ih_case_get_coords(c, d, 1, G.xu[d].data()))) return lib_fail(rc, "coordinates");
    }
    return IH_OK;
}

// the grid the case was created with
int case_grid(ih_case *c, Grid &G) {
    ih_grid_desc gd{};
    int rc = ih_case_get_grid(c, &gd);
    if (rc) return lib_fail(rc, "grid");
    return load_grid(c, G, &gd);
}

// filespace_props for the first block (src/mod_inout.f90:2713-2815)
void filespace_props(const Grid &G, int vel_dir, const int stride[3], int Sw[3], int Mw[3], int dim[3]) {
    for (int d = 0; d < 3; ++d) {
        const int M = G.N[d];
        Sw[d] = 2 + LS;
        Mw[d] = (M - 1) / stride[d] + 1 + LS;
        if (G.bcL[d] != -1) {
            if (vel_dir == d + 1) {
                Sw[d] = G.bcL[d] == -2 ? 1 : 0;
                Mw[d] = G.bcU[d] == -2 ? M - 1 : M;
            } else {
                Sw[d] = 1;
                Mw[d] = (M - 1) / stride[d] + 1;
            }
        }
        dim[d] = Mw[d] - Sw[d] + 1;
    }
}

void range_of(const Grid &G, int vel_dir, int SS[3], int NN[3]) {
    for (int d = 0; d < 3; ++d) {
        SS[d] = vel_dir == 0 ? G.Sp[d] : G.SB[vel_dir - 1][d];
        NN[d] = vel_dir == 0 ? G.Np[d] : G.NB[vel_dir - 1][d];
    }
}

// RAII for HDF5 identifiers
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    operator hid_t() const { return id; }
    bool bad() const { return id < 0; }
};

// write_hdf_infoREAL / INT / LOG: a 1-D attribute of the dataset (attr = true)
// or a 1-D dataset of the file, in the file as the native type
bool put_info(hid_t loc, bool attr, const char *name, hid_t type, const void *v, hsize_t n) {
    Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (sp.bad()) return false;
    if (attr) {
        Hid a(H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        return !a.bad() && H5Awrite(a, type, v) >= 0;
    }
    Hid ds(H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return !ds.bad() && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
}

bool put_r(hid_t loc, const char *name, double v) { return put_info(loc, true, name, H5T_NATIVE_DOUBLE, &v, 1); }
bool put_i(hid_t loc, const char *name, int v) { return put_info(loc, true, name, H5T_NATIVE_INT, &v, 1); }
bool put_l(hid_t loc, const char *name, int v) {
    const int b = v ? 1 : 0;
    return put_info(loc, true, name, H5T_NATIVE_INT, &b, 1);
}
bool put_i3(hid_t loc, const char *name, const int v[3]) { return put_info(loc, true, name, H5T_NATIVE_INT, v, 3); }

bool get_attr(hid_t loc, const char *name, hid_t type, void *v) {
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    return !a.bad() && H5Aread(a, type, v) >= 0;
}

std::string num_to_string(int digits, int num) {
    char b[16];
    std::snprintf(b, sizeof b, "%0*d", digits, num);
    return b;
}

bool num_ok(int digits, int num) {
    int lim = 1;
    for (int q = 0; q < digits; ++q) lim *= 10;
    return num >= 0 && num < lim;
}

}  // namespace

extern "C" {

int ih_io_abi_version(void) { return IH_IO_ABI_VERSION; }

int ih_io_last_error(char *buf, size_t n) {
    if (!buf || n == 0) return IH_EINVAL;
    std::snprintf(buf, n, "%s", g_io_err.c_str());
    return IH_OK;
}

int ih_io_write_hdf(ih_case *c, const char *filename, const char *dsetname, ih_field f, int vel_dir,
                    const int stride[3], const ih_io_state *st) {
    if (!c || !filename || !dsetname || !stride || !st || vel_dir < 0 || vel_dir > 3)
        return io_fail(IH_EINVAL, "bad argument");
    if (st->version != IH_IO_ABI_VERSION) return io_fail(IH_EINVAL, "ih_io_state.version %d", st->version);
    Grid G;
    int rc = case_grid(c, G);
    if (rc) return rc;
    for (int d = 0; d < 3; ++d)
        if (stride[d] < 1) return io_fail(IH_EINVAL, "stride %d", stride[d]);
    // the two checks of write_hdf (src/mod_inout.f90:538-550)
    if (vel_dir != 0 && stride[vel_dir - 1] != 1)
        return io_fail(IH_EINVAL, "Cannot write velocity field with stride /= 1 in the corresponding direction");
    for (int d = 0; d < 3; ++d)
        if ((G.N[d] - 1) % stride[d] != 0) return io_fail(IH_EINVAL, "Cannot write field with this stride");
    int Sw[3], Mw[3], dim[3], SS[3], NN[3], dd[3];
    filespace_props(G, vel_dir, stride, Sw, Mw, dim);
    range_of(G, vel_dir, SS, NN);
    int hi[3];
    for (int d = 0; d < 3; ++d) {
        dd[d] = (NN[d] - SS[d]) / stride[d] + 1;
        hi[d] = SS[d] + (dd[d] - 1) * stride[d];
        if (dd[d] > dim[d]) return io_fail(IH_EINVAL, "range exceeds the file extent along %d", d + 1);
    }
    std::vector<double> buf((size_t)dd[0] * dd[1] * dd[2]);
    if ((rc = ih_field_get_box(c, f, SS, hi, stride, buf.data()))) return lib_fail(rc, "field box");

    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    const std::string path = std::string(filename) + ".h5";
    Hid file(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.bad()) return io_fail(IH_EINVAL, "cannot create %s", path.c_str());
    // Fortran (dim1, dim2, dim3) is C (dim3, dim2, dim1): x fastest in both
    const hsize_t fdims[3] = {(hsize_t)dim[2], (hsize_t)dim[1], (hsize_t)dim[0]};
    const hsize_t cnt[3] = {(hsize_t)dd[2], (hsize_t)dd[1], (hsize_t)dd[0]};
    const hsize_t off[3] = {0, 0, 0};
    Hid fsp(H5Screate_simple(3, fdims, nullptr), H5Sclose);
    Hid msp(H5Screate_simple(3, cnt, nullptr), H5Sclose);
    Hid ds(H5Dcreate2(file, dsetname, H5T_NATIVE_DOUBLE, fsp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (fsp.bad() || msp.bad() || ds.bad()) return io_fail(IH_EINVAL, "cannot create dataset %s", dsetname);
    // the attributes, names and order of src/mod_inout.f90:596-656
    const int M[3] = {G.N[0], G.N[1], G.N[2]};
    const int NBv[3] = {1, 1, 1}, lsv[3] = {LS, LS, LS};
    const int bcl[3] = {G.bcL[0], G.bcL[1], G.bcL[2]}, bcu[3] = {G.bcU[0], G.bcU[1], G.bcU[2]};
    bool ok = put_i(ds, "restart", st->restart) && put_i(ds, "write_count", st->write_count) &&
              put_r(ds, "time_out_vect", st->time_out_vect) && put_r(ds, "time_out_scal", st->time_out_scal) &&
              put_r(ds, "dtime_out_vect", st->dtime_out_vect) && put_r(ds, "dtime_out_scal", st->dtime_out_scal) &&
              put_l(ds, "write_out_vect", st->write_out_vect) && put_l(ds, "write_out_scal", st->write_out_scal) &&
              put_l(ds, "new_dtime", st->new_dtime) && put_i3(ds, "M1  M2  M3 ", M) && put_i3(ds, "S1w S2w S3w", Sw) &&
              put_i3(ds, "M1w M2w M3w", Mw) && put_i3(ds, "NB1 NB2 NB3", NBv) && put_i3(ds, "ls1 ls2 ls3", lsv) &&
              put_r(ds, "time", st->time) && put_r(ds, "dtime", st->dtime) && put_i(ds, "timestep", st->timestep) &&
              put_r(ds, "Re", st->Re) && put_r(ds, "CFL   ", st->CFL) && put_r(ds, "thetaL", st->thetaL) &&
              put_r(ds, "epsU  ", st->epsU) && put_r(ds, "p_end_systole_in", st->p_end_systole_in) &&
              put_r(ds, "p_end_systole_out", st->p_end_systole_out) &&
              put_r(ds, "p_end_systole_out_2", st->p_end_systole_out_2) && put_l(ds, "Euler_yes", st->Euler_yes) &&
              put_l(ds, "twostep_yes", st->twostep_yes) && put_l(ds, "mapping_yes", st->mapping_yes) &&
              put_l(ds, "upwind_yes", st->upwind_yes) && put_i3(ds, "BC_iL", bcl) && put_i3(ds, "BC_iU", bcu);
    const char *ncbn[9] = {"ncb1c", "ncb1g", "ncb1d", "ncb2c", "ncb2g", "ncb2d", "ncb3c", "ncb3g", "ncb3d"};
    for (int q = 0; ok && q < 9; ++q) {
        const int *v = q % 3 == 0 ? NCBC : (q % 3 == 1 ? NCBG : NCBD);
        ok = put_info(ds, true, ncbn[q], H5T_NATIVE_INT, v, 4);
    }
    const char *pn[3] = {"y1p", "y2p", "y3p"}, *un[3] = {"y1u", "y2v", "y3w"};
    for (int d = 0; ok && d < 3; ++d)  // y?p(1:M), y?u(0:M)
        ok = put_info(ds, true, pn[d], H5T_NATIVE_DOUBLE, G.xp[d].data() + 4, (hsize_t)M[d]);
    for (int d = 0; ok && d < 3; ++d)
        ok = put_info(ds, true, un[d], H5T_NATIVE_DOUBLE, G.xu[d].data() + 3, (hsize_t)M[d] + 1);
    ok = ok && put_r(ds, "U_ref", st->U_ref) && put_r(ds, "L_ref", st->L_ref);
    if (!ok) return io_fail(IH_EINVAL, "cannot write the attributes of %s", dsetname);
    if (H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off, nullptr, cnt, nullptr) < 0 ||
        H5Dwrite(ds, H5T_NATIVE_DOUBLE, msp, fsp, H5P_DEFAULT, buf.data()) < 0)
        return io_fail(IH_EINVAL, "cannot write dataset %s", dsetname);
    // VectorX/Y/Z: coordinates of the written points, y?u along vel_dir
    const char *vn[3] = {"VectorX", "VectorY", "VectorZ"};
    for (int d = 0; d < 3; ++d) {
        const std::vector<double> &x = vel_dir == d + 1 ? G.xu[d] : G.xp[d];
        std::vector<double> v((size_t)dim[d]);
        for (int q = 0; q < dim[d]; ++q) v[(size_t)q] = x[(size_t)(Sw[d] + q * stride[d] + 3)];
        if (!put_info(file, false, vn[d], H5T_NATIVE_DOUBLE, v.data(), (hsize_t)dim[d]))
            return io_fail(IH_EINVAL, "cannot write %s", vn[d]);
    }
    return IH_OK;
}

int ih_io_read_hdf(ih_case *c, const char *filename, const char *dsetname, ih_field f, int vel_dir,
                   ih_io_state *st) {
    if (!c || !filename || !dsetname || vel_dir < 0 || vel_dir > 3) return io_fail(IH_EINVAL, "bad argument");
    Grid G;
    int rc = case_grid(c, G);
    if (rc) return rc;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    const std::string path = std::string(filename) + ".h5";
    Hid file(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.bad()) return io_fail(IH_EINVAL, "Cannot open file %s", path.c_str());
    Hid ds(H5Dopen2(file, dsetname, H5P_DEFAULT), H5Dclose);
    if (ds.bad()) return io_fail(IH_EINVAL, "Cannot open dataset %s", dsetname);
    ih_io_state s{};
    int Sw[3], Mw[3];
    bool ok = get_attr(ds, "time", H5T_NATIVE_DOUBLE, &s.time) && get_attr(ds, "dtime", H5T_NATIVE_DOUBLE, &s.dtime) &&
              get_attr(ds, "time_out_vect", H5T_NATIVE_DOUBLE, &s.time_out_vect) &&
              get_attr(ds, "time_out_scal", H5T_NATIVE_DOUBLE, &s.time_out_scal) &&
              get_attr(ds, "timestep", H5T_NATIVE_INT, &s.timestep) &&
              get_attr(ds, "write_count", H5T_NATIVE_INT, &s.write_count) &&
              get_attr(ds, "write_out_vect", H5T_NATIVE_INT, &s.write_out_vect) &&
              get_attr(ds, "write_out_scal", H5T_NATIVE_INT, &s.write_out_scal) &&
              get_attr(ds, "new_dtime", H5T_NATIVE_INT, &s.new_dtime) &&
              get_attr(ds, "S1w S2w S3w", H5T_NATIVE_INT, Sw) && get_attr(ds, "M1w M2w M3w", H5T_NATIVE_INT, Mw);
    if (!ok) return io_fail(IH_EINVAL, "%s: missing attributes", dsetname);
    // the overlap of the file's S?w..M?w with the range (src/mod_inout.f90:1604-1690, i0 = 0)
    int SS[3], NN[3], Sr[3], Nr[3], dd[3];
    range_of(G, vel_dir, SS, NN);
    hsize_t fdims[3];
    {
        Hid fsp(H5Dget_space(ds), H5Sclose);
        if (fsp.bad() || H5Sget_simple_extent_ndims(fsp) != 3) return io_fail(IH_EINVAL, "%s is not 3-D", dsetname);
        H5Sget_simple_extent_dims(fsp, fdims, nullptr);
    }
    bool any = true;
    for (int d = 0; d < 3; ++d) {
        if (Sw[d] <= NN[d] && Mw[d] >= SS[d]) {
            Sr[d] = Sw[d] >= SS[d] ? Sw[d] : SS[d];
            Nr[d] = Mw[d] <= NN[d] ? Mw[d] : NN[d];
            dd[d] = Nr[d] - Sr[d] + 1;
        } else {
            any = false;
        }
        if (any && (hsize_t)dd[d] > fdims[2 - d]) return io_fail(IH_EINVAL, "%s: extent below its attributes", dsetname);
    }
    // phi = 0. over the whole array before the hyperslab read (src/mod_inout.f90:1675):
    // ghosts and points outside the file's box do not keep earlier values
    if ((rc = ih_field_fill(c, f, 0.0))) return lib_fail(rc, "field fill");
    if (any) {
        const hsize_t cnt[3] = {(hsize_t)dd[2], (hsize_t)dd[1], (hsize_t)dd[0]};
        const hsize_t off[3] = {0, 0, 0};
        std::vector<double> buf((size_t)dd[0] * dd[1] * dd[2]);
        Hid fsp(H5Dget_space(ds), H5Sclose);
        Hid msp(H5Screate_simple(3, cnt, nullptr), H5Sclose);
        if (H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off, nullptr, cnt, nullptr) < 0 ||
            H5Dread(ds, H5T_NATIVE_DOUBLE, msp, fsp, H5P_DEFAULT, buf.data()) < 0)
            return io_fail(IH_EINVAL, "cannot read dataset %s", dsetname);
        const int one[3] = {1, 1, 1};
        if ((rc = ih_field_set_box(c, f, Sr, Nr, one, buf.data()))) return lib_fail(rc, "field box");
    }
    if (st) {
        st->time = s.time;
        st->dtime = s.dtime;
        st->time_out_vect = s.time_out_vect;
        st->time_out_scal = s.time_out_scal;
        st->timestep = s.timestep;
        st->write_count = s.write_count;
        st->write_out_vect = s.write_out_vect;
        st->write_out_scal = s.write_out_scal;
        st->new_dtime = s.new_dtime;
    }
    return IH_OK;
}

int ih_io_write_restart(ih_case *c, const char *prefix, const ih_io_state *st) {
    if (!c || !prefix || !st) return io_fail(IH_EINVAL, "bad argument");
    if (!num_ok(3, st->restart)) return io_fail(IH_EINVAL, "restart %d does not fit three digits", st->restart);
    const std::string p(prefix), r = num_to_string(3, st->restart);
    const int one[3] = {1, 1, 1};
    const char *nm[3] = {"velX_restart", "velY_restart", "velZ_restart"};
    for (int m = 0; m < 3; ++m) {
        int rc = ih_io_write_hdf(c, (p + nm[m] + r).c_str(), nm[m], (ih_field)(IH_VEL1 + m), m + 1, one, st);
        if (rc) return rc;
    }
    return ih_io_write_hdf(c, (p + "pre_restart" + r).c_str(), "pre_restart", IH_PRE, 0, one, st);
}

int ih_io_read_restart(ih_case *c, const char *prefix, int restart, ih_io_state *st) {
    if (!c || !prefix || !st) return io_fail(IH_EINVAL, "bad argument");
    if (!num_ok(3, restart)) return io_fail(IH_EINVAL, "restart %d does not fit three digits", restart);
    const std::string p(prefix), r = num_to_string(3, restart);
    const double dtv = st->dtime_out_vect, dts = st->dtime_out_scal;
    const char *nm[3] = {"velX_restart", "velY_restart", "velZ_restart"};
    for (int m = 0; m < 3; ++m) {
        int rc = ih_io_read_hdf(c, (p + nm[m] + r).c_str(), nm[m], (ih_field)(IH_VEL1 + m), m + 1, st);
        if (rc) return rc;
    }
    int rc = ih_io_read_hdf(c, (p + "pre_restart" + r).c_str(), "pre_restart", IH_PRE, 0, st);
    if (rc) return rc;
    // the output intervals stored with velX, then the configured ones (src/mod_inout.f90:452-487)
    const std::string path = p + "velX_restart" + r + ".h5";
    Hid file(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.bad()) return io_fail(IH_EINVAL, "Cannot open file %s", path.c_str());
    Hid ds(H5Dopen2(file, "velX_restart", H5P_DEFAULT), H5Dclose);
    double old_s = 0.0, old_v = 0.0;
    if (ds.bad() || !get_attr(ds, "dtime_out_scal", H5T_NATIVE_DOUBLE, &old_s) ||
        !get_attr(ds, "dtime_out_vect", H5T_NATIVE_DOUBLE, &old_v) ||
        !get_attr(ds, "p_end_systole_in", H5T_NATIVE_DOUBLE, &st->p_end_systole_in) ||
        !get_attr(ds, "p_end_systole_out", H5T_NATIVE_DOUBLE, &st->p_end_systole_out) ||
        !get_attr(ds, "p_end_systole_out_2", H5T_NATIVE_DOUBLE, &st->p_end_systole_out_2))
        return io_fail(IH_EINVAL, "%s: missing attributes", path.c_str());
    st->time_out_vect = (st->time_out_vect - old_v) + dtv;
    st->time_out_scal = (st->time_out_scal - old_s) + dts;
    if (st->time_out_vect <= st->time) st->time_out_vect = st->time + st->dtime;
    if (st->time_out_scal <= st->time) st->time_out_scal = st->time + st->dtime;
    st->restart = restart;
    return IH_OK;
}

int ih_io_write_fields(ih_case *c, const char *prefix, const int sizes[3], const int strides[9], ih_io_state *st) {
    if (!c || !prefix || !sizes || !strides || !st) return io_fail(IH_EINVAL, "bad argument");
    if (!num_ok(5, st->write_count)) return io_fail(IH_EINVAL, "write_count %d does not fit five digits", st->write_count);
    const std::string p(prefix), cc = num_to_string(5, st->write_count);
    const char *tag[3] = {"large", "med", "small"};
    const char *vn[3] = {"velX", "velY", "velZ"};
    for (int m = 0; m < 3; ++m) {
        // exchange(m, m, vel_m), then nl_m = cI?p-interpolation to the pressure grid
        int rc = ih_exchange(c, m + 1, m + 1, (ih_field)(IH_VEL1 + m));
        if (!rc) rc = ih_interpolate2_vel_pre(c, m + 1, (ih_field)(IH_VEL1 + m), IH_TMP1);
        if (rc) return lib_fail(rc, "velocity interpolation");
        for (int q = 0; q < 3; ++q) {
            if (!sizes[q]) continue;
            rc = ih_io_write_hdf(c, (p + vn[m] + "_" + tag[q] + "_" + cc).c_str(), vn[m], IH_TMP1, 0, strides + 3 * q,
                                 st);
            if (rc) return rc;
        }
    }
    for (int q = 0; q < 3; ++q) {
        if (!sizes[q]) continue;
        int rc = ih_io_write_hdf(c, (p + "pre_" + tag[q] + "_" + cc).c_str(), "pre", IH_PRE, 0, strides + 3 * q, st);
        if (rc) return rc;
    }
    st->write_count = st->write_count + 1;
    st->time_out_vect = st->time_out_vect + st->dtime_out_vect;
    st->write_out_vect = 0;
    return IH_OK;
}

int ih_io_write_xdmf(ih_case *c, const char *prefix, const int sizes[3], const int strides[9],
                     const ih_io_state *st) {
    if (!c || !prefix || !sizes || !strides || !st) return io_fail(IH_EINVAL, "bad argument");
    if (!num_ok(5, st->write_count)) return io_fail(IH_EINVAL, "write_count %d does not fit five digits", st->write_count);
    Grid G;
    int rc = case_grid(c, G);
    if (rc) return rc;
    const std::string cc = num_to_string(5, st->write_count);
    const char *tag[3] = {"large", "med", "small"};
    for (int q = 0; q < 3; ++q) {
        if (!sizes[q]) continue;
        int Sw[3], Mw[3], dim[3];
        filespace_props(G, 0, strides + 3 * q, Sw, Mw, dim);
        const std::string t = tag[q];
        const std::string path = std::string(prefix) + "IMPACT_xdmf3d_" + t + "_" + cc + ".xmf";
        FILE *x = std::fopen(path.c_str(), "w");
        if (!x) return io_fail(IH_EINVAL, "cannot create %s", path.c_str());
        // write_3D_xdmf (src/mod_inout.f90:3162-3290), unscaled
        const std::string pre = "pre_" + t + "_" + cc + ".h5";
        std::fprintf(x, "<?xml version=\"1.0\" ?>\n");
        std::fprintf(x, "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
        std::fprintf(x, "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"2.0\">\n");
        std::fprintf(x, "  <Domain>\n");
        std::fprintf(x, "    <Grid Name=\"IMPACT mesh\" GridType=\"Uniform\">\n");
        std::fprintf(x, "      <Topology TopologyType=\"3DRectMesh\" Dimensions=\"%d %d %d\"/>\n", dim[2], dim[1], dim[0]);
        std::fprintf(x, "        <Geometry GeometryType=\"VXVYVZ\">\n");
        const char *vec[3] = {"VectorX", "VectorY", "VectorZ"};
        for (int d = 0; d < 3; ++d) {
            std::fprintf(x, "          <DataItem Format=\"HDF\" Dimensions=\"%d\" NumberType=\"Float\" Precision=\"8\">\n",
                         dim[d]);
            std::fprintf(x, "            %s:/%s\n", pre.c_str(), vec[d]);
            std::fprintf(x, "          </DataItem>\n");
        }
        std::fprintf(x, "        </Geometry>\n");
        std::fprintf(x, "        <Time TimeType=\"Single\" Value=\"%15.9f\"/>\n", st->time);
        const char *an[4] = {"Pressure", "Velocity_X", "Velocity_Y", "Velocity_Z"};
        const char *fn[4] = {"pre", "velX", "velY", "velZ"};
        for (int a = 0; a < 4; ++a) {
            std::fprintf(x, "        <Attribute Name=\"%s\" Center=\"Node\">\n", an[a]);
            std::fprintf(x,
                         "          <DataItem Format=\"HDF\" Dimensions=\"%d %d %d\" NumberType=\"Float\" "
                         "Precision=\"8\">\n",
                         dim[2], dim[1], dim[0]);
            std::fprintf(x, "            %s_%s_%s.h5:/%s\n", fn[a], t.c_str(), cc.c_str(), fn[a]);
            std::fprintf(x, "          </DataItem>\n");
            std::fprintf(x, "        </Attribute>\n");
        }
        std::fprintf(x, "    </Grid>\n");
        std::fprintf(x, "  </Domain>\n");
        std::fprintf(x, "</Xdmf>\n");
        if (std::fclose(x) != 0) return io_fail(IH_EINVAL, "cannot write %s", path.c_str());
    }
    return IH_OK;
}

}  // extern "C"
