// CT geometry on the host: the label volume of a legacy VTK image, inlet and
// outlet detection, and the echo waveform in reference units.
//
//   read_vtk_3D       src/usr_func.f90:187-555
//   convert_waveform  src/usr_func.f90:599-651 (CT-based branch)
//   get_pulse_number_and_length, get_shifted_time_in_pulse, the waveform
//   lookup of apply_CT_forcing  src/usr_func.f90:753-760, 893-903, 930-937
//
// Labels are uint8 on the case's field layout.  The reference keeps them as
// REAL(8) and compares them with ==, < 0.5 and the band [0.5, 2.0), so
// integer labels carry the same decisions; an image value that is not an
// integer in 0..255 is rejected.
#include "ih_host.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ih {

namespace {

bool next_token(FILE *f, std::string &t) {
    t.clear();
    int ch;
    do {
        ch = std::fgetc(f);
    } while (ch != EOF && (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == ','));
    if (ch == EOF) return false;
    while (ch != EOF && !(ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == ',')) {
        t.push_back((char)ch);
        ch = std::fgetc(f);
    }
    return true;
}

bool rest_of_line(FILE *f) {
    int ch;
    while ((ch = std::fgetc(f)) != EOF && ch != '\n') {
    }
    return ch != EOF;
}

}  // namespace

bool vtk_read_ascii(const char *path, int dims[3], std::vector<double> &values, std::string &err) {
    FILE *f = std::fopen(path, "r");
    if (!f) {
        err = "cannot open the VTK file";
        return false;
    }
    std::string t;
    bool ok = rest_of_line(f) && rest_of_line(f);  // version line, title
    ok = ok && next_token(f, t) && t == "ASCII";
    ok = ok && next_token(f, t) && next_token(f, t);  // DATASET <type>
    ok = ok && next_token(f, t) && t == "DIMENSIONS";
    for (int d = 0; ok && d < 3; ++d) ok = next_token(f, t) && (dims[d] = std::atoi(t.c_str())) > 0;
    ok = ok && next_token(f, t) && t == "SPACING" && rest_of_line(f);
    ok = ok && next_token(f, t) && t == "ORIGIN" && rest_of_line(f);
    long np = 0;
    ok = ok && next_token(f, t) && t == "POINT_DATA" && next_token(f, t) && (np = std::atol(t.c_str())) > 0;
    ok = ok && next_token(f, t) && t == "SCALARS" && rest_of_line(f);
    ok = ok && next_token(f, t) && t == "LOOKUP_TABLE" && rest_of_line(f);
    ok = ok && np == (long)dims[0] * dims[1] * dims[2];
    if (!ok) {
        std::fclose(f);
        err = "not a legacy ASCII VTK image (DIMENSIONS, SPACING, ORIGIN, POINT_DATA, SCALARS, LOOKUP_TABLE)";
        return false;
    }
    values.resize((size_t)np);
    for (long q = 0; q < np; ++q) {
        if (!next_token(f, t)) {
            std::fclose(f);
            err = "the VTK file holds fewer values than POINT_DATA says";
            return false;
        }
        values[(size_t)q] = std::strtod(t.c_str(), nullptr);
    }
    std::fclose(f);
    return true;
}

void waveform_convert(const double *t_ms, const double *q, int npts, int M1, std::vector<double> &t_out,
                      std::vector<double> &q_out, double ref[3]) {
    // CT pixel width 0.0004 m, upsampling 1, peak inflow 1 in reference units
    const double pixel = 0.0004, pixel_star = (1.0 / M1) * 1.0;
    double vmax = q[0];
    for (int i = 1; i < npts; ++i) vmax = q[i] > vmax ? q[i] : vmax;
    const double velocity = vmax / 1.0;
    const double length = pixel / pixel_star;
    const double time = length / velocity;
    t_out.resize((size_t)npts);
    q_out.resize((size_t)npts);
    for (int i = 0; i < npts; ++i) {
        t_out[(size_t)i] = t_ms[i] / (1000.0 * time);
        q_out[(size_t)i] = q[i] / velocity;
    }
    ref[0] = velocity;
    ref[1] = length;
    ref[2] = time;
}

// the inflow of apply_CT_forcing at `time`: shifted into the current pulse, the
// last interval [t_i, t_i+1] holding it gives Q_i
double waveform_inflow(const std::vector<double> &t, const std::vector<double> &q, double time, bool &found) {
    const size_t n = t.size();
    const double pulse = t[n - 1];
    const double number = std::floor(time / pulse) + 1.0;
    const double shifted = time - (number - 1.0) * pulse;
    double v = 0.0;
    found = false;
    for (size_t i = 0; i + 1 < n; ++i)
        if (shifted >= t[i] && shifted <= t[i + 1]) {
            v = q[i] * 1.0;  // flow_rate_correction_factor
            found = true;
        }
    return v;
}

}  // namespace ih
