// CT geometry on the host: the label volume of a legacy VTK image, inlet and
// outlet detection, and the echo waveform in reference units.
//
//   read_vtk_3D       src/usr_func.f90:187-555
//   convert_waveform  src/usr_func.f90:599-651 (both branches' scales)
//   get_pressure_waveform_echo_WKM  src/usr_func.f90:654-750
//   get_pressure_and_flow_rate_WKM  src/usr_func.f90:762-855
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

void waveform_convert(const double *t_ms, const double *q, int npts, int M1, bool idealized,
                      std::vector<double> &t_out, std::vector<double> &q_out, double ref[3]) {
    // CT pixel width 0.0004 m, upsampling 1, peak inflow 1 in reference units;
    // the idealized aorta overrides the scales with 0.2 m and 2.7 m/s (:637-640)
    const double pixel = 0.0004, pixel_star = (1.0 / M1) * 1.0;
    double vmax = q[0];
    for (int i = 1; i < npts; ++i) vmax = q[i] > vmax ? q[i] : vmax;
    double velocity = vmax / 1.0;
    double length = pixel / pixel_star;
    if (idealized) {
        length = 0.2;
        velocity = 2.7;
    }
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

namespace {
constexpr double kBloodDensity = 1060.;  // src/mod_vars.f90:236
}

// get_pressure_and_flow_rate_WKM: the reference's closed form of a two-element
// Windkessel under a half-sine systolic inflow (T_beat 1 s, systole 0.4 s,
// 80 mmHg at the start), with its constants per branch (1 inflow, 2 outflow,
// 3 the supra-aortic outflow).  The physiological time is time x reference
// time, not wrapped into a beat.  state[branch-1] (p_end_systole_*) is written
// in systole and read in diastole.  p and q come back nondimensional.  The
// expressions keep the reference's operation order; the sine term of the
// systolic pressure is the reference's, which lacks the Q R T_systole factor
// of the ODE's particular solution.
void wk_pressure_flow(int branch, double time, const double ref[3], double state[3], double &p_out, double &q_out) {
    const double pi = 4. * std::atan(1.);
    const double T_systole = 0.4 * 1.;
    const double p_diastole = 80.;
    double C, R, Qp;
    if (branch == 1) {
        C = 1.5496 / 1000000.;
        R = 0.9758 * 1000000.;
        Qp = 0.0004;
    } else if (branch == 2) {
        C = 1.5 / 1000000.;
        R = 1.12 * 1000000.;
        Qp = 0.0004 * 0.85;
    } else {
        C = 0.059 / 1000000.;
        R = 19.68 * 1000000.;
        Qp = 0.0004 * 0.05;
    }
    const double tp = time * ref[2];
    double p, q;
    if (tp <= T_systole) {
        const double den = std::pow(T_systole, 2.) + std::pow(C, 2.) * std::pow(pi, 2.) * std::pow(R, 2.);
        const double homog = p_diastole + Qp * T_systole * R * C * pi * R / den;
        const double forced = T_systole * Qp * R * (C * pi * R * std::cos(pi * tp / T_systole)) -
                              T_systole * std::sin(pi * tp / T_systole);
        p = homog * std::exp(-(1. * tp) / (R * C)) - forced / den;
        q = Qp * std::sin(pi * (tp) / T_systole);
        state[branch - 1] = p;
    } else {
        p = state[branch - 1] * std::exp(-(1. * (tp - T_systole)) / (R * C));
        q = 0.;
    }
    p_out = p / (kBloodDensity * std::pow(ref[0], 2.)) * 133.322;
    q_out = q / (std::pow(ref[1], 3.) / ref[2]);
}

// get_pressure_waveform_echo_WKM: from 80 mmHg, the step p (1 - dt/(C R)) -
// dt/C q_in for each waveform interval whose converted times hold the
// pulse-shifted time (dt the interval in seconds, q_in the interval's start
// velocity times q_max / max velocity), for the inlet and the outlet models;
// nondimensional.
void wk_echo(const std::vector<double> &t_ms, const std::vector<double> &q_raw, const std::vector<double> &t_conv,
             double time, double reference_velocity, double &pre_inlet, double &pre_outlet) {
    const size_t n = t_conv.size();
    const double pulse = t_conv[n - 1];
    const double number = std::floor(time / pulse) + 1.0;
    const double shifted = time - (number - 1.0) * pulse;
    double vel_max = q_raw[0];
    for (size_t i = 1; i < n; ++i) vel_max = q_raw[i] > vel_max ? q_raw[i] : vel_max;
    for (int side = 0; side < 2; ++side) {
        const double R = side == 0 ? 0.78 * 1000000. : 0.898 * 1000000.;
        const double C = side == 0 ? 2.166 / 1000000. : 2.066 / 1000000.;
        const double q_max = side == 0 ? 0.0004 * 1.0 : 0.0004 * 0.85 * 1.0;
        const double area = q_max / vel_max;
        double pre = 80.;
        for (size_t i = 0; i + 1 < n; ++i)
            if (shifted >= t_conv[i] && shifted <= t_conv[i + 1]) {
                const double dt = t_ms[i + 1] / 1000. - t_ms[i] / 1000.;
                const double q_in = q_raw[i] * area;
                pre = pre * (1. - dt / (C * R)) - dt / C * q_in;
            }
        pre = pre / (kBloodDensity * std::pow(reference_velocity, 2.)) * 133.322;
        (side == 0 ? pre_inlet : pre_outlet) = pre;
    }
}

}  // namespace ih
