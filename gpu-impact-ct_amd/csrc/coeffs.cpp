// Host-side grid and finite-difference coefficients for one block.
//
// Coordinates follow get_coords/coordinates (src/usr_geometry.f90:93-157,
// src/mod_geometry.f90:67-200) with the tanh map of coord_tanh
// (src/usr_func.f90:1964-2000).  Stencil widths and positions follow
// diff_coeffs (src/mod_coeffs.f90:200-596) with mapping: weights are found on
// the index grid and scaled by the metric dx/dxi.  The index-grid weights are
// computed here exactly, as rational Lagrange-derivative weights solved in
// integer arithmetic, and rounded once to double; the reference reads the same
// rationals from its tables (src/mod_coeffs.f90:598-772).  cdg follows
// get_stencil (src/mod_coeffs.f90:1048-1258); the multigrid levels follow the
// hierarchy of src/mod_setup.f90:96-143, the level coordinates of
// src/mod_geometry.f90:104-166, get_stencil on every level, interp_coeffs
// (src/mod_coeffs.f90:1683-1746) and restr_coeffs (:1879-2035).
#include "ih_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace ih {

namespace {

using i128 = __int128;

i128 gcd128(i128 a, i128 b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b != 0) {
        i128 t = a % b;
        a = b;
        b = t;
    }
    return a;
}

struct Frac {
    i128 n = 0, d = 1;
    Frac() = default;
    Frac(i128 nn, i128 dd) : n(nn), d(dd) { norm(); }
    void norm() {
        if (d < 0) {
            n = -n;
            d = -d;
        }
        const i128 g = gcd128(n, d);
        if (g > 1) {
            n /= g;
            d /= g;
        }
    }
    Frac operator+(const Frac &o) const { return Frac(n * o.d + o.n * d, d * o.d); }
    Frac operator-(const Frac &o) const { return Frac(n * o.d - o.n * d, d * o.d); }
    Frac operator*(const Frac &o) const { return Frac(n * o.n, d * o.d); }
    Frac operator/(const Frac &o) const { return Frac(n * o.d, d * o.n); }
    bool zero() const { return n == 0; }
    double to_double() const { return (double)(long long)n / (double)(long long)d; }
};

// Weights w_q of the n-point stencil at offsets (twice_x0 + 2q) / 2 that
// reproduce the abl-th derivative at 0 exactly for polynomials of degree < n.
bool exact_weights(int abl, int n, int twice_x0, double *out) {
    Frac A[7][8];
    for (int p = 0; p < n; ++p) {
        for (int q = 0; q < n; ++q) {
            Frac x(twice_x0 + 2 * q, 2), v(1, 1);
            for (int e = 0; e < p; ++e) v = v * x;
            A[p][q] = v;
        }
        i128 fact = 1;
        for (int e = 2; e <= p; ++e) fact *= e;
        A[p][n] = (p == abl) ? Frac(fact, 1) : Frac(0, 1);
    }
    for (int col = 0; col < n; ++col) {
        int piv = -1;
        for (int r = col; r < n; ++r)
            if (!A[r][col].zero()) {
                piv = r;
                break;
            }
        if (piv < 0) return false;
        if (piv != col)
            for (int c = 0; c <= n; ++c) std::swap(A[piv][c], A[col][c]);
        for (int r = 0; r < n; ++r) {
            if (r == col || A[r][col].zero()) continue;
            const Frac f = A[r][col] / A[col][col];
            for (int c = col; c <= n; ++c) A[r][c] = A[r][c] - f * A[col][c];
        }
    }
    for (int q = 0; q < n; ++q) out[q] = (A[q][n] / A[q][q]).to_double();
    return true;
}

void coord_map(double Lmax, double iimax, double fac, double ii, double *xx) {
    if (fac == 0.0) {
        *xx = (ii - 1.0) * Lmax / (iimax - 1.0);
    } else {
        const double cmax = std::tanh(fac);
        const double cmin = -std::tanh(fac);
        const double yy = (ii - 1.0) / (iimax - 1.0) * (cmax - cmin) + cmin;
        *xx = Lmax * (std::atanh(yy) - std::atanh(cmin)) / (fac + fac);
    }
}

// cc(bL:bU, 0:Nmax) column-major; xC indexed (bL:Nmax+bU) -> xC[i - bL]
bool diff_coeffs(int abl, int upwind, const int *ncb, int dim_ncb, const double *xC, int BCL, int BCU,
                 int grid_type, int Nmax, int bL, int bU, std::vector<double> &cc, std::string &err) {
    const int w = bU - bL + 1;
    const int iStart = (grid_type == 0 || grid_type == 3) ? 1 : 0;
    cc.assign((size_t)w * (Nmax + 1), 0.0);
    for (int i = iStart; i <= Nmax; ++i) {
        int n;
        if (BCL > 0 && i <= iStart - 1 + dim_ncb)
            n = ncb[i - iStart];
        else if (BCU > 0 && i >= Nmax + 1 - dim_ncb)
            n = ncb[Nmax - i];
        else
            n = ncb[dim_ncb - 1];
        if (!(n > 0 && n > abl)) continue;
        int right;
        if (grid_type == 0 || grid_type == 1) {
            if (BCL > 0 && i <= iStart - 1 + n / 2)
                right = n - i + iStart - 1;
            else if (BCU > 0 && i >= Nmax + 1 - n / 2)
                right = Nmax - i;
            else
                right = (n - 1) / 2;
        } else if (grid_type == 2) {
            if (BCL > 0 && i <= n / 2)
                right = n - i - iStart;
            else if (BCU > 0 && i >= Nmax - n / 2)
                right = Nmax - i;
            else
                right = n / 2;
        } else {
            if (BCL > 0 && i <= n / 2)
                right = n - i - iStart;
            else if (BCU > 0 && i >= Nmax + 1 - n / 2)
                right = Nmax - i;
            else
                right = n / 2 - 1;
        }
        int left = right - n + 1;
        if (upwind == -1 && !(BCL > 0 && i < iStart + n / 2)) {
            n -= 1;
            left += 1;
        }
        if (upwind == 1 && !(BCU > 0 && i > Nmax - n / 2)) {
            n -= 1;
            right -= 1;
        }
        if (right > bU || left < bL || n > 7) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "stencil at i=%d does not fit (left %d right %d n %d)", i, left, right, n);
            err = buf;
            return false;
        }
        // offsets on the index grid, in halves
        const int twice_x0 = 2 * left + (grid_type == 2 ? -1 : (grid_type == 3 ? 1 : 0));
        if (abl != 0) {
            double cxi[2][7];
            double dxi[2] = {0.0, 0.0};
            for (int kk = 1; kk <= abl; ++kk) {
                if (!exact_weights(kk, n, twice_x0, cxi[kk - 1])) {
                    err = "singular stencil";
                    return false;
                }
                for (int ii = left; ii <= right; ++ii) dxi[kk - 1] = dxi[kk - 1] + cxi[kk - 1][ii - left] * xC[(i + ii) - bL];
            }
            for (int ii = left; ii <= right; ++ii) {
                double v;
                if (abl == 1) {
                    v = cxi[0][ii - left] / dxi[0];
                } else {
                    const double d1sq = dxi[0] * dxi[0];
                    const double d1cu = dxi[0] * dxi[0] * dxi[0];
                    v = cxi[1][ii - left] / d1sq - cxi[0][ii - left] * dxi[1] / d1cu;
                }
                cc[(size_t)(ii - bL) + (size_t)w * i] = v;
            }
        } else {
            double t[7];
            if (!exact_weights(0, n, twice_x0, t)) {
                err = "singular stencil";
                return false;
            }
            for (int ii = left; ii <= right; ++ii) cc[(size_t)(ii - bL) + (size_t)w * i] = t[ii - left];
        }
    }
    return true;
}

}  // namespace

void host_coords(HostGrid &hg) {
    for (int d = 0; d < 3; ++d) {
        const int M = hg.N[d];
        const double L = hg.L[d];
        const double s = hg.stretch[d];
        std::vector<double> &yp = hg.xp[d];
        std::vector<double> &yu = hg.xu[d];
        yp.assign((size_t)M + 7, 0.0);
        yu.assign((size_t)M + 7, 0.0);
        for (int i = 1; i <= M; ++i) coord_map(L, (double)M, s, (double)i, &yp[i + 3]);
        for (int i = 0; i <= M; ++i) coord_map(L, (double)M, s, (double)i + 0.5, &yu[i + 3]);
        if (hg.bc[d][0] == -1) {
            for (int i = -3; i <= 0; ++i) {
                yp[i + 3] = yp[(M - 1 + i) + 3] - L;
                yu[i + 3] = yu[(M - 1 + i) + 3] - L;
            }
            for (int i = M; i <= M + 3; ++i) {
                yp[i + 3] = yp[(1 + i - M) + 3] + L;
                yu[i + 3] = yu[(1 + i - M) + 3] + L;
            }
        }
        if (hg.bc[d][0] > 0) {
            for (int i = -3; i <= 0; ++i) yp[i + 3] = 2.0 * yp[1 + 3] - yp[(2 - i) + 3];
            for (int i = -3; i <= -1; ++i) yu[i + 3] = 2.0 * yp[1 + 3] - yu[(1 - i) + 3];
        }
        if (hg.bc[d][1] > 0) {
            for (int i = M + 1; i <= M + 3; ++i) yp[i + 3] = 2.0 * yp[M + 3] - yp[(2 * M - i) + 3];
            for (int i = M + 1; i <= M + 3; ++i) yu[i + 3] = 2.0 * yp[M + 3] - yu[(2 * M - 1 - i) + 3];
        }
    }
}

bool host_coeffs(const HostGrid &hg, HostCoeffs &hc, std::string &err) {
    static const int ncbc[4] = {4, 5, 5, 7};  // src/mod_dims.f90:64-84
    static const int ncbg[4] = {3, 4, 4, 6};
    static const int ncbd[4] = {4, 4, 6, 6};
    for (int d = 0; d < 3; ++d) {
        const int N = hg.N[d];
        const int BL = hg.bc[d][0], BU = hg.bc[d][1];
        const double *xp = hg.xp[d].data();
        const double *xu = hg.xu[d].data();
        bool ok = true;
        ok = ok && diff_coeffs(1, -1, ncbc, 4, xp, BL, BU, 0, N, -3, 3, hc.NpD[d], err);
        ok = ok && diff_coeffs(1, 1, ncbc, 4, xp, BL, BU, 0, N, -3, 3, hc.NpU[d], err);
        ok = ok && diff_coeffs(1, -1, ncbc, 4, xu, BL, BU, 1, N, -3, 3, hc.NuD[d], err);
        ok = ok && diff_coeffs(1, 1, ncbc, 4, xu, BL, BU, 1, N, -3, 3, hc.NuU[d], err);
        ok = ok && diff_coeffs(2, 0, ncbc, 4, xp, BL, BU, 0, N, -3, 3, hc.p2[d], err);
        ok = ok && diff_coeffs(2, 0, ncbc, 4, xu, BL, BU, 1, N, -3, 3, hc.u2[d], err);
        ok = ok && diff_coeffs(1, 0, ncbg, 4, xp + 1, BL, BU, 2, N, -2, 3, hc.G[d], err);
        ok = ok && diff_coeffs(0, 0, ncbg, 4, xp + 1, BL, BU, 2, N, -2, 3, hc.Ip[d], err);
        ok = ok && diff_coeffs(1, 0, ncbd, 4, xu, BL, BU, 3, N, -3, 2, hc.D[d], err);
        ok = ok && diff_coeffs(0, 0, ncbd, 4, xu, BL, BU, 3, N, -3, 2, hc.Iu[d], err);
        if (!ok) return false;

        // cdg (get_stencil, finest grid)
        std::vector<double> cD(2 * (size_t)(N + 1), 0.0), cG(2 * (size_t)(N + 1), 0.0);
        auto CD = [&](int ii, int i) -> double & { return cD[(size_t)(ii + 1) + 2 * (size_t)i]; };
        auto CG = [&](int ii, int i) -> double & { return cG[(size_t)ii + 2 * (size_t)i]; };
        auto XU = [&](int i) { return xu[i + 3]; };
        auto XP = [&](int i) { return xp[i + 3]; };
        for (int i = 1; i <= N; ++i) {
            CD(-1, i) = -1.0 / (XU(i) - XU(i - 1));
            CD(0, i) = 1.0 / (XU(i) - XU(i - 1));
        }
        for (int i = 1; i <= N - 1; ++i) {
            CG(0, i) = -1.0 / (XP(i + 1) - XP(i));
            CG(1, i) = 1.0 / (XP(i + 1) - XP(i));
        }
        if (BL > 0) {
            CG(0, 0) = CG(1, 0) = 0.0;
            CD(-1, 1) = 0.0;
            CD(0, 1) = 1.0 / (XU(1) - XU(0));
        } else {
            CG(0, 0) = -1.0 / (XP(1) - XP(0));
            CG(1, 0) = 1.0 / (XP(1) - XP(0));
        }
        if (BU > 0) {
            CG(0, N) = CG(1, N) = 0.0;
            CD(0, N) = 0.0;
            CD(-1, N) = -1.0 / (XU(N) - XU(N - 1));
        } else {
            CG(0, N) = -1.0 / (XP(N + 1) - XP(N));
            CG(1, N) = 1.0 / (XP(N + 1) - XP(N));
        }
        std::vector<double> &dg = hc.dg[d];
        dg.assign(3 * (size_t)N, 0.0);
        for (int i = 1; i <= N; ++i) {
            dg[0 + 3 * (size_t)(i - 1)] = CD(-1, i) * CG(0, i - 1);
            dg[1 + 3 * (size_t)(i - 1)] = CD(-1, i) * CG(1, i - 1) + CD(0, i) * CG(0, i);
            dg[2 + 3 * (size_t)(i - 1)] = CD(0, i) * CG(1, i);
        }
        if (BL > 0)
            for (int q = 0; q < 3; ++q) dg[q] = 2.0 * dg[q];
        if (BU > 0)
            for (int q = 0; q < 3; ++q) dg[q + 3 * (size_t)(N - 1)] = 2.0 * dg[q + 3 * (size_t)(N - 1)];
    }
    return true;
}

namespace {

// cdg of one direction from the level's coordinates xp (0..N+1), xu (0..N) at
// [i+3]; at walls the finest grid's first and last u-cell widths (yu0) enter
// on every level, as get_stencil writes them (src/mod_coeffs.f90:1093, 1101)
void level_dg(int N, int bl, int bu, const std::vector<double> &xp, const std::vector<double> &xu,
              const std::vector<double> &yu0, int M, std::vector<double> &dg) {
    std::vector<double> cDm(N + 1, 0.0), cD0(N + 1, 0.0), cG0(N + 1, 0.0), cG1(N + 1, 0.0);
    for (int i = 1; i <= N; ++i) {
        const double w = xu[i + 3] - xu[i + 2];
        cDm[i] = -1.0 / w;
        cD0[i] = 1.0 / w;
    }
    for (int i = 1; i <= N - 1; ++i) {
        const double w = xp[i + 4] - xp[i + 3];
        cG0[i] = -1.0 / w;
        cG1[i] = 1.0 / w;
    }
    if (bl > 0) {
        cG0[0] = cG1[0] = 0.0;
        cDm[1] = 0.0;
        cD0[1] = 1.0 / (yu0[1 + 3] - yu0[0 + 3]);
    } else {
        cG0[0] = -1.0 / (xp[1 + 3] - xp[0 + 3]);
        cG1[0] = 1.0 / (xp[1 + 3] - xp[0 + 3]);
    }
    if (bu > 0) {
        cG0[N] = cG1[N] = 0.0;
        cD0[N] = 0.0;
        cDm[N] = -1.0 / (yu0[M + 3] - yu0[M - 1 + 3]);
    } else {
        cG0[N] = -1.0 / (xp[N + 4] - xp[N + 3]);
        cG1[N] = 1.0 / (xp[N + 4] - xp[N + 3]);
    }
    dg.assign(3 * (size_t)N, 0.0);
    for (int i = 1; i <= N; ++i) {
        double *r = &dg[3 * (size_t)(i - 1)];
        r[0] = cDm[i] * cG0[i - 1];
        r[1] = cDm[i] * cG1[i - 1] + cD0[i] * cG0[i];
        r[2] = cD0[i] * cG1[i];
    }
    for (int q = 0; q < 3; ++q) {
        if (bl > 0) dg[q] = 2.0 * dg[q];
        if (bu > 0) dg[3 * (size_t)(N - 1) + q] = 2.0 * dg[3 * (size_t)(N - 1) + q];
    }
}

}  // namespace

void host_mg_levels(const HostGrid &hg, int n_grids_limit, std::vector<HostMgLevel> &out) {
    // sizes: a direction halves while N - 1 is even and N >= 5; the number of
    // grids is the last level on which any direction halved
    int NN[MG_MAX_LEVELS + 1][3];
    int G = 1;
    for (int d = 0; d < 3; ++d) NN[1][d] = hg.N[d];
    for (int g = 2; g <= MG_MAX_LEVELS; ++g)
        for (int d = 0; d < 3; ++d) {
            const int p = NN[g - 1][d];
            const bool halve = (p - 1) % 2 == 0 && p >= 5;
            NN[g][d] = halve ? (p - 1) / 2 + 1 : p;
            if (halve) G = g;
        }
    G = std::min(G, std::max(1, n_grids_limit));
    out.assign(G, HostMgLevel{});
    for (int g = 1; g <= G; ++g) {
        HostMgLevel &lv = out[g - 1];
        for (int d = 0; d < 3; ++d) {
            const int N = NN[g][d], M = hg.N[d];
            const int bl = hg.bc[d][0], bu = hg.bc[d][1];
            const double Lx = hg.L[d];
            const std::vector<double> &yp0 = hg.xp[d];
            lv.N[d] = N;
            lv.dn[d] = g < G ? (N - 1) / (NN[g + 1][d] - 1) : 1;
            std::vector<double> xp, xu;
            if (g == 1) {
                xp = hg.xp[d];
                xu = hg.xu[d];
            } else {
                const int di = (M - 1) / (N - 1);
                xp.assign((size_t)N + 7, 0.0);
                xu.assign((size_t)N + 7, 0.0);
                for (int i = 1; i <= N; ++i) xp[i + 3] = yp0[(1 + di * (i - 1)) + 3];
                for (int i = 1; i <= N - 1; ++i) xu[i + 3] = yp0[(1 + di * i - di / 2) + 3];
                if (bl == -1) {
                    std::vector<double> p0(xp), u0(xu);  // array assignments read the old values
                    for (int b = -3; b <= 0; ++b) {
                        xp[b + 3] = p0[(N - 1 + b) + 3] - Lx;
                        xu[b + 3] = u0[(N - 1 + b) + 3] - Lx;
                    }
                    std::vector<double> p1(xp), u1(xu);
                    for (int t = 0; t <= 3; ++t) {
                        xp[(N + t) + 3] = p1[(1 + t) + 3] + Lx;
                        xu[(N + t) + 3] = u1[(1 + t) + 3] + Lx;
                    }
                }
                if (bl > 0) {
                    for (int b = -3; b <= 0; ++b) xp[b + 3] = 2.0 * xp[1 + 3] - xp[(2 - b) + 3];
                    for (int b = -3; b <= 0; ++b) xu[b + 3] = 2.0 * xp[1 + 3] - xu[(1 - b) + 3];
                }
                if (bu > 0) {
                    for (int t = 1; t <= 3; ++t) xp[(N + t) + 3] = 2.0 * xp[N + 3] - xp[(N - t) + 3];
                    for (int t = 0; t <= 3; ++t) xu[(N + t) + 3] = 2.0 * xp[N + 3] - xu[(N - 1 - t) + 3];
                }
            }
            level_dg(N, bl, bu, xp, xu, hg.xu[d], M, lv.dg[d]);
            lv.cR[d].assign(3 * (size_t)N, 0.0);
            if (g >= 2) {
                for (int ii = 1; ii <= N; ++ii) {
                    double *r = &lv.cR[d][3 * (size_t)(ii - 1)];
                    const bool ident = (bl > 0 && ii <= 2) || (bu > 0 && ii >= N - 1);
                    r[0] = ident ? 0.0 : 1.0 / 4.0;
                    r[1] = ident ? 1.0 : 2.0 / 4.0;
                    r[2] = ident ? 0.0 : 1.0 / 4.0;
                }
            }
            lv.cI[d].assign(2 * (size_t)N, 0.0);
            if (g < G) {
                const int di = (M - 1) / (N - 1);
                for (int i = 2; i <= N - 1; i += 2) {
                    const int i0 = 1 + (i - 1) * di;
                    const double d10 = yp0[i0 + 3] - yp0[(i0 - di) + 3];
                    const double d12 = yp0[(i0 + di) + 3] - yp0[(i0 - di) + 3];
                    lv.cI[d][2 * (size_t)(i - 1)] = 1.0 - d10 / d12;
                    lv.cI[d][2 * (size_t)(i - 1) + 1] = d10 / d12;
                }
            }
        }
    }
}

}  // namespace ih
