// Consumers of a wavelength bar's T-matrix on the host (include/scadyn_amd.h):
// orientation-averaged cross sections (tr_T, src/mie.f90:728-759) and the
// radiation force and torque along the incident direction of an x-polarised
// plane wave travelling along +z, the particle in its reference orientation
// (forcetorque, src/forces.f90:81-189, with F_z / T_z, :215-257, and the
// plane-wave coefficients of planewave, src/mie.f90:204-228).
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "internal.hpp"

namespace scadyn {

typedef std::complex<double> zc;

namespace {

constexpr double PI_C = 3.141592653589793;
constexpr double EPS0 = 8.854187817e-12;  // src/common.f90:21, :34
constexpr double MU0 = PI_C * 4e-7;
constexpr double C0 = 2.99792458e8;

inline int nm_of(int nmax) { return (nmax + 1) * (nmax + 1) - 1; }

// column-major nm x nm complex block (interleaved doubles) times a vector
std::vector<zc> block_mul(const double *T, int nm, const std::vector<zc> &v) {
    std::vector<zc> out(nm, zc(0.0, 0.0));
    for (int j = 0; j < nm; ++j) {
        const zc vj = v[j];
        if (vj == zc(0.0, 0.0)) continue;
        for (int i = 0; i < nm; ++i) out[i] += zc(T[2 * ((size_t)j * nm + i)], T[2 * ((size_t)j * nm + i) + 1]) * vj;
    }
    return out;
}

// the force sum of one field (a, b incident, p, q outgoing), index n(n+1)+m-1
double force_z(int nmax, const std::vector<zc> &a, const std::vector<zc> &bb, const std::vector<zc> &p,
               const std::vector<zc> &qq) {
    const int nm = nm_of(nmax);
    const zc I(0.0, 1.0);
    double f = 0.0;
    int ind = 0;
    for (int n = 1; n <= nmax; ++n)
        for (int m = -n; m <= n; ++m, ++ind) {
            const zc b = bb[ind] * I, q = qq[ind] * I;
            double g = (2.0 * m / (n * (n + 1.0))) * std::imag(std::conj(a[ind]) * b - std::conj(p[ind]) * q);
            const int up = ind + 2 * n + 2;  // (n + 1, m)
            if (up < nm) {
                const zc bu = bb[up] * I, qu = qq[up] * I;
                g -= (2.0 / (n + 1.0)) *
                     std::sqrt((n * (n + 2.0) * (n - m + 1.0) * (n + m + 1.0)) / ((2.0 * n + 1.0) * (2.0 * n + 3.0))) *
                     std::imag(a[ind] * std::conj(a[up]) + b * std::conj(bu) - p[ind] * std::conj(p[up]) -
                               q * std::conj(qu));
            }
            f += g;
        }
    return f;
}

double torque_z(int nmax, const std::vector<zc> &a, const std::vector<zc> &b, const std::vector<zc> &p,
                const std::vector<zc> &q) {
    double t = 0.0;
    int ind = 0;
    for (int n = 1; n <= nmax; ++n)
        for (int m = -n; m <= n; ++m, ++ind)
            t += m * (std::norm(a[ind]) + std::norm(b[ind]) - std::norm(p[ind]) - std::norm(q[ind]));
    return t;
}

}  // namespace
}  // namespace scadyn

extern "C" {

int scadyn_cross_sections(int32_t nmax, const double *Taa, const double *Tab, const double *Tba, const double *Tbb,
                          double k, double crs[4]) {
    using namespace scadyn;
    if (nmax < 1 || !Taa || !Tab || !Tba || !Tbb || !crs || !(k > 0.0)) {
        set_error(SCADYN_ERR_ARG, "bad argument");
        return SCADYN_ERR_ARG;
    }
    const int nm = nm_of(nmax);
    double T1 = 0.0, T2 = 0.0;
    const double *blk[4] = {Taa, Tbb, Tab, Tba};
    for (int q = 0; q < 4; ++q)
        for (int mu = 0; mu < nm; ++mu) {
            T1 += blk[q][2 * ((size_t)mu * nm + mu)];
            for (int nu = 0; nu < nm; ++nu) {
                const double *v = blk[q] + 2 * ((size_t)nu * nm + mu);
                T2 += v[0] * v[0] + v[1] * v[1];
            }
        }
    crs[0] = -2.0 * PI_C * T1 / (k * k);
    crs[1] = 2.0 * PI_C * T2 / (k * k);
    crs[2] = crs[0] - crs[1];
    crs[3] = T1 + T2;
    return SCADYN_OK;
}

int scadyn_force_torque_z(int32_t nmax, const double *Taa, const double *Tab, const double *Tba, const double *Tbb,
                          double k, double E, double a, double out[4]) {
    using namespace scadyn;
    if (nmax < 1 || !Taa || !Tab || !Tba || !Tbb || !out || !(k > 0.0) || !(a > 0.0)) {
        set_error(SCADYN_ERR_ARG, "bad argument");
        return SCADYN_ERR_ARG;
    }
    const int nm = nm_of(nmax);
    // planewave (x-polarised, +z, E0 = 1), scaled as forcetorque does
    const double scale = E / std::sqrt(2.0 * std::sqrt(MU0 / EPS0) * k * k) / 2.0;
    std::vector<zc> av(nm, zc(0.0, 0.0)), bv(nm, zc(0.0, 0.0));
    const zc I(0.0, 1.0);
    for (int n = 1; n <= nmax; ++n)
        for (int m = -1; m <= 1; m += 2) {
            const int ind = n * (n + 1) + m - 1;
            const double C = std::sqrt(PI_C * (2.0 * n + 1.0));
            av[ind] = -(m > 0 ? 1.0 : -1.0) * C * std::pow(I, n - 1) * scale;
            bv[ind] = C * std::pow(I, n + 1) * scale;
        }
    // p = 2 (Taa a + Tab b) + a, q = 2 (Tbb b + Tba a) + b
    std::vector<zc> p = block_mul(Taa, nm, av), q = block_mul(Tbb, nm, bv);
    const std::vector<zc> p2 = block_mul(Tab, nm, bv), q2 = block_mul(Tba, nm, av);
    for (int i = 0; i < nm; ++i) {
        p[i] = 2.0 * (p[i] + p2[i]) + av[i];
        q[i] = 2.0 * (q[i] + q2[i]) + bv[i];
    }
    const double fz = force_z(nmax, av, bv, p, q), tz = torque_z(nmax, av, bv, p, q);
    out[0] = fz / C0;                                                       // force (N)
    out[1] = tz / (C0 * k);                                                 // torque (N m)
    out[2] = out[0] / (EPS0 * E * E / 2.0) / PI_C / (a * a);                // Q_f (src/forces.f90:59-63)
    out[3] = out[1] / (0.5 * PI_C / k) / (EPS0 * E * E) / (a * a);          // Q_t
    return SCADYN_OK;
}

}  // extern "C"
