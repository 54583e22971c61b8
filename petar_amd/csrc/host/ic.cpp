// ic.cpp -- initial conditions for the soft-force workloads.
//
// pth_ic_plummer_equal follows the recipe of makePlummerModel
// (src/particle_distribution_generator.hpp:173-250) step by step, with the
// same MT19937 stream, so the same seed gives the same particles.  The
// other generators (Kroupa IMF, King W0, primordial binaries) have no
// counterpart in the reference, which reads such models from external tools
// (sample/star_cluster.sh:3); they follow the published definitions cited
// at each function.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../../include/petar_host.h"
#include "mt19937.hpp"

namespace pth {
thread_local std::string g_host_err;
int host_fail(int code, const std::string &msg)
{
    g_host_err = msg;
    return code;
}
}  // namespace pth

using pth::host_fail;
using pth::MT19937;

extern "C" const char *pth_host_last_error(void) { return pth::g_host_err.c_str(); }

struct pth_mt {
    MT19937 g;
};
extern "C" pth_mt *pth_mt_new(unsigned int seed)
{
    pth_mt *m = new pth_mt;
    m->g.seed_(seed);
    return m;
}
extern "C" void pth_mt_free(pth_mt *mt) { delete mt; }
extern "C" unsigned int pth_mt_int32(pth_mt *mt) { return mt->g.next_u32(); }
extern "C" double pth_mt_res53(pth_mt *mt) { return mt->g.res53(); }
extern "C" double pth_mt_real2(pth_mt *mt) { return mt->g.real2(); }

namespace {

const double PI = 3.14159265358979323846;

// One Plummer particle in model units (scale radius 1), the draw sequence of
// src/particle_distribution_generator.hpp:193-219.
void plummer_draw(MT19937 &mt, double r_cutoff, double *x, double *v)
{
    double r_tmp = 9999.9;
    while (r_tmp > r_cutoff) {
        double m_tmp = mt.res53();
        r_tmp = 1.0 / std::sqrt(std::pow(m_tmp, (-2.0 / 3.0)) - 1.0);
    }
    double phi = 2.0 * PI * mt.res53();
    double cth = 2.0 * (mt.real2() - 0.5);
    double sth = std::sqrt(1.0 - cth * cth);
    x[0] = r_tmp * sth * std::cos(phi);
    x[1] = r_tmp * sth * std::sin(phi);
    x[2] = r_tmp * cth;
    while (true) {
        const double v_max = 0.1;
        const double v_try = mt.res53();
        const double v_crit = v_max * mt.res53();
        if (v_crit < v_try * v_try * std::pow((1.0 - v_try * v_try), 3.5)) {
            const double ve = std::sqrt(2.0) * std::pow((r_tmp * r_tmp + 1.0), -0.25);
            phi = 2.0 * PI * mt.res53();
            cth = 2.0 * (mt.res53() - 0.5);
            sth = std::sqrt(1.0 - cth * cth);
            v[0] = ve * v_try * sth * std::cos(phi);
            v[1] = ve * v_try * sth * std::sin(phi);
            v[2] = ve * v_try * cth;
            break;
        }
    }
}

void remove_cm(long n, const double *mass, double *a)
{
    double c[3] = {0, 0, 0}, m = 0;
    for (long i = 0; i < n; i++) {
        m += mass[i];
        for (int d = 0; d < 3; d++) c[d] += mass[i] * a[3 * i + d];
    }
    for (int d = 0; d < 3; d++) c[d] /= m;
    for (long i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) a[3 * i + d] -= c[d];
}

// King (1966) model: W'' + (2/r) W' = -9 rho(W)/rho(W0) in core-radius
// units, rho(W) ~ e^W erf(sqrt W) - sqrt(4W/pi) (1 + 2W/3).
double king_rho(double w)
{
    if (w <= 0) return 0.0;
    return std::exp(w) * std::erf(std::sqrt(w)) - std::sqrt(4.0 * w / PI) * (1.0 + 2.0 * w / 3.0);
}

struct KingTable {
    std::vector<double> r, w, m;  // radius, potential, enclosed (relative) mass
    double u_model = 0;           // potential energy, G = 9/(4 pi), rho0 = 1
    double m_model = 0;
};

KingTable king_solve(double w0)
{
    KingTable t;
    const double rho0 = king_rho(w0);
    auto f = [&](double r, double w, double dw, double &ddw) {
        ddw = -9.0 * king_rho(w) / rho0 - 2.0 / r * dw;
    };
    double h = 1e-4;
    double r = 1e-5;
    double w = w0 - 1.5 * r * r, dw = -3.0 * r;
    double mass = 4.0 * PI / 3.0 * r * r * r;
    const double G = 9.0 / (4.0 * PI);
    double u = 0.0;
    t.r.push_back(0.0);
    t.w.push_back(w0);
    t.m.push_back(0.0);
    while (w > 0.0) {
        // RK4 on (w, dw); mass and energy by the trapezoid rule
        double k1w = dw, k1d;
        f(r, w, dw, k1d);
        double k2w = dw + 0.5 * h * k1d, k2d;
        f(r + 0.5 * h, w + 0.5 * h * k1w, k2w, k2d);
        double k3w = dw + 0.5 * h * k2d, k3d;
        f(r + 0.5 * h, w + 0.5 * h * k2w, k3w, k3d);
        double k4w = dw + h * k3d, k4d;
        f(r + h, w + h * k3w, k4w, k4d);
        double wn = w + h / 6.0 * (k1w + 2 * k2w + 2 * k3w + k4w);
        double dwn = dw + h / 6.0 * (k1d + 2 * k2d + 2 * k3d + k4d);
        double rn = r + h;
        double rho_a = king_rho(w) / rho0, rho_b = king_rho(std::max(wn, 0.0)) / rho0;
        double dm = 0.5 * h * 4.0 * PI * (r * r * rho_a + rn * rn * rho_b);
        double mn = mass + dm;
        u -= G * 0.5 * (mass / r + mn / rn) * dm;
        if (wn <= 0.0) {
            // linear interpolation to the tidal radius
            double frac = w / (w - wn);
            rn = r + frac * h;
            wn = 0.0;
        }
        r = rn;
        w = wn;
        dw = dwn;
        mass = mn;
        t.r.push_back(r);
        t.w.push_back(w);
        t.m.push_back(mass);
        if (r > 1e4) break;
        h = 1e-3 * std::max(r, 0.1);
    }
    t.m_model = mass;
    t.u_model = u;
    return t;
}

// Random rotation from a uniform quaternion (Shoemake 1992).
void random_rotation(MT19937 &mt, double R[9])
{
    double u1 = mt.res53(), u2 = mt.res53(), u3 = mt.res53();
    double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    double q0 = a * std::sin(2 * PI * u2), q1 = a * std::cos(2 * PI * u2);
    double q2 = b * std::sin(2 * PI * u3), q3 = b * std::cos(2 * PI * u3);
    // quaternion (w=q3, x=q0, y=q1, z=q2)
    double w = q3, x = q0, y = q1, z = q2;
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w);     R[2] = 2 * (x * z + y * w);
    R[3] = 2 * (x * y + z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
    R[6] = 2 * (x * z - y * w);     R[7] = 2 * (y * z + x * w);     R[8] = 1 - 2 * (x * x + y * y);
}

}  // namespace

extern "C" int pth_ic_plummer_equal(long n, double mass_glb, double eng, unsigned int seed,
                                    double *mass, double *pos, double *vel)
{
    if (n <= 0) return host_fail(-1, "pth_ic_plummer_equal: n must be positive");
    if (!(eng < 0.0)) return host_fail(-1, "pth_ic_plummer_equal: energy must be negative");
    // src/particle_distribution_generator.hpp:183 (cutoff 22.8 N-body units)
    const double r_cutoff = 22.8 / (-3.0 * PI * mass_glb * mass_glb / (64.0 * -0.25));
    MT19937 mt(seed);
    std::vector<double> vtmp(vel ? 0 : 3);
    for (long i = 0; i < n; i++) {
        mass[i] = mass_glb / n;
        double *v = vel ? vel + 3 * i : vtmp.data();
        plummer_draw(mt, r_cutoff, pos + 3 * i, v);
    }
    remove_cm(n, mass, pos);
    if (vel) remove_cm(n, mass, vel);
    const double r_scale = -3.0 * PI * mass_glb * mass_glb / (64.0 * eng);
    const double coef = 1.0 / std::sqrt(r_scale);
    for (long i = 0; i < 3 * n; i++) pos[i] *= r_scale;
    if (vel)
        for (long i = 0; i < 3 * n; i++) vel[i] *= coef;
    return 0;
}

extern "C" int pth_ic_kroupa_masses(long n, double m_lo, double m_hi, double m_tot,
                                    unsigned int seed, double *mass)
{
    // Kroupa (2001, MNRAS 322, 231): dN/dm ~ m^-1.3 (m < 0.5), m^-2.3 (m >= 0.5),
    // continuous at 0.5; sampled by inverting the piecewise CDF.
    if (n <= 0 || !(m_lo > 0) || !(m_hi > m_lo))
        return host_fail(-1, "pth_ic_kroupa_masses: bad arguments");
    const double mb = 0.5;
    const double a1 = 1.3, a2 = 2.3;
    auto seg_int = [](double a, double lo, double hi) {  // int m^-a dm
        return (std::pow(hi, 1 - a) - std::pow(lo, 1 - a)) / (1 - a);
    };
    double lo1 = m_lo, hi1 = std::min(m_hi, mb);
    double lo2 = std::max(m_lo, mb), hi2 = m_hi;
    double c2 = std::pow(mb, a2 - a1);  // continuity: m^-a1 = c2 m^-a2 at mb
    double w1 = (hi1 > lo1) ? seg_int(a1, lo1, hi1) : 0.0;
    double w2 = (hi2 > lo2) ? c2 * seg_int(a2, lo2, hi2) : 0.0;
    MT19937 mt(seed);
    double sum = 0;
    for (long i = 0; i < n; i++) {
        double u = mt.res53() * (w1 + w2);
        double m;
        if (u < w1) {
            double t = std::pow(lo1, 1 - a1) + u * (1 - a1);
            m = std::pow(t, 1.0 / (1 - a1));
        } else {
            double v = (u - w1) / c2;
            double t = std::pow(lo2, 1 - a2) + v * (1 - a2);
            m = std::pow(t, 1.0 / (1 - a2));
        }
        m = std::min(std::max(m, m_lo), m_hi);
        mass[i] = m;
        sum += m;
    }
    if (m_tot > 0)
        for (long i = 0; i < n; i++) mass[i] *= m_tot / sum;
    return 0;
}

extern "C" int pth_ic_cluster(long n, int model, double w0, int kroupa, double f_bin,
                              double a_lo, double a_hi, int segregate, unsigned int seed,
                              double *mass, double *pos, double *vel)
{
    if (n <= 1) return host_fail(-1, "pth_ic_cluster: n must be > 1");
    if (model != 0 && model != 1) return host_fail(-1, "pth_ic_cluster: model must be 0 or 1");
    if (model == 1 && !(w0 > 0.0 && w0 < 16.0))
        return host_fail(-1, "pth_ic_cluster: King W0 must be in (0,16)");
    if (!(f_bin >= 0.0 && f_bin <= 1.0)) return host_fail(-1, "pth_ic_cluster: f_bin in [0,1]");
    if (f_bin > 0 && !(a_lo > 0 && a_hi >= a_lo))
        return host_fail(-1, "pth_ic_cluster: bad semi-major axis range");

    const long n_pair = (long)std::floor(f_bin * (double)n / 2.0);
    const long n_sys = n - n_pair;

    // star masses (solar units), later normalised to total 1
    std::vector<double> mstar(n);
    if (kroupa) {
        int rc = pth_ic_kroupa_masses(n, 0.08, 150.0, 1.0, seed ^ 0x9e3779b9u, mstar.data());
        if (rc) return rc;
    } else {
        for (long i = 0; i < n; i++) mstar[i] = 1.0 / n;
    }
    // system masses: the first n_pair systems are binaries of consecutive stars
    std::vector<double> msys(n_sys);
    for (long s = 0; s < n_sys; s++)
        msys[s] = (s < n_pair) ? mstar[2 * s] + mstar[2 * s + 1] : mstar[n_pair + s];

    // system positions/velocities in model units
    std::vector<double> xs(3 * n_sys), vs(3 * n_sys);
    MT19937 mt(seed);
    double len_scale = 1.0, vel_scale = 1.0;
    if (model == 0) {
        const double r_cutoff = 22.8 / (3.0 * PI / 16.0);
        for (long s = 0; s < n_sys; s++) plummer_draw(mt, r_cutoff, &xs[3 * s], &vs[3 * s]);
        len_scale = 3.0 * PI / 16.0;  // E = -1/4, M = 1
        vel_scale = 1.0 / std::sqrt(len_scale);
    } else {
        KingTable kt = king_solve(w0);
        const double mtot = kt.m.back();
        const double G = 9.0 / (4.0 * PI);
        for (long s = 0; s < n_sys; s++) {
            double u = mt.res53() * mtot;
            long k = std::upper_bound(kt.m.begin(), kt.m.end(), u) - kt.m.begin();
            k = std::min(std::max(k, 1L), (long)kt.m.size() - 1);
            double f = (u - kt.m[k - 1]) / std::max(kt.m[k] - kt.m[k - 1], 1e-300);
            double r = kt.r[k - 1] + f * (kt.r[k] - kt.r[k - 1]);
            double wloc = kt.w[k - 1] + f * (kt.w[k] - kt.w[k - 1]);
            double phi = 2.0 * PI * mt.res53();
            double cth = 2.0 * (mt.res53() - 0.5);
            double sth = std::sqrt(1.0 - cth * cth);
            xs[3 * s + 0] = r * sth * std::cos(phi);
            xs[3 * s + 1] = r * sth * std::sin(phi);
            xs[3 * s + 2] = r * cth;
            // speed from the lowered Maxwellian, p(v) ~ v^2 (exp(W - v^2/2) - 1), sigma = 1
            double vmax = std::sqrt(2.0 * std::max(wloc, 0.0));
            double v = 0.0;
            if (vmax > 0) {
                double pmax = 0;
                for (int t = 1; t <= 64; t++) {
                    double vv = vmax * t / 64.0;
                    pmax = std::max(pmax, vv * vv * (std::exp(wloc - 0.5 * vv * vv) - 1.0));
                }
                pmax *= 1.1;
                for (int it = 0; it < 100000; it++) {
                    double vt = vmax * mt.res53();
                    double pt = vt * vt * (std::exp(wloc - 0.5 * vt * vt) - 1.0);
                    if (mt.res53() * pmax < pt) { v = vt; break; }
                }
            }
            phi = 2.0 * PI * mt.res53();
            cth = 2.0 * (mt.res53() - 0.5);
            sth = std::sqrt(1.0 - cth * cth);
            vs[3 * s + 0] = v * sth * std::cos(phi);
            vs[3 * s + 1] = v * sth * std::sin(phi);
            vs[3 * s + 2] = v * cth;
        }
        // N-body units: G = 1, M = 1, virial U = -1/2
        double u_nb = kt.u_model / G / (kt.m_model * kt.m_model);
        len_scale = -2.0 * u_nb;
        vel_scale = 0.0;  // set below from the kinetic energy
    }

    // mass segregation: heaviest systems at the smallest radii
    std::vector<long> sys_of(n_sys);
    std::iota(sys_of.begin(), sys_of.end(), 0L);
    if (segregate) {
        std::vector<long> by_r(n_sys), by_m(n_sys);
        std::iota(by_r.begin(), by_r.end(), 0L);
        std::iota(by_m.begin(), by_m.end(), 0L);
        auto r2 = [&](long s) {
            return xs[3 * s] * xs[3 * s] + xs[3 * s + 1] * xs[3 * s + 1] + xs[3 * s + 2] * xs[3 * s + 2];
        };
        std::stable_sort(by_r.begin(), by_r.end(), [&](long a, long b) { return r2(a) < r2(b); });
        std::stable_sort(by_m.begin(), by_m.end(), [&](long a, long b) { return msys[a] > msys[b]; });
        for (long k = 0; k < n_sys; k++) sys_of[by_m[k]] = by_r[k];
    }

    // place systems (N-body units), total mass 1
    double msum = 0;
    for (long i = 0; i < n; i++) msum += mstar[i];
    std::vector<double> xcm(3 * n_sys), vcm(3 * n_sys), mcm(n_sys);
    for (long s = 0; s < n_sys; s++) {
        long src = sys_of[s];
        mcm[s] = msys[s] / msum;
        for (int d = 0; d < 3; d++) {
            xcm[3 * s + d] = xs[3 * src + d] * len_scale;
            vcm[3 * s + d] = vs[3 * src + d];
        }
    }
    remove_cm(n_sys, mcm.data(), xcm.data());
    remove_cm(n_sys, mcm.data(), vcm.data());
    if (model == 1) {
        // scale velocities to virial equilibrium, T = 1/4
        double t = 0;
        for (long s = 0; s < n_sys; s++)
            t += 0.5 * mcm[s] * (vcm[3 * s] * vcm[3 * s] + vcm[3 * s + 1] * vcm[3 * s + 1] +
                                 vcm[3 * s + 2] * vcm[3 * s + 2]);
        vel_scale = (t > 0) ? std::sqrt(0.25 / t) : 0.0;
    }
    for (long s = 0; s < n_sys; s++)
        for (int d = 0; d < 3; d++) vcm[3 * s + d] *= vel_scale;

    // expand binaries (Kepler orbits, G = 1)
    MT19937 mb(seed + 0x51ed270bu);
    long p = 0;
    for (long s = 0; s < n_sys; s++) {
        if (s < n_pair) {
            double m1 = mstar[2 * s] / msum, m2 = mstar[2 * s + 1] / msum, mt12 = m1 + m2;
            double la = std::log(a_lo) + (std::log(a_hi) - std::log(a_lo)) * mb.res53();
            double a = std::exp(la);
            double e = std::sqrt(mb.res53());  // thermal f(e) = 2e
            if (e > 0.999) e = 0.999;
            double mean_an = 2.0 * PI * mb.res53();
            double E = (e > 0.8) ? PI : mean_an;
            for (int it = 0; it < 60; it++) {
                double dE = (E - e * std::sin(E) - mean_an) / (1.0 - e * std::cos(E));
                E -= dE;
                if (std::fabs(dE) < 1e-15) break;
            }
            double R[9];
            random_rotation(mb, R);
            double xr[3] = {a * (std::cos(E) - e), a * std::sqrt(1 - e * e) * std::sin(E), 0.0};
            double nmo = std::sqrt(mt12 / (a * a * a));
            double edot = nmo / (1.0 - e * std::cos(E));
            double vr[3] = {-a * std::sin(E) * edot, a * std::sqrt(1 - e * e) * std::cos(E) * edot, 0.0};
            double xw[3], vw[3];
            for (int d = 0; d < 3; d++) {
                xw[d] = R[3 * d] * xr[0] + R[3 * d + 1] * xr[1] + R[3 * d + 2] * xr[2];
                vw[d] = R[3 * d] * vr[0] + R[3 * d + 1] * vr[1] + R[3 * d + 2] * vr[2];
            }
            mass[p] = m1;
            mass[p + 1] = m2;
            for (int d = 0; d < 3; d++) {
                pos[3 * p + d] = xcm[3 * s + d] + m2 / mt12 * xw[d];
                pos[3 * (p + 1) + d] = xcm[3 * s + d] - m1 / mt12 * xw[d];
                if (vel) {
                    vel[3 * p + d] = vcm[3 * s + d] + m2 / mt12 * vw[d];
                    vel[3 * (p + 1) + d] = vcm[3 * s + d] - m1 / mt12 * vw[d];
                }
            }
            p += 2;
        } else {
            mass[p] = mcm[s];
            for (int d = 0; d < 3; d++) {
                pos[3 * p + d] = xcm[3 * s + d];
                if (vel) vel[3 * p + d] = vcm[3 * s + d];
            }
            p += 1;
        }
    }
    return 0;
}
