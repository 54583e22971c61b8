/*
 * oracle.c -- FP64 CPU restatement of PeTar's soft-force path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function cites the
 * reference lines it restates.  Parity pins (what each function is checked
 * against, in tests/test_oracle.py):
 *   orc_epj_walks            reference phantomquad kernel built from source
 *                            (oracle/_ref), two-body closed forms, self pair,
 *                            r < r_out linear regime, Newton's third law.
 *   orc_spj_walks            reference phantomquad traceless-form kernel
 *                            (oracle/_ref), multipole convergence against
 *                            brute force over member particles (error falls
 *                            as (s/R)^3), monopole limit.
 *   orc_soft_force_walks     sum of the two above (checked as such).
 *   orc_nnb_ambiguous_walks  brute-force band count on tiny inputs.
 *   orc_direct_*             theta=0 tree lists (every j an EPJ) and
 *                            closed-form two-body cases.
 *   orc_moments              parallel-axis theorem and point-mass cases.
 *   orc_changeover_*         src/changeover.hpp cannot be compiled here (it
 *                            includes SDAR's Common/Float.h, absent from the
 *                            image); pinned by the closed forms W(x<=0)=1,
 *                            W(x>=1)=0, continuity at both ends, the
 *                            identity W_acc = W_pot - r dW_pot/dr, and
 *                            W_acc1 = dW_acc/dt by finite differences.
 *   orc_changeover_correct_one  pinned by the pairwise identity
 *                            linear-cutoff + correction = (1-W) Newton.
 *   orc_soft_force_walks_co  two-body closed forms inside r_in, between r_in
 *                            and r_out and beyond the neighbour radius, the
 *                            larger-r_out rule, the self term, and pot_soft
 *                            as the potential of acc (finite differences).
 *                            Which pairs FDPS's neighbour search returns at
 *                            the search radius is parity unpinned (FDPS is
 *                            not available).
 *   orc_ngb_walks            all-pairs FP64 brute force (a k-d tree query
 *                            widened to the largest r_search, filtered with
 *                            the same FP64 expression) on clusters with
 *                            binaries: equal lists prove both the
 *                            restatement and that the symmetric-search
 *                            lists hold every neighbour.
 */
#include "oracle.h"
#include <math.h>
#include <stddef.h>

static inline double dmax(double a, double b) { return a > b ? a : b; }

static int nnb_pred(int mode, double r2_eps, double r2_raw, double rs2)
{
    /* src/force_gpu_cuda.cu:62-64 (mode 0),
       src/phantomquad_for_p3t_x86.hpp:516-525 (mode 1, blend on rs2 < r2),
       src/soft_force.hpp:57-64 (mode 2). */
    if (mode == 1) return r2_eps <= rs2;
    if (mode == 2) return r2_raw < rs2;
    return r2_eps < rs2;
}

/* One EPJ pair in FP64, src/force_gpu_cuda.cu:50-77 (dx = xj - xi). */
static inline void epj_pair(const double *xi, double rsi, const double *xj, double mj, double rsj,
                            double eps2, double rout2, int mode,
                            double *ax, double *ay, double *az, double *p, long long *nb)
{
    double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
    double r2raw = dx * dx + dy * dy + dz * dz;
    double r2 = eps2 + r2raw;
    double rs = dmax(rsi, rsj);
    if (nnb_pred(mode, r2, r2raw, rs * rs)) (*nb)++;
    double r2c = (r2 > rout2) ? r2 : rout2;
    double rinv = 1.0 / sqrt(r2c);
    double pij = mj * rinv;
    double mri3 = rinv * rinv * pij;
    *ax += mri3 * dx;
    *ay += mri3 * dy;
    *az += mri3 * dz;
    *p -= pij;
}

/* One quadrupole SPJ pair in FP64, src/force_gpu_cuda.cu:276-323 and
   src/soft_force.hpp:160-200 (dx = xi - xs, raw second moment Q). */
static inline void spj_pair(const double *xi, const double *xs, double m, const double *q,
                            double eps2, double *ax, double *ay, double *az, double *p)
{
    double dx = xi[0] - xs[0], dy = xi[1] - xs[1], dz = xi[2] - xs[2];
    double r2 = eps2 + dx * dx + dy * dy + dz * dz;
    double rinv = 1.0 / sqrt(r2);
    double qxx = q[0], qyy = q[1], qzz = q[2], qxy = q[3], qxz = q[4], qyz = q[5];
    double qrx = qxx * dx + qxy * dy + qxz * dz;
    double qry = qxy * dx + qyy * dy + qyz * dz;
    double qrz = qxz * dx + qyz * dy + qzz * dz;
    double tr = qxx + qyy + qzz;
    double qrr = qrx * dx + qry * dy + qrz * dz;
    double rinv2 = rinv * rinv;
    double rinv3 = rinv2 * rinv;
    double rinv5 = rinv2 * rinv3 * 1.5;
    double qrr_r5 = rinv5 * qrr;
    double qrr_r7 = rinv2 * qrr_r5;
    double A = m * rinv3 - tr * rinv5 + 5.0 * qrr_r7;
    double B = -2.0 * rinv5;
    *ax -= A * dx + B * qrx;
    *ay -= A * dy + B * qry;
    *az -= A * dz + B * qrz;
    *p -= m * rinv - 0.5 * tr * rinv3 + qrr_r5;
}

/* One monopole SPJ pair in FP64, src/soft_force.hpp:125-158
   (CalcForceEpSpMonoNoSimd; rij = xi - xs, no cutoff, no neighbour count). */
static inline void spj_mono_pair(const double *xi, const double *xs, double m, double eps2,
                                 double *ax, double *ay, double *az, double *p)
{
    double rx = xi[0] - xs[0], ry = xi[1] - xs[1], rz = xi[2] - xs[2];
    double r3_inv = rx * rx + ry * ry + rz * rz + eps2;
    double r_inv = 1.0 / sqrt(r3_inv);
    r3_inv = r_inv * r_inv;
    r_inv *= m;
    r3_inv *= r_inv;
    *ax -= r3_inv * rx;
    *ay -= r3_inv * ry;
    *az -= r3_inv * rz;
    *p -= r_inv;
}

static void prefix(int n, const int *cnt, long long *off)
{
    off[0] = 0;
    for (int k = 0; k < n; k++) off[k + 1] = off[k] + cnt[k];
}

#include <stdlib.h>

int orc_epj_walks(int n_walk, const int *n_epi, const int *n_epj,
                  const double *epi_pos, const double *epi_rs, const int *id_epj,
                  const double *epj_pos, const double *epj_mass, const double *epj_rs,
                  double eps2, double rout2, double G, int nnb_mode,
                  double *acc, double *pot, long long *nnb)
{
    long long *io = malloc(sizeof(long long) * (n_walk + 1));
    long long *jo = malloc(sizeof(long long) * (n_walk + 1));
    if (!io || !jo) { free(io); free(jo); return -1; }
    prefix(n_walk, n_epi, io);
    prefix(n_walk, n_epj, jo);
#pragma omp parallel for schedule(dynamic)
    for (int w = 0; w < n_walk; w++) {
        for (long long i = io[w]; i < io[w + 1]; i++) {
            double ax = 0, ay = 0, az = 0, p = 0;
            long long nb = 0;
            for (long long k = jo[w]; k < jo[w + 1]; k++) {
                int j = id_epj[k];
                epj_pair(epi_pos + 3 * i, epi_rs[i], epj_pos + 3 * (long long)j, epj_mass[j],
                         epj_rs[j], eps2, rout2, nnb_mode, &ax, &ay, &az, &p, &nb);
            }
            acc[3 * i + 0] = G * ax;
            acc[3 * i + 1] = G * ay;
            acc[3 * i + 2] = G * az;
            pot[i] = G * p;
            nnb[i] = nb;
        }
    }
    free(io);
    free(jo);
    return 0;
}

int orc_spj_walks(int n_walk, const int *n_epi, const int *n_spj,
                  const double *epi_pos, const int *id_spj,
                  const double *spj_pos, const double *spj_mass, const double *spj_quad,
                  double eps2, double G, double *acc, double *pot)
{
    long long *io = malloc(sizeof(long long) * (n_walk + 1));
    long long *so = malloc(sizeof(long long) * (n_walk + 1));
    if (!io || !so) { free(io); free(so); return -1; }
    prefix(n_walk, n_epi, io);
    prefix(n_walk, n_spj, so);
#pragma omp parallel for schedule(dynamic)
    for (int w = 0; w < n_walk; w++) {
        for (long long i = io[w]; i < io[w + 1]; i++) {
            double ax = 0, ay = 0, az = 0, p = 0;
            for (long long k = so[w]; k < so[w + 1]; k++) {
                long long s = id_spj[k];
                if (spj_quad)
                    spj_pair(epi_pos + 3 * i, spj_pos + 3 * s, spj_mass[s], spj_quad + 6 * s, eps2,
                             &ax, &ay, &az, &p);
                else
                    spj_mono_pair(epi_pos + 3 * i, spj_pos + 3 * s, spj_mass[s], eps2, &ax, &ay, &az, &p);
            }
            acc[3 * i + 0] = G * ax;
            acc[3 * i + 1] = G * ay;
            acc[3 * i + 2] = G * az;
            pot[i] = G * p;
        }
    }
    free(io);
    free(so);
    return 0;
}

int orc_soft_force_walks(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                         const double *epi_pos, const double *epi_rs,
                         const int *id_epj, const int *id_spj,
                         const double *epj_pos, const double *epj_mass, const double *epj_rs,
                         const double *spj_pos, const double *spj_mass, const double *spj_quad,
                         double eps2, double rout2, double G, int nnb_mode,
                         double *acc, double *pot, long long *nnb)
{
    /* src/force_gpu_cuda.cu:692-693: ep_ep assigns, ep_sp then adds into
       the same force record. */
    long long ni = 0;
    for (int w = 0; w < n_walk; w++) ni += n_epi[w];
    double *acc_s = malloc(sizeof(double) * 3 * (ni > 0 ? ni : 1));
    double *pot_s = malloc(sizeof(double) * (ni > 0 ? ni : 1));
    if (!acc_s || !pot_s) { free(acc_s); free(pot_s); return -1; }
    int rc = orc_epj_walks(n_walk, n_epi, n_epj, epi_pos, epi_rs, id_epj, epj_pos, epj_mass,
                           epj_rs, eps2, rout2, G, nnb_mode, acc, pot, nnb);
    if (rc == 0)
        rc = orc_spj_walks(n_walk, n_epi, n_spj, epi_pos, id_spj, spj_pos, spj_mass, spj_quad,
                           eps2, G, acc_s, pot_s);
    if (rc == 0) {
        for (long long i = 0; i < ni; i++) {
            acc[3 * i + 0] += acc_s[3 * i + 0];
            acc[3 * i + 1] += acc_s[3 * i + 1];
            acc[3 * i + 2] += acc_s[3 * i + 2];
            pot[i] += pot_s[i];
        }
    }
    free(acc_s);
    free(pot_s);
    return rc;
}

int orc_nnb_ambiguous_walks(int n_walk, const int *n_epi, const int *n_epj,
                            const double *epi_pos, const double *epi_rs, const int *id_epj,
                            const double *epj_pos, const double *epj_rs,
                            const double *centres, double eps2, double rel,
                            long long *n_amb)
{
    long long *io = malloc(sizeof(long long) * (n_walk + 1));
    long long *jo = malloc(sizeof(long long) * (n_walk + 1));
    if (!io || !jo) { free(io); free(jo); return -1; }
    prefix(n_walk, n_epi, io);
    prefix(n_walk, n_epj, jo);
#pragma omp parallel for schedule(dynamic)
    for (int w = 0; w < n_walk; w++) {
        double c[3] = {0, 0, 0};
        if (centres) { c[0] = centres[3 * w]; c[1] = centres[3 * w + 1]; c[2] = centres[3 * w + 2]; }
        for (long long i = io[w]; i < io[w + 1]; i++) {
            const double *xi = epi_pos + 3 * i;
            double di = sqrt((xi[0] - c[0]) * (xi[0] - c[0]) + (xi[1] - c[1]) * (xi[1] - c[1]) +
                             (xi[2] - c[2]) * (xi[2] - c[2]));
            long long na = 0;
            for (long long k = jo[w]; k < jo[w + 1]; k++) {
                const double *xj = epj_pos + 3 * (long long)id_epj[k];
                double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
                double r2raw = dx * dx + dy * dy + dz * dz;
                double r2 = r2raw + eps2;
                double rs = dmax(epi_rs[i], epj_rs[id_epj[k]]);
                double rs2 = rs * rs;
                double dj = sqrt((xj[0] - c[0]) * (xj[0] - c[0]) + (xj[1] - c[1]) * (xj[1] - c[1]) +
                                 (xj[2] - c[2]) * (xj[2] - c[2]));
                double band = rel * (rs2 + r2 + 2.0 * sqrt(r2raw) * (di + dj));
                if (fabs(r2 - rs2) <= band) na++;
            }
            n_amb[i] = na;
        }
    }
    free(io);
    free(jo);
    return 0;
}

/* One row of the all-pairs sum: every j against particle i.  The pair is
   epj_pair's arithmetic term for term (same order, no contraction); only the
   summation over j is split into vector-lane partial sums (an OpenMP simd
   reduction, which permits exactly that reordering and no other).  Cloned
   for AVX-512, AVX2 and plain x86-64, picked at load time. */
static inline __attribute__((always_inline)) void direct_row_body(long long n, const double *x, const double *y, const double *z,
                                   const double *m, const double *rsj, double xi, double yi, double zi,
                                   double rsi, double eps2, double rout2, const int mode, double out[4],
                                   long long *nb_out)
{
    double ax = 0, ay = 0, az = 0, p = 0;
    long long nb = 0;
#pragma omp simd reduction(+ : ax, ay, az, p, nb)
    for (long long j = 0; j < n; j++) {
        const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
        const double r2raw = dx * dx + dy * dy + dz * dz;
        const double r2 = eps2 + r2raw;
        const double rs = rsi > rsj[j] ? rsi : rsj[j];
        const double rs2 = rs * rs;
        nb += (mode == 1) ? (r2 <= rs2) : (mode == 2) ? (r2raw < rs2) : (r2 < rs2);
        const double r2c = (r2 > rout2) ? r2 : rout2;
        const double rinv = 1.0 / sqrt(r2c);
        const double pij = m[j] * rinv;
        const double mri3 = rinv * rinv * pij;
        ax += mri3 * dx;
        ay += mri3 * dy;
        az += mri3 * dz;
        p -= pij;
    }
    out[0] = ax;
    out[1] = ay;
    out[2] = az;
    out[3] = p;
    *nb_out = nb;
}

__attribute__((target_clones("avx512f", "avx2", "default")))
static void direct_row(long long n, const double *x, const double *y, const double *z, const double *m,
                       const double *rsj, double xi, double yi, double zi, double rsi, double eps2,
                       double rout2, int mode, double out[4], long long *nb_out)
{
    if (mode == 1) direct_row_body(n, x, y, z, m, rsj, xi, yi, zi, rsi, eps2, rout2, 1, out, nb_out);
    else if (mode == 2) direct_row_body(n, x, y, z, m, rsj, xi, yi, zi, rsi, eps2, rout2, 2, out, nb_out);
    else direct_row_body(n, x, y, z, m, rsj, xi, yi, zi, rsi, eps2, rout2, 0, out, nb_out);
}

int orc_direct_linear_cutoff(int n, const double *pos, const double *mass, const double *rs,
                             int n_sample, const int *idx_sample,
                             double eps2, double rout2, double G, int nnb_mode,
                             double *acc, double *pot, long long *nnb)
{
    /* coordinates as three planes, so each partial sum reads unit-stride */
    double *xs = malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    if (!xs) return -1;
    double *ys = xs + n, *zs = ys + n;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++) {
        xs[j] = pos[3 * (size_t)j];
        ys[j] = pos[3 * (size_t)j + 1];
        zs[j] = pos[3 * (size_t)j + 2];
    }
    /* tiles of DS_TI sampled rows against chunks of DS_JB j (which stay in
       the core's L2 while the tile's rows pass over them): each row's sum
       runs over the chunks in order */
    enum { DS_TI = 16, DS_JB = 4096 };
    const int n_tile = (n_sample + DS_TI - 1) / DS_TI;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < n_tile; t++) {
        const int s0 = t * DS_TI, s1 = (s0 + DS_TI < n_sample) ? s0 + DS_TI : n_sample;
        double A[DS_TI][4];
        long long NB[DS_TI];
        for (int s = s0; s < s1; s++) {
            A[s - s0][0] = A[s - s0][1] = A[s - s0][2] = A[s - s0][3] = 0.0;
            NB[s - s0] = 0;
        }
        for (long long jb = 0; jb < n; jb += DS_JB) {
            const long long nj = (n - jb < DS_JB) ? n - jb : DS_JB;
            for (int s = s0; s < s1; s++) {
                const long long i = idx_sample ? idx_sample[s] : s;
                double r[4];
                long long nb = 0;
                direct_row(nj, xs + jb, ys + jb, zs + jb, mass + jb, rs + jb, xs[i], ys[i], zs[i], rs[i], eps2,
                           rout2, nnb_mode, r, &nb);
                for (int d = 0; d < 4; d++) A[s - s0][d] += r[d];
                NB[s - s0] += nb;
            }
        }
        for (int s = s0; s < s1; s++) {
            acc[3 * s + 0] = G * A[s - s0][0];
            acc[3 * s + 1] = G * A[s - s0][1];
            acc[3 * s + 2] = G * A[s - s0][2];
            pot[s] = G * A[s - s0][3];
            nnb[s] = NB[s - s0];
        }
    }
    free(xs);
    return 0;
}

int orc_direct_newton(int n, const double *pos, const double *mass,
                      int n_sample, const int *idx_sample, double eps2, double G,
                      double *acc, double *pot)
{
    /* src/soft_force.hpp:202-236 (CalcForcePPNoSimd) with the self pair
       skipped instead of softened. */
#pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < n_sample; s++) {
        long long i = idx_sample ? idx_sample[s] : s;
        double ax = 0, ay = 0, az = 0, p = 0;
        for (long long j = 0; j < n; j++) {
            if (j == i) continue;
            double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                   dz = pos[3 * i + 2] - pos[3 * j + 2];
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double rinv = 1.0 / sqrt(r2);
            double mr = mass[j] * rinv;
            double mr3 = mr * rinv * rinv;
            ax -= mr3 * dx;
            ay -= mr3 * dy;
            az -= mr3 * dz;
            p -= mr;
        }
        acc[3 * s + 0] = G * ax;
        acc[3 * s + 1] = G * ay;
        acc[3 * s + 2] = G * az;
        pot[s] = G * p;
    }
    return 0;
}

void orc_moments(int n, const double *pos, const double *mass,
                 double *m_out, double *com_out, double *quad_out)
{
    double m = 0, c[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        m += mass[i];
        for (int d = 0; d < 3; d++) c[d] += mass[i] * pos[3 * i + d];
    }
    for (int d = 0; d < 3; d++) c[d] = (m > 0) ? c[d] / m : 0.0;
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        double x = pos[3 * i] - c[0], y = pos[3 * i + 1] - c[1], z = pos[3 * i + 2] - c[2];
        q[0] += mass[i] * x * x;
        q[1] += mass[i] * y * y;
        q[2] += mass[i] * z * z;
        q[3] += mass[i] * x * y;
        q[4] += mass[i] * x * z;
        q[5] += mass[i] * y * z;
    }
    *m_out = m;
    for (int d = 0; d < 3; d++) com_out[d] = c[d];
    for (int d = 0; d < 6; d++) quad_out[d] = q[d];
}

/* ---- changeover, src/changeover.hpp:44-77, 294-381 ---------------------- */

void orc_changeover_setr(double m_fac, double r_in, double r_out,
                         double *r_in_out, double *r_out_out)
{
    /* src/changeover.hpp:44-59: radii scale with max(m_fac^(1/3), 1). */
    double f = pow(m_fac, 1.0 / 3.0);
    if (f < 1.0) f = 1.0;
    *r_in_out = f * r_in;
    *r_out_out = f * r_out;
}

double orc_changeover_wpot(double r, double r_in, double r_out)
{
    /* src/changeover.hpp:294-309 */
    double norm = 1.0 / (r_out - r_in);
    double coff = (r_out - r_in) / (r_out + r_in);
    double pot_off = (1.0 + coff) / r_out;
    double x = (r - r_in) * norm;
    double k = 1.0;
    if (x >= 1.0) k = pot_off * r;
    else if (x > 0.0) {
        double x2 = x * x, x3 = x2 * x, x5 = x2 * x3;
        k -= coff * x5 * (5.0 * x3 - 20.0 * x2 + 28.0 * x - 14.0);
    }
    return k;
}

double orc_changeover_wacc(double r, double r_in, double r_out)
{
    /* src/changeover.hpp:318-334 */
    double norm = 1.0 / (r_out - r_in);
    double coff = (r_out - r_in) / (r_out + r_in);
    double x = (r - r_in) * norm;
    x = (x < 1.0) ? x : 1.0;
    x = (x > 0.0) ? x : 0.0;
    double x_1 = x - 1.0, x_2 = x_1 * x_1, x_4 = x_2 * x_2;
    double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
    return x_4 * (1.0 + 4.0 * x + 10.0 * x2 + 20.0 * x3 + 35.0 * coff * x4);
}

double orc_changeover_wacc1(double r, double rdot, double r_in, double r_out)
{
    /* src/changeover.hpp:346-361 */
    double norm = 1.0 / (r_out - r_in);
    double coff = (r_out - r_in) / (r_out + r_in);
    double x = (r - r_in) * norm;
    double xdot = norm * rdot;
    double kdot = 0.0;
    if (x > 0.0 && x < 1.0) {
        double x3 = x * x * x, x_1 = x - 1.0, x_3 = x_1 * x_1 * x_1;
        kdot = coff * 280.0 * x3 * (r_in * norm + x) * x_3 * xdot;
    }
    return kdot;
}

double orc_changeover_wpot_two(double r, double rin1, double rout1, double rin2, double rout2)
{
    /* src/changeover.hpp:366-369: the larger r_out wins, ties go to the second. */
    return (rout1 > rout2) ? orc_changeover_wpot(r, rin1, rout1) : orc_changeover_wpot(r, rin2, rout2);
}

double orc_changeover_wacc_two(double r, double rin1, double rout1, double rin2, double rout2)
{
    /* src/changeover.hpp:372-375 */
    return (rout1 > rout2) ? orc_changeover_wacc(r, rin1, rout1) : orc_changeover_wacc(r, rin2, rout2);
}

/* Particle types of PeTar's group data (ArtificialParticleInformation,
   src/artificial_particles.hpp:20-102: g[0] = mass_backup, g[1] = status):
   single (status 0, mass_backup 0), member (status < 0), otherwise an
   artificial or c.m. particle (status > 0).  A member's c.m. address
   (src/ptcl.hpp:215-222) is unset (getParticleCMAddress() < 0) when its
   status is -PS::LARGE_FLOAT; FDPS's constant is not available here, and
   every set address is an array index, so "unset" is taken as status below
   -1e30. */
static int grp_single(const double *g) { return !g || (g[1] == 0.0 && g[0] == 0.0); }
static int grp_member(const double *g) { return g && g[1] < 0.0; }
static int grp_member_no_cm(const double *g) { return g && g[1] < -1e30; }

void orc_changeover_correct_one(const double *pos_i, double mass_i, double rin_i, double rout_i,
                                long long id_i, int n_ngb, const double *pos_j,
                                const double *mass_j, const double *rin_j, const double *rout_j,
                                const long long *id_j, double eps2, double r_out_glb, double G,
                                double *acc, double *pot_soft, double *pot_tot, const double *grp_i,
                                const double *grp_j)
{
    /* self-potential term, src/hard.hpp:1669-1673: singles and members
       without a c.m. address (binaries without artificial particles) */
    if (grp_single(grp_i) || grp_member_no_cm(grp_i)) {
        double pot_cor = G * mass_i / r_out_glb;
        *pot_tot += pot_cor;
        *pot_soft += pot_cor;
    }
    for (int k = 0; k < n_ngb; k++) {
        if (id_j[k] == id_i) continue; /* src/hard.hpp:1677 */
        /* FP64 branch of src/hard.hpp:1404-1476 (single j) */
        double dr[3] = {pos_i[0] - pos_j[3 * k], pos_i[1] - pos_j[3 * k + 1], pos_i[2] - pos_j[3 * k + 2]};
        double dr2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        double dr2_eps = dr2 + eps2;
        double drinv = 1.0 / sqrt(dr2_eps);
        double gmor = G * mass_j[k] * drinv;
        double drinv2 = drinv * drinv;
        double gmor3 = gmor * drinv2;
        double dr_eps = drinv * dr2_eps;
        double kacc = 1.0 - orc_changeover_wacc_two(dr_eps, rin_i, rout_i, rin_j[k], rout_j[k]);
        double r_out2 = r_out_glb * r_out_glb;
        double dr2_max = (dr2_eps > r_out2) ? dr2_eps : r_out2;
        double drinv_max = 1.0 / sqrt(dr2_max);
        double gmor_max = G * mass_j[k] * drinv_max;
        double drinv2_max = drinv_max * drinv_max;
        double gmor3_max = gmor_max * drinv2_max;
        for (int d = 0; d < 3; d++) acc[d] -= (gmor3 * kacc - gmor3_max) * dr[d];
        double kpot = 1.0 - orc_changeover_wpot_two(dr_eps, rin_i, rout_i, rin_j[k], rout_j[k]);
        const double *gj = grp_j ? grp_j + 2 * k : NULL;
        if (grp_single(gj)) {
            *pot_soft -= gmor * kpot - gmor_max;
            *pot_tot -= gmor - gmor_max;
        } else if (grp_member(gj)) {
            /* src/hard.hpp:1459-1464: the member's mass is zero in the tree;
               its potential uses the backup mass */
            gmor = G * gj[0] * drinv;
            *pot_soft -= gmor * kpot - gmor_max;
            *pot_tot -= gmor - gmor_max;
        } else {
            /* src/hard.hpp:1466-1470: artificial particles leave no
               potential (the linear-cutoff term is cancelled) */
            *pot_soft += gmor_max;
            *pot_tot += gmor_max;
        }
    }
}

/* Changeover mode over walks: the linear-cutoff result of
   orc_soft_force_walks (nnb_mode 0), then for every i the FP64 correction
   of src/hard.hpp:1655-1691 (self-potential term, then
   src/hard.hpp:1404-1476 per neighbour, all j singles) over the neighbour set
   taken from the walk's own EPJ list: j with |x_i - x_j|^2 <
   (f max(rs_i, rs_j))^2 (FDPS's symmetric search on getRSearch() = f r_search,
   src/soft_ptcl.hpp:97-101,349-353), the correction itself skipping id_j ==
   id_i.  The i's mass and changeover radii come from its own EPJ entry (same
   id) in the list.  n_amb[i] counts pairs with |r^2 - (f rs)^2| <= amb_rel
   (f rs)^2 (membership undecided at FP32 rounding).  Returns -1 if some i has
   no own entry, else 0.  pot_soft and pot_tot are both assigned. */
int orc_soft_force_walks_co(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                            const double *epi_pos, const double *epi_rs, const long long *epi_id,
                            const int *id_epj, const int *id_spj,
                            const double *epj_pos, const double *epj_mass, const double *epj_rs,
                            const long long *epj_id, const double *epj_rin, const double *epj_rout,
                            const double *spj_pos, const double *spj_mass, const double *spj_quad,
                            double eps2, double rout2, double G, double ngb_factor, double amb_rel,
                            double *acc, double *pot_soft, double *pot_tot, long long *nnb,
                            long long *n_amb, const double *epj_group)
{
    int rc = orc_soft_force_walks(n_walk, n_epi, n_epj, n_spj, epi_pos, epi_rs, id_epj, id_spj, epj_pos,
                                  epj_mass, epj_rs, spj_pos, spj_mass, spj_quad, eps2, rout2, G, 0, acc,
                                  pot_soft, nnb);
    if (rc) return rc;
    long long *io = malloc(sizeof(long long) * (n_walk + 1));
    long long *eo = malloc(sizeof(long long) * (n_walk + 1));
    prefix(n_walk, n_epi, io);
    prefix(n_walk, n_epj, eo);
    const double r_out = sqrt(rout2);
    int missing = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : missing)
    for (int w = 0; w < n_walk; w++) {
        const int ne = n_epj[w];
        double *pj = malloc(sizeof(double) * 3 * (ne > 0 ? ne : 1));
        double *mj = malloc(sizeof(double) * (ne > 0 ? ne : 1));
        double *ri = malloc(sizeof(double) * (ne > 0 ? ne : 1));
        double *ro = malloc(sizeof(double) * (ne > 0 ? ne : 1));
        long long *idj = malloc(sizeof(long long) * (ne > 0 ? ne : 1));
        double *gj = malloc(sizeof(double) * 2 * (ne > 0 ? ne : 1));
        const int *lst = id_epj + eo[w];
        for (long long i = io[w]; i < io[w + 1]; i++) {
            const double *xi = epi_pos + 3 * i;
            pot_tot[i] = pot_soft[i];
            n_amb[i] = 0;
            int self = -1;
            for (int k = 0; k < ne; k++)
                if (epj_id[lst[k]] == epi_id[i]) {
                    self = lst[k];
                    break;
                }
            if (self < 0) {
                missing = 1;
                continue;
            }
            int nn = 0;
            for (int k = 0; k < ne; k++) {
                const int j = lst[k];
                const double dx = xi[0] - epj_pos[3 * j], dy = xi[1] - epj_pos[3 * j + 1],
                             dz = xi[2] - epj_pos[3 * j + 2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double rs = ngb_factor * dmax(epi_rs[i], epj_rs[j]);
                const double rs2 = rs * rs;
                if (fabs(r2 - rs2) <= amb_rel * rs2 && epj_id[j] != epi_id[i]) n_amb[i]++;
                if (!(r2 < rs2)) continue;
                pj[3 * nn] = epj_pos[3 * j];
                pj[3 * nn + 1] = epj_pos[3 * j + 1];
                pj[3 * nn + 2] = epj_pos[3 * j + 2];
                mj[nn] = epj_mass[j];
                ri[nn] = epj_rin[j];
                ro[nn] = epj_rout[j];
                idj[nn] = epj_id[j];
                gj[2 * nn] = epj_group ? epj_group[2 * j] : 0.0;
                gj[2 * nn + 1] = epj_group ? epj_group[2 * j + 1] : 0.0;
                nn++;
            }
            orc_changeover_correct_one(xi, epj_mass[self], epj_rin[self], epj_rout[self], epi_id[i], nn, pj, mj,
                                       ri, ro, idj, eps2, r_out, G, acc + 3 * i, pot_soft + i, pot_tot + i,
                                       epj_group ? epj_group + 2 * (long long)self : NULL, gj);
        }
        free(pj);
        free(mj);
        free(ri);
        free(ro);
        free(idj);
        free(gj);
    }
    free(io);
    free(eo);
    return missing ? -1 : 0;
}


/* Neighbour lists over walks, src/soft_force.hpp:11-34
 * (SearchNeighborEpEpNoSimd: rij = xi - xj, r2 = rij*rij in FP64,
 * r_search = max(rs_i, rs_j), neighbour when r2 < r_search^2; self
 * included) applied to each walk's EPJ list, as FDPS's
 * getNeighborListOneParticle returns the list PeTar's cluster search reads
 * (src/cluster_list.hpp:461-493).  count[i] is assigned for the walk-ordered
 * i; with ngb != NULL the entries (EPJ indices, list order) are also written
 * at ngb[offset[i]..].  Returns the total count. */
long long orc_ngb_walks(int n_walk, const int *n_epi, const int *n_epj, const double *epi_pos,
                        const double *epi_rs, const int *id_epj, const double *epj_pos,
                        const double *epj_rs, long long *count, const long long *offset, int *ngb)
{
    long long *io = (long long *)malloc(sizeof(long long) * (size_t)(n_walk + 1));
    long long *jo = (long long *)malloc(sizeof(long long) * (size_t)(n_walk + 1));
    io[0] = jo[0] = 0;
    for (int w = 0; w < n_walk; w++) {
        io[w + 1] = io[w] + n_epi[w];
        jo[w + 1] = jo[w] + n_epj[w];
    }
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
    for (int w = 0; w < n_walk; w++) {
        for (long long i = io[w]; i < io[w + 1]; i++) {
            const double *xi = epi_pos + 3 * i;
            long long c = 0;
            for (long long k = jo[w]; k < jo[w + 1]; k++) {
                const int j = id_epj[k];
                const double *xj = epj_pos + 3 * (size_t)j;
                const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double rs = (epi_rs[i] < epj_rs[j]) ? epj_rs[j] : epi_rs[i];
                if (r2 < rs * rs) {
                    if (ngb) ngb[offset[i] + c] = j;
                    c++;
                }
            }
            count[i] = c;
            total += c;
        }
    }
    free(io);
    free(jo);
    return total;
}

