// ref_phantomquad.cpp -- driver that runs the reference x86 SIMD soft-force
// kernel (src/phantomquad_for_p3t_x86.hpp, compiled from /root/reference by
// oracle/Makefile) over a set of walks.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ as a checker for the oracle and by
// bench.py as the CPU baseline ("kind": "reference").  Nothing of the product
// links it.
//
// The per-walk logic restates the FDPS functor wrappers that PeTar uses with
// this kernel; the wrappers themselves (src/soft_force.hpp) need FDPS and
// SDAR headers that the image lacks, so they are restated, not compiled:
//   EP-EP  src/soft_force.hpp:350-425  (j with mass>0 only, i with type==1
//          only, pg.set_r_crit2(r_out^2), x G, +=, n_ngb += (S32)(nngb*1.00001))
//   EP-SP  src/soft_force.hpp:491-554  (i with type==1 only, x G, +=)
//   EP-SP monopole (do_spj == 2) src/soft_force.hpp:427-480: the plain
//          gravity kernel (run_epj) on the superparticles' mass and centre
// The kernel header only needs the PS::S32 name for its diagnostics; the
// driver defines it here before the include (no FDPS header is stood in).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <vector>
#include <omp.h>

namespace PS {
typedef int S32;
}

// src/petar.hpp:12-15 selects one Newton-Raphson step for EPJ and none for SPJ.
#ifndef RSQRT_NR_EPJ_X2
#define RSQRT_NR_EPJ_X2
#endif
#include "phantomquad_for_p3t_x86.hpp"

namespace {

struct Pool {
    std::vector<PhantomGrapeQuad *> pg;
    Pool()
    {
        int n = omp_get_max_threads();
        pg.resize(n, nullptr);
    }
    PhantomGrapeQuad *get(int t)
    {
        if (!pg[t]) {
            // 64-byte alignment for the kernel's aligned vector loads
            void *mem = std::aligned_alloc(64, (sizeof(PhantomGrapeQuad) + 63) / 64 * 64);
            if (!mem) return nullptr;
            pg[t] = new (mem) PhantomGrapeQuad();
        }
        return pg[t];
    }
    ~Pool()
    {
        for (auto p : pg)
            if (p) {
                p->~PhantomGrapeQuad();
                std::free(p);
            }
    }
};

Pool &pool()
{
    static Pool p;
    return p;
}

}  // namespace

extern "C" {

int ref_isa_bits()
{
#ifdef USE__AVX512
    return 512;
#else
    return 256;
#endif
}

int ref_njmax() { return PhantomGrapeQuad::NJMAX; }
int ref_nimax() { return PhantomGrapeQuad::NIMAX; }

// acc/pot/nnb are assigned (zeroed, then the wrappers' += applied).
// epi_type may be NULL (all type 1).  nthreads <= 0 uses the OpenMP default.
int ref_soft_force_walks(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                         const double *epi_pos, const double *epi_rs, const int *epi_type,
                         const int *id_epj, const int *id_spj,
                         const double *epj_pos, const double *epj_mass, const double *epj_rs,
                         const double *spj_pos, const double *spj_mass, const double *spj_quad,
                         double eps2, double rout2, double G, int do_epj, int do_spj,
                         int nthreads, double *acc, double *pot, long long *nnb)
{
    std::vector<long long> io(n_walk + 1), jo(n_walk + 1), so(n_walk + 1);
    io[0] = jo[0] = so[0] = 0;
    for (int w = 0; w < n_walk; w++) {
        if (n_epi[w] > PhantomGrapeQuad::NIMAX) return -2;
        io[w + 1] = io[w] + n_epi[w];
        jo[w + 1] = jo[w] + n_epj[w];
        so[w + 1] = so[w] + n_spj[w];
    }
    long long ni = io[n_walk];
    std::memset(acc, 0, sizeof(double) * 3 * ni);
    std::memset(pot, 0, sizeof(double) * ni);
    std::memset(nnb, 0, sizeof(long long) * ni);
    Pool &P = pool();
    int nt = (nthreads > 0) ? nthreads : omp_get_max_threads();
    if (nt > (int)P.pg.size()) nt = (int)P.pg.size();
    int fail = 0;
#pragma omp parallel num_threads(nt)
    {
        PhantomGrapeQuad *pg = P.get(omp_get_thread_num());
        if (!pg) {
#pragma omp atomic write
            fail = 1;
        }
        std::vector<int> ilist, jlist;
#pragma omp for schedule(dynamic)
        for (int w = 0; w < n_walk; w++) {
            if (!pg) continue;
            const int n_ip = n_epi[w];
            const double *xi = epi_pos + 3 * io[w];
            ilist.clear();
            if (do_epj) {
                // ---- EP-EP, src/soft_force.hpp:350-425
                jlist.clear();
                for (long long k = jo[w]; k < jo[w + 1]; k++)
                    if (epj_mass[id_epj[k]] > 0) jlist.push_back(id_epj[k]);
                const int n_jp_local = (int)jlist.size();
                pg->set_eps2(eps2);
                pg->set_r_crit2(rout2);
                int n_ip_local = 0;
                for (int i = 0; i < n_ip; i++) {
                    if (epi_type && epi_type[io[w] + i] != 1) continue;
                    ilist.push_back(i);
                    pg->set_xi_one(n_ip_local, xi[3 * i], xi[3 * i + 1], xi[3 * i + 2],
                                   epi_rs[io[w] + i]);
                    n_ip_local++;
                }
                const int loop_max = (n_jp_local - 1) / PhantomGrapeQuad::NJMAX + 1;
                for (int loop = 0; loop < loop_max; loop++) {
                    const int ih = PhantomGrapeQuad::NJMAX * loop;
                    const int n_jp_tmp = (n_jp_local - ih < PhantomGrapeQuad::NJMAX)
                                             ? n_jp_local - ih : PhantomGrapeQuad::NJMAX;
                    for (int t = 0; t < n_jp_tmp; t++) {
                        const int j = jlist[ih + t];
                        pg->set_epj_one(t, epj_pos[3 * (long long)j], epj_pos[3 * (long long)j + 1],
                                        epj_pos[3 * (long long)j + 2], epj_mass[j], epj_rs[j]);
                    }
                    pg->run_epj_for_p3t_with_linear_cutoff(n_ip, n_jp_tmp);
                    for (int k = 0; k < n_ip_local; k++) {
                        const long long i = io[w] + ilist[k];
                        double p = 0, a[3] = {0, 0, 0}, ng = 0;
                        pg->accum_accp_one(k, a[0], a[1], a[2], p, ng);
                        acc[3 * i + 0] += G * a[0];
                        acc[3 * i + 1] += G * a[1];
                        acc[3 * i + 2] += G * a[2];
                        pot[i] += G * p;
                        nnb[i] += (int)(ng * 1.00001);
                    }
                }
            }
            if (do_spj == 2) {
                // ---- EP-SP monopole, src/soft_force.hpp:427-480
                const int n_jp = n_spj[w];
                pg->set_eps2(eps2);
                ilist.clear();
                int n_ip_local = 0;
                for (int i = 0; i < n_ip; i++) {
                    if (epi_type && epi_type[io[w] + i] != 1) continue;
                    ilist.push_back(i);
                    pg->set_xi_one(n_ip_local, xi[3 * i], xi[3 * i + 1], xi[3 * i + 2], 0.0);
                    n_ip_local++;
                }
                const int loop_max = (n_jp - 1) / PhantomGrapeQuad::NJMAX + 1;
                for (int loop = 0; loop < loop_max; loop++) {
                    const int ih = PhantomGrapeQuad::NJMAX * loop;
                    const int n_jp_tmp = (n_jp - ih < PhantomGrapeQuad::NJMAX)
                                             ? n_jp - ih : PhantomGrapeQuad::NJMAX;
                    for (int t = 0; t < n_jp_tmp; t++) {
                        const long long s = id_spj[so[w] + ih + t];
                        pg->set_epj_one(t, spj_pos[3 * s], spj_pos[3 * s + 1], spj_pos[3 * s + 2], spj_mass[s], 0.0);
                    }
                    pg->run_epj(n_ip, n_jp_tmp);
                    for (int k = 0; k < n_ip_local; k++) {
                        const long long i = io[w] + ilist[k];
                        double p = 0, a[3] = {0, 0, 0};
                        pg->accum_accp_one(k, a[0], a[1], a[2], p);
                        acc[3 * i + 0] += G * a[0];
                        acc[3 * i + 1] += G * a[1];
                        acc[3 * i + 2] += G * a[2];
                        pot[i] += G * p;
                    }
                }
            } else if (do_spj) {
                // ---- EP-SP quadrupole, src/soft_force.hpp:491-554
                const int n_jp = n_spj[w];
                pg->set_eps2(eps2);
                ilist.clear();
                int n_ip_local = 0;
                for (int i = 0; i < n_ip; i++) {
                    if (epi_type && epi_type[io[w] + i] != 1) continue;
                    ilist.push_back(i);
                    pg->set_xi_one(n_ip_local, xi[3 * i], xi[3 * i + 1], xi[3 * i + 2], 0.0);
                    n_ip_local++;
                }
                const int loop_max = (n_jp - 1) / PhantomGrapeQuad::NJMAX + 1;
                for (int loop = 0; loop < loop_max; loop++) {
                    const int ih = PhantomGrapeQuad::NJMAX * loop;
                    const int n_jp_tmp = (n_jp - ih < PhantomGrapeQuad::NJMAX)
                                             ? n_jp - ih : PhantomGrapeQuad::NJMAX;
                    for (int t = 0; t < n_jp_tmp; t++) {
                        const long long s = id_spj[so[w] + ih + t];
                        const double *q = spj_quad + 6 * s;  // xx yy zz xy xz yz
                        pg->set_spj_one(t, spj_pos[3 * s], spj_pos[3 * s + 1], spj_pos[3 * s + 2],
                                        spj_mass[s], q[0], q[1], q[2], q[3], q[5], q[4]);
                    }
                    pg->run_spj(n_ip, n_jp_tmp);
                    for (int k = 0; k < n_ip_local; k++) {
                        const long long i = io[w] + ilist[k];
                        double p = 0, a[3] = {0, 0, 0};
                        pg->accum_accp_one(k, a[0], a[1], a[2], p);
                        acc[3 * i + 0] += G * a[0];
                        acc[3 * i + 1] += G * a[1];
                        acc[3 * i + 2] += G * a[2];
                        pot[i] += G * p;
                    }
                }
            }
        }
    }
    return fail ? -1 : 0;
}

}  // extern "C"
