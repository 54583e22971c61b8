// fdps_replay.cpp -- drives include/petar_hip_fdps.hpp the way FDPS's
// calcForceAllAndWriteBackMultiWalkIndex drives the reference GPU functor
// (src/petar.hpp:887-899): one send_flag == true call with every EPJ and SPJ,
// then batches of at most n_walk_limit walks, each dispatched (send_flag ==
// false) and retrieved into ForceSoft records, with the names petar.hpp uses
// (CalcForceWithLinearCutoffCUDAMultiWalk, RetrieveForceCUDA).
//
// The walks come from the repository's own tree (libpetar_host).  The
// program writes the inputs it used and the retrieved forces to a binary
// file that tests/test_gpu_dropin.py checks against the FP64 oracle.
//
// Built with -DPETAR_HIP_CHANGEOVER (fdps_replay_co) it runs the adapter in
// changeover mode: changeover radii and r_search scale with
// max((m/<m>)^(1/3), 1) per particle (src/changeover.hpp:44-59), the
// retrieved ForceSoft carries pot_tot, and the EPJ group data carries every
// particle type: of each 50 particles (by input index) 5 are group members
// whose c.m. holds their mass (tree mass 0, backup mass, status -(address)),
// 2 members of groups without artificial particles (status -LARGE_FLOAT)
// and 3 artificial particles (status > 0).
//
// With a timed-pass count the FDPS call sequence (send, then every batch
// dispatched and retrieved) is repeated that many times after one untimed
// pass, and each pass's wall time and interaction rate are printed: the
// drop-in rate with every list crossing PCIe, from host records in and to
// host records out.
//
// Built without USE_QUAD (fdps_replay_mono) the superparticles are
// PS::SPJMonopoleInAndOut records, as in a PeTar build configured with
// --disable-quad.  Built with GPU_PROFILE (fdps_replay_prof) it prints
// gpu_counter and gpu_profile after the passes, as PeTar's profile output
// does (src/petar.hpp:2123-2130).
//
// usage: fdps_replay N seed theta n_walk_limit r_out out_file [f_bin [timed_passes]]
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <iostream>

#include "particle_simulator.hpp"
#include "soft_types.hpp"
#include "petar_hip_fdps.hpp"
#include "petar_host.h"

static void die(const char *what)
{
    std::fprintf(stderr, "fdps_replay: %s: %s\n", what, pth_host_last_error());
    std::exit(2);
}

template <class T> static void put(FILE *f, const std::vector<T> &v)
{
    if (!v.empty() && std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) die("write");
}

int main(int argc, char **argv)
{
    if (argc < 7 || argc > 9) {
        std::fprintf(stderr, "usage: %s N seed theta n_walk_limit r_out out_file [f_bin [timed_passes]]\n",
                     argv[0]);
        return 1;
    }
    const double f_bin = argc > 7 ? std::atof(argv[7]) : 0.3;
    const int timed_passes = argc > 8 ? std::atoi(argv[8]) : 0;
    const long n = std::atol(argv[1]);
    const unsigned seed = (unsigned)std::atoi(argv[2]);
    const double theta = std::atof(argv[3]);
    const int n_walk_limit = std::atoi(argv[4]);
    const double r_out = std::atof(argv[5]);
    const double r_search = 1.5 * r_out;

    std::vector<double> mass(n), pos(3 * n);
    if (pth_ic_cluster(n, 0, 0.0, 1, f_bin, 1e-5, 1e-2, 0, seed, mass.data(), pos.data(), nullptr)) die("ic");
#ifdef PETAR_HIP_CHANGEOVER
    double mbar = 0.0;
    for (double m : mass) mbar += m;
    mbar /= (double)n;
    auto fac = [&](double m) { return std::max(std::cbrt(m / mbar), 1.0); };
#else
    auto fac = [](double) { return 1.0; };
#endif
    std::vector<double> rs(n);
    for (long i = 0; i < n; i++) rs[i] = r_search * fac(mass[i]);
    pth_tree *tree = nullptr;
    if (pth_tree_build(&tree, n, pos.data(), mass.data(), rs.data(), 20, 512, 0)) die("tree");
    if (pth_tree_walk(tree, theta)) die("walk");
    long n_node, n_group, n_le, n_ls;
    pth_tree_sizes(tree, &n_node, &n_group, &n_le, &n_ls);
    std::vector<long> order(n), nb(n_node), ne(n_node), epi_begin(n_group);
    std::vector<int> lvl(n_node), n_epi(n_group), n_epj(n_group), n_spj(n_group);
    std::vector<double> smass(n_node), scom(3 * n_node), squad(6 * n_node);
    std::vector<int> id_epj(n_le), id_spj(n_ls);
    pth_tree_order(tree, order.data());
    pth_tree_nodes(tree, smass.data(), scom.data(), squad.data(), nb.data(), ne.data(), lvl.data());
    pth_tree_groups(tree, epi_begin.data(), n_epi.data(), n_epj.data(), n_spj.data());
    pth_tree_lists(tree, id_epj.data(), id_spj.data());
    pth_tree_free(tree);

    // FDPS-side records (tree order)
    std::vector<EPJSoft> epj(n);
    std::vector<double> epj_pos(3 * n), epj_mass(n), epj_rs(n), epj_rin(n), epj_rout(n), epj_group(2 * n, 0.0);
    std::vector<long long> epj_id(n);
    for (long k = 0; k < n; k++) {
        const long i = order[k];
        epj[k] = EPJSoft{};
        epj[k].id = i + 1;
        epj[k].mass = mass[i];
        epj[k].pos = PS::F64vec(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        epj[k].r_in = 0.1 * r_out * fac(mass[i]);
        epj[k].r_out = r_out * fac(mass[i]);
        epj[k].r_search = rs[i];
        epj[k].r_scale_next = 1.0;
#ifdef PETAR_HIP_CHANGEOVER
        const long t = i % 50;
        if (t < 5) {  // member: the c.m. carries its mass
            epj[k].group_data.d[0] = mass[i];
            epj[k].group_data.d[1] = -(double)(i + 1);
            epj[k].mass = 0.0;
        } else if (t < 7) {  // member of a group without artificial particles
            epj[k].group_data.d[0] = mass[i];
            epj[k].group_data.d[1] = -3.4028234663852886e38 / 16;
        } else if (t < 10) {  // artificial particle
            epj[k].group_data.d[1] = 2.0 + (double)(t - 7);
        }
        epj_group[2 * k] = epj[k].group_data.d[0];
        epj_group[2 * k + 1] = epj[k].group_data.d[1];
#endif
        for (int d = 0; d < 3; d++) epj_pos[3 * k + d] = pos[3 * i + d];
        epj_mass[k] = epj[k].mass;
        epj_rs[k] = rs[i];
        epj_rin[k] = epj[k].r_in;
        epj_rout[k] = epj[k].r_out;
        epj_id[k] = epj[k].id;
    }
    std::vector<SPJSoft> spj(n_node);
    for (long s = 0; s < n_node; s++) {
        spj[s].mass = smass[s];
        spj[s].pos = PS::F64vec(scom[3 * s], scom[3 * s + 1], scom[3 * s + 2]);
#if PETAR_HIP_SPJ_QUAD
        const double *q = &squad[6 * s];
        spj[s].quad = PS::F64mat{q[0], q[1], q[2], q[3], q[4], q[5]};
#endif
    }
    std::vector<std::vector<EPISoft>> epi(n_group);
    std::vector<std::vector<ForceSoft>> force(n_group);
    std::vector<long> eoff(n_group + 1, 0), soff(n_group + 1, 0);
    for (long g = 0; g < n_group; g++) {
        eoff[g + 1] = eoff[g] + n_epj[g];
        soff[g + 1] = soff[g] + n_spj[g];
        epi[g].resize(n_epi[g]);
        force[g].resize(n_epi[g]);
        for (int i = 0; i < n_epi[g]; i++) {
            const EPJSoft &p = epj[epi_begin[g] + i];
            epi[g][i].id = p.id;
            epi[g][i].pos = p.pos;
            epi[g][i].r_search = p.r_search;
            epi[g][i].rank_org = 0;
            epi[g][i].type = 1;
            force[g][i].pot = 99.0;  // retrieve must overwrite
#ifdef PETAR_HIP_CHANGEOVER
            force[g][i].pot_tot = 99.0;
#endif
        }
    }

    // FDPS multi-walk call sequence
    const PS::F64 eps2 = 0.0, rout2 = r_out * r_out, G = 1.0;
#ifdef PETAR_HIP_REPLAY_COPIES
    // a build without PARTICLE_SIMULATOR_GPU_MULIT_WALK_INDEX: FDPS hands each
    // walk copies of its EPJ and SPJ records (calcForceAllAndWriteBackMultiWalk)
    std::vector<std::vector<EPJSoft>> wepj(n_group);
    std::vector<std::vector<SPJSoft>> wspj(n_group);
    for (long g = 0; g < n_group; g++) {
        for (int j = 0; j < n_epj[g]; j++) wepj[g].push_back(epj[id_epj[eoff[g] + j]]);
        for (int j = 0; j < n_spj[g]; j++) wspj[g].push_back(spj[id_spj[soff[g] + j]]);
    }
    auto sequence = [&]() {
        for (long g0 = 0; g0 < n_group; g0 += n_walk_limit) {
            const int nw = (int)std::min<long>(n_walk_limit, n_group - g0);
            std::vector<const EPISoft *> pe(nw);
            std::vector<const EPJSoft *> pj(nw);
            std::vector<const SPJSoft *> ps(nw);
            std::vector<ForceSoft *> pf(nw);
            for (int w = 0; w < nw; w++) {
                pe[w] = epi[g0 + w].data();
                pj[w] = wepj[g0 + w].data();
                ps[w] = wspj[g0 + w].data();
                pf[w] = force[g0 + w].data();
            }
            CalcForceWithLinearCutoffCUDA fw(0, eps2, rout2, G);
            fw(0, nw, pe.data(), n_epi.data() + g0, pj.data(), n_epj.data() + g0, ps.data(), n_spj.data() + g0);
            RetrieveForceCUDA(0, nw, n_epi.data() + g0, pf.data());
        }
    };
#else
    auto sequence = [&]() {
        CalcForceWithLinearCutoffCUDAMultiWalk f(0, eps2, rout2, G);
        f(0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, epj.data(), (PS::S32)n, spj.data(),
          (PS::S32)n_node, true);
        for (long g0 = 0; g0 < n_group; g0 += n_walk_limit) {
            const int nw = (int)std::min<long>(n_walk_limit, n_group - g0);
            std::vector<const EPISoft *> pe(nw);
            std::vector<const PS::S32 *> pj(nw), ps(nw);
            std::vector<ForceSoft *> pf(nw);
            for (int w = 0; w < nw; w++) {
                pe[w] = epi[g0 + w].data();
                pj[w] = id_epj.data() + eoff[g0 + w];
                ps[w] = id_spj.data() + soff[g0 + w];
                pf[w] = force[g0 + w].data();
            }
            CalcForceWithLinearCutoffCUDAMultiWalk fw(0, eps2, rout2, G);
            fw(0, nw, pe.data(), n_epi.data() + g0, pj.data(), n_epj.data() + g0, ps.data(), n_spj.data() + g0,
               epj.data(), (PS::S32)n, spj.data(), (PS::S32)n_node, false);
            RetrieveForceCUDA(0, nw, n_epi.data() + g0, pf.data());
        }
    };
#endif
    sequence();
    double inter = 0.0;
    for (long g = 0; g < n_group; g++) inter += (double)n_epi[g] * ((double)n_epj[g] + (double)n_spj[g]);
    for (int k = 0; k < timed_passes; k++) {
        const auto t0 = std::chrono::steady_clock::now();
        sequence();
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::printf("fdps_replay: pass %d: %.3f ms, %.2f Gint/s (%.4g interactions, %ld walks in batches of %d)\n",
                    k, ms, inter / (ms * 1e6), inter, n_group, n_walk_limit);
        std::fflush(stdout);
    }

#ifdef GPU_PROFILE
    // what PeTar prints: names, then values per pass (src/petar.hpp:2123-2130)
    std::printf("fdps_replay: gpu_counter n_walk %lld n_epi %lld n_epj %lld n_spj %lld n_call %lld\n",
                gpu_counter.n_walk.n, gpu_counter.n_epi.n, gpu_counter.n_epj.n, gpu_counter.n_spj.n,
                gpu_counter.n_call.n);
    std::printf("fdps_replay: gpu_profile copy %.6e send %.6e receive %.6e calc_force %.6e\n", gpu_profile.copy.time,
                gpu_profile.send.time, gpu_profile.recv.time, gpu_profile.calc.time);
    gpu_profile.dumpName(std::cout);
    gpu_counter.dumpName(std::cout);
    std::cout << std::endl;
    gpu_profile.dump(std::cout, 1 + timed_passes);
    gpu_counter.dump(std::cout, 1 + timed_passes);
    std::cout << std::endl;
    gpu_profile.clear();
    gpu_counter.clear();
    if (gpu_profile.calc.time != 0.0 || gpu_counter.n_call.n != 0) die("clear");
#endif

    std::vector<double> acc, pot, pot_tot;
    std::vector<long long> nngb;
    for (long g = 0; g < n_group; g++)
        for (auto &fr : force[g]) {
            acc.push_back(fr.acc.x);
            acc.push_back(fr.acc.y);
            acc.push_back(fr.acc.z);
            pot.push_back(fr.pot);
            nngb.push_back(fr.n_ngb);
#ifdef PETAR_HIP_CHANGEOVER
            pot_tot.push_back(fr.pot_tot);
#else
            pot_tot.push_back(fr.pot);
#endif
        }
    FILE *out = std::fopen(argv[6], "wb");
    if (!out) die("open output");
    std::vector<long long> hdr{n, n_node, n_group, n_le, n_ls};
    put(out, hdr);
    put(out, epj_pos);
    put(out, epj_mass);
    put(out, epj_rs);
    put(out, scom);
    put(out, smass);
    put(out, squad);
    put(out, epi_begin);
    put(out, n_epi);
    put(out, n_epj);
    put(out, n_spj);
    put(out, id_epj);
    put(out, id_spj);
    put(out, acc);
    put(out, pot);
    put(out, nngb);
    put(out, epj_id);
    put(out, epj_rin);
    put(out, epj_rout);
    put(out, pot_tot);
    put(out, epj_group);
    std::fclose(out);
    std::printf("fdps_replay: N=%ld walks=%ld batches=%ld written %s\n", n, n_group,
                (n_group + n_walk_limit - 1) / n_walk_limit, argv[6]);
    return 0;
}
