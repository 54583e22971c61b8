// petar_hip_fdps.hpp -- header-only drop-in for PeTar's GPU soft-force hooks.
//
// Replaces src/force_gpu_cuda.hpp (declarations at :101-133, :137-162 and
// :165-168): the multi-walk-index functor (PARTICLE_SIMULATOR_GPU_MULIT_WALK_INDEX,
// always set by Makefile.in:175) and the functor of a build without it.  Include it where PeTar includes
// force_gpu_cuda.hpp, after the FDPS header and soft_ptcl.hpp (it needs
// PS::S32/F64, EPISoft, EPJSoft, ForceSoft and PS::SPJQuadrupoleInAndOut),
// and link libpetar_hip.so.  The call at src/petar.hpp:894-899,
//
//   tree_soft.calcForceAllAndWriteBackMultiWalkIndex(
//       CalcForceWithLinearCutoffCUDAMultiWalk(my_rank, eps2, rout2, G),
//       RetrieveForceCUDA, tag_max, system_soft, dinfo, n_walk_limit);
//
// compiles unchanged: the two names are provided as aliases of the HIP
// versions below.  Semantics kept from the reference: one process-wide
// device context (device = my_rank % number of devices) created at the first
// call, all EPJ/SPJ sent with send_flag == true, one batch of walks per
// dispatch, retrieve assigns acc/pot/n_ngb.  The reference asserts on bad
// input and ignores device errors; here every failure stops the program with
// the library's error text (PeTar's fail-stop behaviour).
//
// With PETAR_HIP_CHANGEOVER defined, the context runs in changeover mode
// (see petar_hip.h): the kernel also applies the changeover weight to the
// neighbour pairs, which replaces PeTar's CPU correction pass
// (treeForceCorrectChangeover, src/petar.hpp:965-989); ForceSoft must then
// have a PS::F64 pot_tot member, and pot receives the soft potential.
//
// As in src/force_gpu_cuda.hpp:95-99, USE_QUAD (set by PeTar's default
// configuration, Makefile.in:28-30) selects quadrupole superparticles
// (SPJSoft = PS::SPJQuadrupoleInAndOut); without it (configure
// --disable-quad) they are monopoles (PS::SPJMonopoleInAndOut: mass and
// centre) and the context runs with spj_monopole = 1.
//
// With GPU_PROFILE defined (every reference GPU build sets it,
// Makefile.in:181) the adapter provides the reference's profile objects,
// gpu_profile (copy, send, receive, calc_force times) and gpu_counter
// (n_walk, n_epi, n_epj, n_spj, n_call), with the members and methods
// src/petar.hpp:2081-2155,2643-2645 uses (src/force_gpu_cuda.hpp:6-93); they
// are built from PeTar's Tprofile / NumCounter (profile.hpp) and filled from
// the library's pth_profile at every retrieve.
#pragma once
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "petar_hip.h"

#ifdef USE_QUAD
#define PETAR_HIP_SPJ_QUAD 1
#else
#define PETAR_HIP_SPJ_QUAD 0
#endif
#ifndef SPJSoft
#if PETAR_HIP_SPJ_QUAD
#define SPJSoft PS::SPJQuadrupoleInAndOut
#else
#define SPJSoft PS::SPJMonopoleInAndOut
#endif
#endif

#ifdef GPU_PROFILE
#include <initializer_list>
#include <ostream>

#include "profile.hpp"

// Times of the soft-force hooks (src/force_gpu_cuda.hpp:6-53): columns
// copy, send, receive, calc_force, in that order (the order PeTar's profile
// readers expect, tools/analysis/profile.py:70-84).
struct GPUProfile {
    Tprofile copy;
    Tprofile send;
    Tprofile recv;
    Tprofile calc;
    const PS::S32 n_profile;
    GPUProfile()
        : copy("copy       "), send("send       "), recv("receive    "), calc("calc_force "), n_profile(4) {}
    void print(std::ostream &fout, const PS::F64 time_sys, const PS::S64 n_loop = 1)
    {
        fout << "Time: " << time_sys << std::endl;
        for (Tprofile *p : {&copy, &send, &recv, &calc}) p->print(fout, n_loop);
    }
    void dump(std::ostream &fout, const PS::S64 n_loop = 1, const PS::S32 width = PROFILE_PRINT_WIDTH) const
    {
        for (const Tprofile *p : {&copy, &send, &recv, &calc}) p->dump(fout, n_loop, width);
    }
    void dumpName(std::ostream &fout, const PS::S32 width = PROFILE_PRINT_WIDTH)
    {
        for (const Tprofile *p : {&copy, &send, &recv, &calc}) p->dumpName(fout, width);
    }
    void clear()
    {
        for (Tprofile *p : {&copy, &send, &recv, &calc}) p->reset();
    }
};

// Counts of the dispatches (src/force_gpu_cuda.hpp:55-92).
struct GPUCounter {
    NumCounter n_walk;
    NumCounter n_epi;
    NumCounter n_epj;
    NumCounter n_spj;
    NumCounter n_call;
    const PS::S32 n_counter;
    GPUCounter()
        : n_walk("n_walk "), n_epi("n_epi  "), n_epj("n_epj  "), n_spj("n_spj  "), n_call("n_call "), n_counter(5) {}
    void dump(std::ostream &fout, const PS::S64 n_loop = 1, const PS::S32 width = PROFILE_PRINT_WIDTH) const
    {
        for (const NumCounter *p : {&n_walk, &n_epi, &n_epj, &n_spj, &n_call}) p->dump(fout, n_loop, width);
    }
    void dumpName(std::ostream &fout, const PS::S32 width = PROFILE_PRINT_WIDTH) const
    {
        for (const NumCounter *p : {&n_walk, &n_epi, &n_epj, &n_spj, &n_call}) p->dumpName(fout, width);
    }
    void clear()
    {
        for (NumCounter *p : {&n_walk, &n_epi, &n_epj, &n_spj, &n_call}) *p = 0;
    }
};

// One object of each per program (the reference defines them in
// force_gpu_cuda.cu:7-9; a header-only adapter defines them here).
#if __cplusplus >= 201703L
inline GPUProfile gpu_profile;
inline GPUCounter gpu_counter;
#else
namespace petar_hip {
template <class T> struct global_object {
    static T value;
};
template <class T> T global_object<T>::value;
}  // namespace petar_hip
static GPUProfile &gpu_profile __attribute__((unused)) = petar_hip::global_object<GPUProfile>::value;
static GPUCounter &gpu_counter __attribute__((unused)) = petar_hip::global_object<GPUCounter>::value;
#endif
#endif  // GPU_PROFILE

namespace petar_hip {

inline pth_ctx *&context_slot()
{
    static pth_ctx *ctx = nullptr;
    return ctx;
}

[[noreturn]] inline void fail(const char *where, pth_ctx *ctx)
{
    std::fprintf(stderr, "petar_hip: %s failed: %s\n", where, ctx ? pth_last_error(ctx) : pth_create_error());
    std::abort();
}

// The process-wide context, created at the first call with the library's
// defaults and the compile-time choices below; later calls update only the
// functor's own parameters (eps2, rcut2, G), so fields a caller set on
// context_slot() with pth_set_params stay as they were.
inline pth_ctx *context(PS::S32 rank, PS::F64 eps2, PS::F64 rcut2, PS::F64 G)
{
    pth_ctx *&c = context_slot();
    pth_params p;
    if (!c) {
        pth_params_default(&p);
#ifdef PETAR_HIP_CHANGEOVER
        p.changeover = 1;
#endif
#ifdef PETAR_HIP_CHUNK_J
        // list entries per work item (0: the library default); applied when
        // the context is created
        p.chunk_j = PETAR_HIP_CHUNK_J;
#endif
        p.spj_monopole = PETAR_HIP_SPJ_QUAD ? 0 : 1;
        p.eps2 = eps2;
        p.rcut2 = rcut2;
        p.G = G;
        if (pth_create(&c, rank, &p) != PTH_OK) fail("pth_create", nullptr);
#ifdef GPU_PROFILE
        if (pth_profile_enable(c, 1) != PTH_OK) fail("pth_profile_enable", c);
#endif
        return c;
    }
    if (pth_get_params(c, &p) != PTH_OK) fail("pth_get_params", c);
    if (p.eps2 != eps2 || p.rcut2 != rcut2 || p.G != G) {
        p.eps2 = eps2;
        p.rcut2 = rcut2;
        p.G = G;
        if (pth_set_params(c, &p) != PTH_OK) fail("pth_set_params", c);
    }
    return c;
}

inline pth_epj_view epj_view()
{
    // group_data.artificial: mass_backup then status (src/artificial_particles.hpp:20-23)
    return pth_epj_view{sizeof(EPJSoft),          offsetof(EPJSoft, pos), offsetof(EPJSoft, mass),
                        offsetof(EPJSoft, r_search), offsetof(EPJSoft, id), offsetof(EPJSoft, r_in),
                        offsetof(EPJSoft, r_out),    (long long)offsetof(EPJSoft, group_data)};
}

inline pth_spj_view spj_view()
{
#if PETAR_HIP_SPJ_QUAD
    const size_t q = offsetof(SPJSoft, quad);
    return pth_spj_view{sizeof(SPJSoft),
                        offsetof(SPJSoft, pos),
                        offsetof(SPJSoft, mass),
                        q + offsetof(PS::F64mat, xx),
                        q + offsetof(PS::F64mat, yy),
                        q + offsetof(PS::F64mat, zz),
                        q + offsetof(PS::F64mat, xy),
                        q + offsetof(PS::F64mat, xz),
                        q + offsetof(PS::F64mat, yz)};
#else
    // monopole superparticles: mass and centre only (spj_monopole = 1)
    return pth_spj_view{sizeof(SPJSoft), offsetof(SPJSoft, pos), offsetof(SPJSoft, mass), 0, 0, 0, 0, 0, 0};
#endif
}

#ifdef GPU_PROFILE
// Adds what the library measured since the last call to gpu_profile and
// gpu_counter.
inline void profile_drain(pth_ctx *c)
{
    pth_profile p;
    if (pth_profile_take(c, &p) != PTH_OK) fail("pth_profile_take", c);
    gpu_profile.copy.time += p.copy_s;
    gpu_profile.send.time += p.send_s;
    gpu_profile.recv.time += p.recv_s;
    gpu_profile.calc.time += p.calc_s;
    gpu_counter.n_walk += p.n_walk;
    gpu_counter.n_epi += p.n_epi;
    gpu_counter.n_epj += p.n_epj;
    gpu_counter.n_spj += p.n_spj;
    gpu_counter.n_call += p.n_call;
}
#endif

inline pth_epi_view epi_view()
{
    return pth_epi_view{sizeof(EPISoft), offsetof(EPISoft, pos), offsetof(EPISoft, r_search),
                        offsetof(EPISoft, id)};
}

inline pth_force_view force_view()
{
#ifdef PETAR_HIP_CHANGEOVER
    // ForceSoft then carries a pot_tot member (pot receives pot_soft)
    const long long pot_tot = (long long)offsetof(ForceSoft, pot_tot);
#else
    const long long pot_tot = -1;
#endif
    return pth_force_view{sizeof(ForceSoft), offsetof(ForceSoft, acc), offsetof(ForceSoft, pot),
                          offsetof(ForceSoft, n_ngb), (int)sizeof(((ForceSoft *)nullptr)->n_ngb), pot_tot};
}

}  // namespace petar_hip

// The dispatch functor (src/force_gpu_cuda.hpp:101-133).
struct CalcForceWithLinearCutoffHIPMultiWalk {
    PS::S32 my_rank;
    PS::F64 eps2;
    PS::F64 rcut2;
    PS::F64 G;

    CalcForceWithLinearCutoffHIPMultiWalk() {}
    CalcForceWithLinearCutoffHIPMultiWalk(PS::S32 _rank, PS::F64 _eps2, PS::F64 _rcut2, PS::F64 _G)
        : my_rank(_rank), eps2(_eps2), rcut2(_rcut2), G(_G) {}

    void initialize(PS::S32 _rank, PS::F64 _eps2, PS::F64 _rcut2, PS::F64 _G)
    {
        my_rank = _rank;
        eps2 = _eps2;
        rcut2 = _rcut2;
        G = _G;
    }

    PS::S32 operator()(const PS::S32 tag, const PS::S32 n_walk, const EPISoft **epi,
                       const PS::S32 *n_epi, const PS::S32 **id_epj, const PS::S32 *n_epj,
                       const PS::S32 **id_spj, const PS::S32 *n_spj, const EPJSoft *epj,
                       const PS::S32 n_epj_tot, const SPJSoft *spj, const PS::S32 n_spj_tot,
                       const bool send_flag)
    {
        pth_ctx *c = petar_hip::context(my_rank, eps2, rcut2, G);
        if (send_flag) {
            const pth_epj_view ev = petar_hip::epj_view();
            const pth_spj_view sv = petar_hip::spj_view();
            if (pth_send_j(c, epj, n_epj_tot, &ev, spj, n_spj_tot, &sv) != PTH_OK)
                petar_hip::fail("pth_send_j", c);
            return 0;
        }
        const pth_epi_view iv = petar_hip::epi_view();
        if (pth_dispatch(c, tag, n_walk, reinterpret_cast<const void *const *>(epi), n_epi, &iv,
                         id_epj, n_epj, id_spj, n_spj) != PTH_OK)
            petar_hip::fail("pth_dispatch", c);
        return 0;
    }
};

// The dispatch functor of a build without PARTICLE_SIMULATOR_GPU_MULIT_WALK_INDEX
// (src/force_gpu_cuda.hpp:137-162, called by FDPS's
// calcForceAllAndWriteBackMultiWalk, src/petar.hpp:901): each walk comes
// with its own copies of its EPJ and SPJ records.
struct CalcForceWithLinearCutoffHIP {
    PS::S32 my_rank;
    PS::F64 eps2;
    PS::F64 rcut2;
    PS::F64 G;

    CalcForceWithLinearCutoffHIP() {}
    CalcForceWithLinearCutoffHIP(PS::S32 _rank, PS::F64 _eps2, PS::F64 _rcut2, PS::F64 _G)
        : my_rank(_rank), eps2(_eps2), rcut2(_rcut2), G(_G) {}

    void initialize(PS::S32 _rank, PS::F64 _eps2, PS::F64 _rcut2, PS::F64 _G)
    {
        my_rank = _rank;
        eps2 = _eps2;
        rcut2 = _rcut2;
        G = _G;
    }

    PS::S32 operator()(const PS::S32 tag, const PS::S32 n_walk, const EPISoft *epi[], const PS::S32 n_epi[],
                       const EPJSoft *epj[], const PS::S32 n_epj[], const SPJSoft *spj[], const PS::S32 n_spj[])
    {
        pth_ctx *c = petar_hip::context(my_rank, eps2, rcut2, G);
        const pth_epi_view iv = petar_hip::epi_view();
        const pth_epj_view ev = petar_hip::epj_view();
        const pth_spj_view sv = petar_hip::spj_view();
        if (pth_dispatch_copies(c, tag, n_walk, reinterpret_cast<const void *const *>(epi), n_epi, &iv,
                                reinterpret_cast<const void *const *>(epj), n_epj, &ev,
                                reinterpret_cast<const void *const *>(spj), n_spj, &sv) != PTH_OK)
            petar_hip::fail("pth_dispatch_copies", c);
        return 0;
    }
};

// The retrieve function (src/force_gpu_cuda.hpp:165-168).
inline PS::S32 RetrieveForceHIP(const PS::S32 tag, const PS::S32 n_walk, const PS::S32 *ni,
                                ForceSoft **force)
{
    pth_ctx *c = petar_hip::context_slot();
    if (!c) petar_hip::fail("RetrieveForceHIP (no dispatch yet)", nullptr);
    const pth_force_view fv = petar_hip::force_view();
    if (pth_retrieve(c, tag, n_walk, ni, reinterpret_cast<void *const *>(force), &fv) != PTH_OK)
        petar_hip::fail("pth_retrieve", c);
#ifdef GPU_PROFILE
    petar_hip::profile_drain(c);
#endif
    return 0;
}

#ifndef PETAR_HIP_NO_REFERENCE_NAMES
// The names src/petar.hpp:894-904 uses.
typedef CalcForceWithLinearCutoffHIPMultiWalk CalcForceWithLinearCutoffCUDAMultiWalk;
typedef CalcForceWithLinearCutoffHIP CalcForceWithLinearCutoffCUDA;
inline PS::S32 RetrieveForceCUDA(const PS::S32 tag, const PS::S32 n_walk, const PS::S32 *ni,
                                 ForceSoft **force)
{
    return RetrieveForceHIP(tag, n_walk, ni, force);
}
#endif
