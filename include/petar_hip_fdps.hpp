// petar_hip_fdps.hpp -- header-only drop-in for PeTar's GPU soft-force hooks.
//
// Replaces src/force_gpu_cuda.hpp (declarations at :101-133 and :165-168)
// for the multi-walk-index build (PARTICLE_SIMULATOR_GPU_MULIT_WALK_INDEX,
// always set by Makefile.in:175).  Include it where PeTar includes
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
#pragma once
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "petar_hip.h"

#ifndef SPJSoft
#define SPJSoft PS::SPJQuadrupoleInAndOut
#endif

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

inline pth_ctx *context(PS::S32 rank, PS::F64 eps2, PS::F64 rcut2, PS::F64 G)
{
    pth_ctx *&c = context_slot();
    pth_params p;
    pth_params_default(&p);
    p.eps2 = eps2;
    p.rcut2 = rcut2;
    p.G = G;
#ifdef PETAR_HIP_CHANGEOVER
    p.changeover = 1;
#endif
#ifdef PETAR_HIP_CHUNK_J
    p.chunk_j = PETAR_HIP_CHUNK_J;  // list entries per work item (0: the library default)
#endif
    if (!c) {
        if (pth_create(&c, rank, &p) != PTH_OK) fail("pth_create", nullptr);
    } else if (pth_set_params(c, &p) != PTH_OK) {
        fail("pth_set_params", c);
    }
    return c;
}

inline pth_epj_view epj_view()
{
    return pth_epj_view{sizeof(EPJSoft), offsetof(EPJSoft, pos), offsetof(EPJSoft, mass),
                        offsetof(EPJSoft, r_search), offsetof(EPJSoft, id), offsetof(EPJSoft, r_in),
                        offsetof(EPJSoft, r_out)};
}

inline pth_spj_view spj_view()
{
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
}

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

// The retrieve function (src/force_gpu_cuda.hpp:165-168).
inline PS::S32 RetrieveForceHIP(const PS::S32 tag, const PS::S32 n_walk, const PS::S32 *ni,
                                ForceSoft **force)
{
    pth_ctx *c = petar_hip::context_slot();
    if (!c) petar_hip::fail("RetrieveForceHIP (no dispatch yet)", nullptr);
    const pth_force_view fv = petar_hip::force_view();
    if (pth_retrieve(c, tag, n_walk, ni, reinterpret_cast<void *const *>(force), &fv) != PTH_OK)
        petar_hip::fail("pth_retrieve", c);
    return 0;
}

#ifndef PETAR_HIP_NO_REFERENCE_NAMES
// The names src/petar.hpp:894-899 uses.
typedef CalcForceWithLinearCutoffHIPMultiWalk CalcForceWithLinearCutoffCUDAMultiWalk;
inline PS::S32 RetrieveForceCUDA(const PS::S32 tag, const PS::S32 n_walk, const PS::S32 *ni,
                                 ForceSoft **force)
{
    return RetrieveForceHIP(tag, n_walk, ni, force);
}
#endif
