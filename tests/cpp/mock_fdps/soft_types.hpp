// The soft-force particle types with the member sets and order of PeTar's
// src/soft_ptcl.hpp (ForceSoft :4-24, EPISoft :271-308, EPJSoft :311-376;
// no KDKDK_4TH / SAVE_NEIGHBOR_ID_IN_FORCE_KERNEL), for the adapter test.
#pragma once
#include "particle_simulator.hpp"
struct ForceSoft {
    PS::F64vec acc;
    PS::F64 pot;
    PS::S64 n_ngb;
#ifdef PETAR_HIP_CHANGEOVER
    PS::F64 pot_tot;  // the member a changeover-mode PeTar build adds
#endif
};
struct EPISoft {
    PS::S64 id;
    PS::F64vec pos;
    PS::F64 r_search;
    PS::S32 rank_org;
    PS::S32 type;
};
struct GroupDataDeliver {
    PS::F64 d[2];
};
struct EPJSoft {
    PS::S64 id;
    PS::F64 mass;
    PS::F64vec pos;
    PS::F64vec vel;
    PS::F64 r_in;
    PS::F64 r_out;
    PS::F64 r_search;
    PS::F64 r_scale_next;
    GroupDataDeliver group_data;
    PS::S32 rank_org;
    PS::S32 adr_org;
};
#ifdef PETAR_HIP_CHANGEOVER
static_assert(sizeof(ForceSoft) == 48, "ForceSoft + pot_tot");
#else
static_assert(sizeof(ForceSoft) == 40, "layout of src/soft_ptcl.hpp");
#endif
static_assert(sizeof(EPISoft) == 48 && sizeof(EPJSoft) == 120, "layouts of src/soft_ptcl.hpp");
