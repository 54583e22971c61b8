// pth_plan.h -- host-side planning of one batch of walks: prefix sums,
// staged-list offsets, work items for the fused form and the reduction
// items of split walks.  Pure host code (no HIP calls), shared by the
// context (pth_context.cpp) and the diagnostic entry point pth_plan_debug.
#pragma once
#include <vector>

#include "pth_device.h"

namespace pth {

// Relative cost of one staged SPJ record against one EPJ slot (VALU
// instructions per pair of the two inner loops, 33 vs 18, rounded).
constexpr int SPJ_COST = 2;
constexpr int XCDS = 8;          // MI355X: 8 XCDs, workgroup b runs on XCD b % 8
constexpr int ITEMS_PER_BLOCK = 4;  // one item per wave, 256-thread workgroups
constexpr int DEFAULT_CHUNK_J = 2048;

struct Plan {
    std::vector<WalkDev> walks;
    std::vector<long long> es_off, ss_off;  // n_walk + 1
    std::vector<Item> items;                // physical (launch) order, dummies have walk < 0
    std::vector<RItem> ritems;
    int n_part = 0;
    int n_real_items = 0;
    long long ni_tot = 0, n_e_slots = 0, n_s_rec = 0;
    long long n_epep = 0, n_epsp = 0;       // sum n_i n_e, sum n_i n_s
};

// Builds the plan; exec_mode PTH_EXEC_FUSED or PTH_EXEC_PLAIN, chunk_j the
// largest cost (in EPJ-slot units) of one item (<= 0: default).
// Returns false if a size overflows the int32 ranges of the device records.
bool make_plan(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj, int exec_mode,
               int chunk_j, Plan &P);

}  // namespace pth
