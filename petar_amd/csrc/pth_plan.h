// pth_plan.h -- host-side planning of one batch of walks: prefix sums,
// staged-list offsets, work items for the fused form and the reduction
// items of split walks.  Pure host code (no HIP calls), shared by the
// context (pth_context.cpp) and the diagnostic entry point pth_plan_debug.
#pragma once
#include <vector>

#include "pth_device.h"

namespace pth {

constexpr int XCDS = 8;             // MI355X: 8 XCDs, workgroup b runs on XCD b % 8
constexpr int ITEMS_PER_BLOCK = 4;  // one item per wave, 256-thread workgroups
// Measured on the precision workload (N = 1M, 7,110 walks): the resident
// autotuner picks 1024, and FDPS's 200-walk dispatch batches run 40 ms per
// step at 512 or 1024 against 49 ms at 2048 (one batch at 2048 holds fewer
// waves than the chip has slots, so each launch is mostly its tail).
constexpr int DEFAULT_CHUNK_J = 1024;

// Issue cost of one pair per lane, in units of one full-rate (4-cycle)
// wave64 VALU instruction, for a tile holding ipl i-particles per lane:
// EPJ (mix of near and far blocks) and SPJ, plus the per-record overhead
// shared by the ipl i-particles of a lane.
inline double epj_unit_cost(int ipl) { return 9.5 + 2.25 / ipl; }
inline double spj_unit_cost(int ipl) { return 18.5 + 5.5 / ipl; }

// A contiguous range of walks planned on its own (items of each class
// arranged over the XCDs within the segment), so segments can run
// concurrently on different streams.
struct PlanSeg {
    int walk_begin = 0, walk_end = 0;
    int seg_begin[4] = {0, 0, 0, 0};  // items of ipl 1, 2, 3, 4
    int seg_count[4] = {0, 0, 0, 0};
    int ritem_begin = 0, ritem_count = 0;
};
constexpr int MAX_SEGMENTS = PTH_MAX_SEGMENTS;
// Fewest pair interactions a pipelined segment of a resident evaluation is
// given by default (smaller evaluations get fewer segments, down to one).
// Measured: the 16,384-particle cluster (1.1e8 interactions) 0.18 ms per
// evaluation on one stream against 0.58 ms over 8 segments on 3 streams; the
// 1,048,576-particle one (1.6e10) 10.9 ms over 8 against 11.6 ms on one.
constexpr long long MIN_SEGMENT_PAIRS = 1000000000ll;

struct Plan {
    std::vector<WalkDev> walks;
    std::vector<long long> es_off, ss_off;  // n_walk + 1
    std::vector<Item> items;                // launch order, empty slots have walk < 0
    int seg_begin[4] = {0, 0, 0, 0};        // items of ipl 1, 2, 3, 4: one launch each
    int seg_count[4] = {0, 0, 0, 0};
    std::vector<RItem> ritems;
    int n_part = 0;                         // partial slots of 64 entries
    int n_real_items = 0;
    long long ni_tot = 0, n_e_slots = 0, n_s_rec = 0;
    long long n_epep = 0, n_epsp = 0;       // sum n_i n_e, sum n_i n_s
    bool mono = false;                      // monopole SPJ staged as far EPJ entries
    std::vector<PlanSeg> segs;              // at least one
    // empties the plan, keeping the vectors' memory for the next one
    void reset()
    {
        walks.clear();
        es_off.clear();
        ss_off.clear();
        items.clear();
        ritems.clear();
        segs.clear();
        for (int k = 0; k < PLAN_CLASSES; k++) seg_begin[k] = seg_count[k] = 0;
        n_part = n_real_items = 0;
        ni_tot = n_e_slots = n_s_rec = n_epep = n_epsp = 0;
        mono = false;
    }
};

// The parameters a plan was made with (exec_mode, chunk_j, max_ipl, n_seg,
// monopole superparticles, fewest interactions per segment).
struct PlanKey {
    int exec_mode = 0, chunk_j = 0, max_ipl = 0, n_seg = 1, mono = 0;
    long long min_seg_pairs = 0;
    bool operator==(const PlanKey &o) const
    {
        return exec_mode == o.exec_mode && chunk_j == o.chunk_j && max_ipl == o.max_ipl && n_seg == o.n_seg &&
               mono == o.mono && min_seg_pairs == o.min_seg_pairs;
    }
};

// Staged EPJ slots of a walk: near and far regions with at least 8 slots of
// padding between them.
inline long long epj_region_slots(long long n_e) { return n_e > 0 ? (n_e + 7) / 8 * 8 + 8 : 0; }

// Builds the plan; exec_mode PTH_EXEC_FUSED or PTH_EXEC_PLAIN, chunk_j the
// largest list length (EPJ entries + 2 x SPJ entries) of one item (<= 0:
// default), max_ipl the largest number of i per lane (1 to 4), n_seg the
// number of walk segments of about equal cost (fused form; the items of a
// walk do not depend on it; with min_seg_pairs > 0, at most one segment per
// min_seg_pairs interactions).  With mono (monopole superparticles) a walk's
// SPJ entries are staged as far EPJ entries after its EPJ list (a monopole
// pair is a far EPJ pair: no cutoff, no neighbour count), so its staged EPJ
// region holds n_epj + n_spj entries and it has no staged SPJ records.
// Returns false if a size overflows the int32 ranges of the device records.
bool make_plan(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj, int exec_mode,
               int chunk_j, int max_ipl, Plan &P, int n_seg = 1, bool mono = false, long long min_seg_pairs = 0);

}  // namespace pth
