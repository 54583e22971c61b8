// pth_device.h -- device record layouts and launch descriptors shared by the
// kernels (pth_kernels.hip) and the host side (pth_context.cpp, pth_plan.cpp).
#pragma once
#include <hip/hip_runtime.h>

// FP32 partial sums are flushed into FP64 every PTH_FLUSH_BLOCKS EPJ blocks
// of 8 (64 pairs) and every PTH_FLUSH_SPJ SPJ records.
#define PTH_FLUSH_BLOCKS 8
#define PTH_FLUSH_SPJ 32

namespace pth {

// Work items come in classes of 1, 2, 3 or 4 i-particles per lane, one
// launch per class (each kernel instance has its own register budget).
constexpr int PLAN_CLASSES = 4;

// Resident EPJ record (32 B): FP64 position, mass, r_search.
struct __align__(32) EpjDev {
    double x, y, z;
    float m, rs;
};

// Resident SPJ record (64 B): FP64 centre of mass, mass, raw second moment.
struct __align__(64) SpjDev {
    double x, y, z;
    float m, qxx, qyy, qzz, qxy, qxz, qyz;
    float pad[3];
};

// Walk-independent part of an SPJ term (32 B), one per superparticle of the
// force call: the mass, the traceless quadrupole q' = 3Q - tr I (xx, yy, zz,
// xy, yz, zx; the rearrangement of src/phantomquad_for_p3t_x86.hpp:144-163)
// and -eps^2 tr.  A staged SPJ record holds only the walk-relative position
// and the index of its entry here.
struct __align__(32) SpjTab {
    float m, qxx, qyy, qzz, qxy, qyz, qzx, mtr;
};

// Changeover data of an EPJ (24 B, changeover mode only): the particle's
// changeover radii (EPJSoft::r_in, r_out), its group type and backup mass
// (EPJSoft::group_data.artificial, src/artificial_particles.hpp:20-102) and
// id.
struct __align__(8) EpjCoDev {
    float rin, rout;
    float mback;  // member: the mass its c.m. carries (src/hard.hpp:1459-1464)
    int type;     // PTH_CO_SINGLE, _MEMBER, _MEMBER_NO_CM or _ARTIFICIAL
    long long id;
};
#define PTH_CO_SINGLE 0
#define PTH_CO_MEMBER 1        // group member with a c.m. address: zero tree mass
#define PTH_CO_MEMBER_NO_CM 2  // member of a group without artificial particles
#define PTH_CO_ARTIFICIAL 3    // artificial or c.m. particle (status > 0)

// Potential mass of a neighbour j in the changeover correction
// (src/hard.hpp:1453-1472): a single's mass, a member's backup mass, and
// zero for an artificial particle (its linear-cutoff term is cancelled and
// nothing added).
__host__ __device__ inline float co_pot_mass(const EpjCoDev &c, float m)
{
    return c.type == PTH_CO_SINGLE ? m : (c.type == PTH_CO_ARTIFICIAL ? 0.f : c.mback);
}
// Mass of the self term -G m / r_out an i loses in changeover mode: singles
// and members without a c.m. address (src/hard.hpp:1665-1673); the others
// keep it.
__host__ __device__ inline float co_self_mass(const EpjCoDev &c, float m)
{
    return (c.type == PTH_CO_SINGLE || c.type == PTH_CO_MEMBER_NO_CM) ? m : 0.f;
}

// Batch i-particle (32 B): FP64 position, r_search.
struct __align__(32) EpiDev {
    double x, y, z;
    float rs;
    int pad;
};

// Changeover mode: an i-group may hold at most this many i (the per-walk
// id table lives in LDS).
#define PTH_CO_MAX_NI 1024
#define PTH_CO_HASH 2048
// floats per staged EPJ block of 8 in the changeover extras:
// r_in[8], r_out[8], x, y, z low parts[8], potential mass[8]
#define PTH_CO_EXTRA 48
// floats per staged near EPJ block of 8 in linear-cutoff mode with
// double-float near pairs: the low-order parts x[8], y[8], z[8]
#define PTH_DF_EXTRA 24

// One walk (i-group) of a batch: i range, list ranges (in the concatenated
// id lists) and sizes.
struct WalkDev {
    int i_off, n_i;
    int e_off, n_e;
    int s_off, n_s;
    int pad[2];
};

// One work item (one wave): a tile of n_i <= 64*ipl i-particles of a walk
// (lane l holds i = i_begin + l + 64 k, k < ipl) against a block range of the
// walk's staged EPJ region and a record range of its staged SPJ.  walk < 0
// marks an empty slot.  part < 0: results go to the output, else to the
// partial slots part .. part+ipl-1 (64 entries each).
struct __align__(16) Item {
    int walk, i_begin, n_i_ipl, e_blk_begin, e_blk_end, s_begin, s_end, part;
};
__host__ __device__ inline int item_n_i(int v) { return v & 0xffff; }
__host__ __device__ inline int item_ipl(int v) { return v >> 16; }
__host__ __device__ inline int item_pack(int n_i, int ipl) { return n_i | (ipl << 16); }

// Sum of n_part chunks of partial slots into outputs [out, out + n_i):
// i = 64 k + lane of the tile is in slot part_first + c*ipl + k of chunk c.
struct RItem {
    int out, n_i, part_first, n_part, ipl, pad[3];
};

// One segment of a batch: its walks, its items per ipl class and its
// reduction items (the launches of one segment run in order on one stream).
struct SegDev {
    int walk_begin = 0, n_walk = 0;
    int seg_begin[4] = {0, 0, 0, 0}, seg_count[4] = {0, 0, 0, 0};  // items of ipl 1, 2, 3, 4
    int ritem_begin = 0, n_ritems = 0;
};
#define PTH_MAX_SEGMENTS 32

struct BatchDev {
    int n_seg = 1;
    SegDev seg[PTH_MAX_SEGMENTS];
    int n_walk = 0;
    long long ni_tot = 0;
    long long n_e_slots = 0;  // staged EPJ slots (per walk: 8*ceil(n_e/8) + 8)
    long long n_s_rec = 0;    // staged SPJ records
    int n_items = 0, n_ritems = 0;
    int plain = 0, nnb_le = 0, j_vgpr = 0;
    int co = 0;        // changeover mode: (1 - K) weight for neighbours, no self term
    int df = 0;        // linear-cutoff mode: near EPJ pairs in double-float coordinates
    int mono = 0;      // monopole superparticles: spj holds EpjDev records, staged as far EPJ
    float eps2 = 0.f, rout2 = 0.f;
    float ngb2 = 0.f;  // changeover mode: squared neighbour-radius factor
    double G = 1.0;
    // changeover mode only
    const long long *epi_id = nullptr;
    const EpjCoDev *epj_co = nullptr;
    float4 *staged_ic = nullptr;  // per i: {self slot (int bits), r_in, r_out, m}
    float4 *staged_il = nullptr;  // per i: low parts of x-c, y-c, z-c (double-float), 0 (co or df)
    float *staged_el = nullptr;   // df: per staged near EPJ block of 8: {x, y, z low parts[8]}
    float *staged_ec = nullptr;   // per staged EPJ block of 8: {r_in[8], r_out[8], x, y, z low parts[8], m_pot[8]}
    int *walk_err = nullptr;      // per walk: number of i without their own EPJ
    double *out_pott = nullptr;
    double *part_pott = nullptr;
    const EpiDev *epi = nullptr;
    const WalkDev *walks = nullptr;
    const long long *es_off = nullptr;  // n_walk+1 staged EPJ slot offsets
    const long long *ss_off = nullptr;  // n_walk+1 staged SPJ record offsets
    const int *id_epj = nullptr;
    const int *id_spj = nullptr;
    const EpjDev *epj = nullptr;
    const SpjDev *spj = nullptr;
    const SpjTab *spj_tab = nullptr;  // quadrupole superparticles: the table of spj
    const Item *items = nullptr;
    const RItem *ritems = nullptr;
    double4 *centre = nullptr;
    int *near_blk = nullptr;  // per walk: first block of the far region (absolute)
    float4 *staged_i = nullptr;
    float *staged_e = nullptr;
    float *staged_s = nullptr;  // per SPJ record {x-c, y-c, z-c, table index (int bits)}
    double4 *out_accp = nullptr;
    int *out_nnb = nullptr;
    double4 *part_accp = nullptr;
    int *part_nnb = nullptr;
    // persistent force launches: PTH_MAX_SEGMENTS x PLAN_CLASSES x 8 item
    // counters (per segment, i-per-lane class and XCD), zeroed by the
    // segment's staging launch
    int persistent = 0;
    int *counters = nullptr;
};

// The SPJ table of n quadrupole superparticles for softening eps2 (the
// batch's FP32 eps2, as the staging uses it).
hipError_t launch_spj_table(const SpjDev *spj, long long n, float eps2, SpjTab *tab, hipStream_t st);
// Floats per staged SPJ record.
#define PTH_SPJ_STAGED 4

// Expands index lists sent as runs of consecutive indices for walks
// [w_begin, w_begin + n_walk): for walk w, runs roff[w] .. roff[w + 1] - 1
// start at position rpos[k] of the walk's list with index rid[k]; the list
// of walk w (n entries from walks[w].e_off for EPJ, or walks[w].s_off for
// SPJ) goes to out_e / out_s.
hipError_t launch_expand_runs(const WalkDev *walks, int w_begin, int n_walk, const int *rid_e, const int *rpos_e,
                              const int *run_eo, const int *rid_s, const int *rpos_s, const int *run_so, int *out_e,
                              int *out_s, hipStream_t st);

// Launches of segment k of a batch (k < b.n_seg).
hipError_t launch_stage(const BatchDev &b, int k, hipStream_t st);
// classes: bit c set launches the items of i-per-lane class c + 1
hipError_t launch_force(const BatchDev &b, int k, hipStream_t st, unsigned classes = 0xFu);
hipError_t launch_reduce(const BatchDev &b, int k, hipStream_t st);
// All segments of a batch in order on one stream.
hipError_t launch_all(const BatchDev &b, hipStream_t st);

}  // namespace pth
