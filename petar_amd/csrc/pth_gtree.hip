// pth_gtree.hip -- the octree, its moments, the i-groups and the interaction
// lists built on the device (pth_gtree_* of include/petar_hip.h), with the
// semantics of the host tree (petar_amd/csrc/host/tree.cpp), which stands in
// for FDPS's TreeForForceLong<...>::QuadrupoleWithSymmetrySearch
// (src/petar.hpp:619-623, 2741) and its multi-walk lists
// (src/petar.hpp:894-899):
//   1. bounding cube (FP64 min/max), 63-bit Morton keys, stable radix sort of
//      (key, index);
//   2. nodes level by level: a node with more than n_leaf_limit particles is
//      split into its non-empty octants (binary search of the key ranges);
//      the nodes of a level follow their parents' order, octants in order;
//   3. moments bottom-up (mass, centre of mass, raw second moment about it,
//      particle box and box widened by r_search), each node by one thread in
//      the host's summation order;
//   4. i-groups: largest cells with at most n_group_limit particles (or
//      leaves), in Morton order;
//   5. walk: one wave per group traverses the tree with a stack in LDS,
//      popping up to 64 nodes at a time; a node is opened when it contains
//      the group, when the group box is closer to its centre of mass than
//      L/theta, or when the boxes overlap with either side widened by
//      r_search; opened leaves give their particles (EPJ), closed nodes give
//      one SPJ.  A counting pass, a scan, then a writing pass.
//   6. optionally (canonical order) a segmented sort of every list into the
//      host walk's depth-first order: EPJ ascending, SPJ by first particle.
// Every decision is the same FP64 expression as on the host (no FMA
// contraction in this file), so the lists are the host's lists as sets, and
// entry for entry after the canonical ordering.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <climits>
#include <cmath>

#include "pth_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int GT_MAX_LEVEL = 21;
constexpr int WALK_STACK = 2048;  // LDS node stack of one walking wave (8 KB: 20 waves per CU;
                                  // 4096 ran the precision workload's walk in 4.1 ms, 2048 and 2560
                                  // in 3.2 ms, 1536 in 4.7 ms, 1024 in 10.3 ms)

struct GNode {
    long long begin, end;
    int level, n_child;
    long long first_child;
};

__device__ __forceinline__ unsigned long long spread3(unsigned long long v)
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// per-block min/max of the coordinates, max of r_search (rs may be null) and
// a flag for non-finite values (which fmin/fmax would skip):
// part[block * 8 + {0..2: min, 3..5: max, 6: max rs, 7: 1 if any non-finite}]
__global__ __launch_bounds__(256) void k_bbox(const double *__restrict__ pos, const double *__restrict__ rs,
                                              long long n, double *__restrict__ part)
{
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rmax = 0.0, bad = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        for (int d = 0; d < 3; d++) {
            const double x = pos[3 * i + d];
            lo[d] = fmin(lo[d], x);
            hi[d] = fmax(hi[d], x);
            if (!isfinite(x)) bad = 1.0;
        }
        if (rs) {
            rmax = fmax(rmax, rs[i]);
            if (!isfinite(rs[i])) bad = 1.0;
        }
    }
    __shared__ double red[4][8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
        }
        rmax = fmax(rmax, __shfl_xor(rmax, off, 64));
        bad = fmax(bad, __shfl_xor(bad, off, 64));
    }
    if (lane == 0) {
        for (int d = 0; d < 3; d++) {
            red[wave][d] = lo[d];
            red[wave][3 + d] = hi[d];
        }
        red[wave][6] = rmax;
        red[wave][7] = bad;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int d = threadIdx.x;
        double v = red[0][d];
        for (int w = 1; w < 4; w++) v = (d < 3) ? fmin(v, red[w][d]) : fmax(v, red[w][d]);
        part[blockIdx.x * 8 + d] = v;
    }
}

__global__ __launch_bounds__(256) void k_keys(const double *__restrict__ pos, long long n, double r0, double r1,
                                              double r2, double scale, unsigned long long *__restrict__ key,
                                              long long *__restrict__ idx)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double r[3] = {r0, r1, r2};
    unsigned long long c[3];
    for (int d = 0; d < 3; d++) {
        const double f = (pos[3 * i + d] - r[d]) * scale;
        long long v = (long long)floor(f);
        v = min(max(v, 0LL), (long long)(1u << GT_MAX_LEVEL) - 1);
        c[d] = (unsigned long long)v;
    }
    key[i] = spread3(c[0]) << 2 | spread3(c[1]) << 1 | spread3(c[2]);
    idx[i] = i;
}

__global__ __launch_bounds__(256) void k_gather(long long n, const long long *__restrict__ order,
                                                const double *__restrict__ pos, const double *__restrict__ mass,
                                                const double *__restrict__ rs, double *__restrict__ spos,
                                                double *__restrict__ smass, double *__restrict__ srs)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const long long i = order[k];
    for (int d = 0; d < 3; d++) spos[3 * k + d] = pos[3 * i + d];
    smass[k] = mass[i];
    srs[k] = rs ? rs[i] : 0.0;
}

// octant boundaries and non-empty child count of nodes [lb, lb + nl)
__global__ __launch_bounds__(256) void k_split_count(const GNode *__restrict__ nodes, long long lb, long long nl,
                                                     const unsigned long long *__restrict__ key, int n_leaf_limit,
                                                     long long *__restrict__ bnd, int *__restrict__ cnt)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k == 0) cnt[nl] = 0;  // so that the scan's entry nl is the level's child count
    if (k >= nl) return;
    const GNode nd = nodes[lb + k];
    long long *b = bnd + 9 * k;
    const long long m = nd.end - nd.begin;
    if (m <= n_leaf_limit || nd.level >= GT_MAX_LEVEL) {
        cnt[k] = 0;
        return;
    }
    const int shift = 3 * (GT_MAX_LEVEL - 1 - nd.level);
    b[0] = nd.begin;
    b[8] = nd.end;
    for (int d = 1; d < 8; d++) {
        // first position in [begin, end) whose octant digit is >= d
        long long lo = nd.begin, hi = nd.end;
        while (lo < hi) {
            const long long mid = (lo + hi) >> 1;
            if ((int)((key[mid] >> shift) & 7u) < d) lo = mid + 1;
            else hi = mid;
        }
        b[d] = lo;
    }
    int c = 0;
    for (int d = 0; d < 8; d++) c += (b[d + 1] > b[d]);
    cnt[k] = c;
}

__global__ __launch_bounds__(256) void k_split_write(GNode *__restrict__ nodes, long long lb, long long nl,
                                                     long long le, const long long *__restrict__ bnd,
                                                     const int *__restrict__ cnt, const int *__restrict__ off)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= nl || cnt[k] == 0) return;
    GNode &nd = nodes[lb + k];
    nd.first_child = le + off[k];
    nd.n_child = cnt[k];
    const long long *b = bnd + 9 * k;
    long long c = nd.first_child;
    for (int d = 0; d < 8; d++)
        if (b[d + 1] > b[d]) nodes[c++] = GNode{b[d], b[d + 1], nd.level + 1, 0, -1};
}

// moments of nodes [a, b) (children are done): the host's sums in its order
__global__ __launch_bounds__(256) void k_moments(const GNode *__restrict__ nodes, long long a, long long b,
                                                 const double *__restrict__ pos, const double *__restrict__ mass,
                                                 const double *__restrict__ rs, double *__restrict__ nmass,
                                                 double *__restrict__ ncom, double *__restrict__ nquad,
                                                 double *__restrict__ box_in, double *__restrict__ box_out)
{
    const long long k = a + (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= b) return;
    const GNode nd = nodes[k];
    double M = 0, c[3] = {0, 0, 0}, q[6] = {0, 0, 0, 0, 0, 0};
    double bi[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    double bo[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    if (nd.n_child == 0) {
        for (long long i = nd.begin; i < nd.end; i++) {
            M += mass[i];
            for (int d = 0; d < 3; d++) {
                const double x = pos[3 * i + d];
                c[d] += mass[i] * x;
                bi[d] = fmin(bi[d], x);
                bi[3 + d] = fmax(bi[3 + d], x);
                bo[d] = fmin(bo[d], x - rs[i]);
                bo[3 + d] = fmax(bo[3 + d], x + rs[i]);
            }
        }
        if (M > 0) {
            for (int d = 0; d < 3; d++) c[d] /= M;
        } else {
            for (int d = 0; d < 3; d++) c[d] = 0.5 * (bi[d] + bi[3 + d]);
        }
        for (long long i = nd.begin; i < nd.end; i++) {
            const double x = pos[3 * i] - c[0], y = pos[3 * i + 1] - c[1], z = pos[3 * i + 2] - c[2], m = mass[i];
            q[0] += m * x * x;
            q[1] += m * y * y;
            q[2] += m * z * z;
            q[3] += m * x * y;
            q[4] += m * x * z;
            q[5] += m * y * z;
        }
    } else {
        for (int ch = 0; ch < nd.n_child; ch++) {
            const long long cc = nd.first_child + ch;
            M += nmass[cc];
            for (int d = 0; d < 3; d++) {
                c[d] += nmass[cc] * ncom[3 * cc + d];
                bi[d] = fmin(bi[d], box_in[6 * cc + d]);
                bi[3 + d] = fmax(bi[3 + d], box_in[6 * cc + 3 + d]);
                bo[d] = fmin(bo[d], box_out[6 * cc + d]);
                bo[3 + d] = fmax(bo[3 + d], box_out[6 * cc + 3 + d]);
            }
        }
        if (M > 0) {
            for (int d = 0; d < 3; d++) c[d] /= M;
        } else {
            for (int d = 0; d < 3; d++) c[d] = 0.5 * (bi[d] + bi[3 + d]);
        }
        for (int ch = 0; ch < nd.n_child; ch++) {
            const long long cc = nd.first_child + ch;
            const double m = nmass[cc];
            const double x = ncom[3 * cc] - c[0], y = ncom[3 * cc + 1] - c[1], z = ncom[3 * cc + 2] - c[2];
            const double *qc = nquad + 6 * cc;
            q[0] += qc[0] + m * x * x;
            q[1] += qc[1] + m * y * y;
            q[2] += qc[2] + m * z * z;
            q[3] += qc[3] + m * x * y;
            q[4] += qc[4] + m * x * z;
            q[5] += qc[5] + m * y * z;
        }
    }
    nmass[k] = M;
    for (int d = 0; d < 3; d++) ncom[3 * k + d] = c[d];
    for (int d = 0; d < 6; d++) {
        nquad[6 * k + d] = q[d];
        box_in[6 * k + d] = bi[d];
        box_out[6 * k + d] = bo[d];
    }
}

// group flags: the root when it qualifies (at most n_group_limit particles,
// or a leaf); k_group_children adds the qualifying children of every
// non-qualifying node, i.e. the largest qualifying cells
__global__ __launch_bounds__(256) void k_group_flags(const GNode *__restrict__ nodes, long long nn, int n_group_limit,
                                                     char *__restrict__ flag)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= nn) return;
    char f = 0;
    if (k == 0) {
        const GNode nd = nodes[0];
        f = (char)((nd.end - nd.begin <= n_group_limit) || nd.n_child == 0);
    }
    flag[k] = f;
}

// marks the children of every non-qualifying internal node that qualify
__global__ __launch_bounds__(256) void k_group_children(const GNode *__restrict__ nodes, long long nn,
                                                        int n_group_limit, char *__restrict__ flag)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= nn) return;
    const GNode nd = nodes[k];
    const bool q = (nd.end - nd.begin <= n_group_limit) || nd.n_child == 0;
    if (q) return;
    for (int ch = 0; ch < nd.n_child; ch++) {
        const GNode c = nodes[nd.first_child + ch];
        if ((c.end - c.begin <= n_group_limit) || c.n_child == 0) flag[nd.first_child + ch] = 1;
    }
}

__global__ __launch_bounds__(64) void k_fill_int(int *__restrict__ p, long long n, int v)
{
    for (long long k = threadIdx.x; k < n; k += 64) p[k] = v;
}

__global__ __launch_bounds__(256) void k_group_keys(const GNode *__restrict__ nodes, const int *__restrict__ gnode,
                                                    long long ng, long long *__restrict__ gbeg)
{
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= ng) return;
    gbeg[g] = nodes[gnode[g]].begin;
}

// per group: first particle, size, list lengths (host copy for the planner)
__global__ __launch_bounds__(256) void k_group_info(const GNode *__restrict__ nodes, const int *__restrict__ groups,
                                                    long long ng, const long long *__restrict__ off_e,
                                                    const long long *__restrict__ off_s, int4 *__restrict__ info)
{
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= ng) return;
    const GNode G = nodes[groups[g]];
    info[g] = make_int4((int)G.begin, (int)(G.end - G.begin), (int)(off_e[g + 1] - off_e[g]),
                        (int)(off_s[g + 1] - off_s[g]));
}

// sort key of an SPJ entry: the node's first particle
__global__ __launch_bounds__(256) void k_spj_keys(const GNode *__restrict__ nodes, const int *__restrict__ id_spj,
                                                  long long n, int *__restrict__ key)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    key[k] = (int)nodes[id_spj[k]].begin;
}

// resident records from the tree (the conversions of pth_resident_create)
__global__ __launch_bounds__(256) void k_res_particles(long long n, const double *__restrict__ spos,
                                                       const double *__restrict__ smass,
                                                       const double *__restrict__ srs, pth::EpiDev *__restrict__ epi,
                                                       pth::EpjDev *__restrict__ epj)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const double x = spos[3 * k], y = spos[3 * k + 1], z = spos[3 * k + 2];
    const float rs = (float)srs[k];
    pth::EpiDev e;
    e.x = x;
    e.y = y;
    e.z = z;
    e.rs = rs;
    e.pad = 0;
    epi[k] = e;
    pth::EpjDev j;
    j.x = x;
    j.y = y;
    j.z = z;
    j.m = (float)smass[k];
    j.rs = rs;
    epj[k] = j;
}

__global__ __launch_bounds__(256) void k_res_nodes(long long nn, const double *__restrict__ nmass,
                                                   const double *__restrict__ ncom, const double *__restrict__ nquad,
                                                   pth::SpjDev *__restrict__ spj)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= nn) return;
    const double *q = nquad + 6 * k;
    pth::SpjDev r;
    r.x = ncom[3 * k];
    r.y = ncom[3 * k + 1];
    r.z = ncom[3 * k + 2];
    r.m = (float)nmass[k];
    r.qxx = (float)q[0];
    r.qyy = (float)q[1];
    r.qzz = (float)q[2];
    r.qxy = (float)q[3];
    r.qxz = (float)q[4];
    r.qyz = (float)q[5];
    r.pad[0] = r.pad[1] = r.pad[2] = 0.f;
    spj[k] = r;
}

// monopole superparticles: the node's mass at its centre of mass
__global__ __launch_bounds__(256) void k_res_nodes_mono(long long nn, const double *__restrict__ nmass,
                                                        const double *__restrict__ ncom,
                                                        pth::EpjDev *__restrict__ spj)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= nn) return;
    pth::EpjDev r;
    r.x = ncom[3 * k];
    r.y = ncom[3 * k + 1];
    r.z = ncom[3 * k + 2];
    r.m = (float)nmass[k];
    r.rs = 0.f;
    spj[k] = r;
}

struct WalkArgs {
    const GNode *nodes;
    const double *ncom, *box_in, *box_out;
    const int *groups;
    long long n_group;
    // tile walks (tbox != null): group g is node groups[g] with the boxes
    // tbox[12 g ..] (particle box, box widened by r_search) instead of the
    // node's own
    const double *tbox;
    double root_len, inv_theta2;
    int theta_pos;
    long long *cnt_e, *cnt_s;                // counting pass (and a padded writing pass)
    const long long *off_e, *off_s;          // writing pass
    int *id_epj, *id_spj;
    // padded writing pass: group g writes into [off[g], off[g + 1]) at most,
    // skipping what does not fit, and stores its exact counts
    int padded;
};

__device__ __forceinline__ bool overlap6(const double *a, const double *b)
{
    for (int d = 0; d < 3; d++)
        if (a[d] > b[3 + d] || b[d] > a[3 + d]) return false;
    return true;
}

// inclusive wave scan of v
__device__ __forceinline__ long long wave_scan(long long v, int lane)
{
    for (int off = 1; off < 64; off <<= 1) {
        const long long t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// One 64-thread workgroup (one wave) per group at a time.
template <bool WRITE>
__global__ __launch_bounds__(64) void k_walk(const WalkArgs a)
{
    __shared__ int stk[WALK_STACK];
    const int lane = threadIdx.x;
    for (long long g = blockIdx.x; g < a.n_group; g += gridDim.x) {
        const int gn = a.groups[g];
        const GNode G = a.nodes[gn];
        const long long gb = G.begin, ge = G.end;
        double bin[6], bout[6];
        for (int d = 0; d < 6; d++) {
            bin[d] = a.tbox ? a.tbox[12 * g + d] : a.box_in[6 * (long long)gn + d];
            bout[d] = a.tbox ? a.tbox[12 * g + 6 + d] : a.box_out[6 * (long long)gn + d];
        }
        long long ne = 0, ns = 0;
        int *oe = nullptr, *os = nullptr;
        long long cap_e = LLONG_MAX, cap_s = LLONG_MAX;
        if (WRITE) {
            oe = a.id_epj + a.off_e[g];
            if (a.id_spj) os = a.id_spj + a.off_s[g];
            if (a.padded) {
                cap_e = a.off_e[g + 1] - a.off_e[g];
                if (a.id_spj) cap_s = a.off_s[g + 1] - a.off_s[g];
            }
        }
        if (lane == 0) stk[0] = 0;
        int size = 1;
        __syncthreads();
        while (size > 0) {
            // pop k nodes (at most 8 pushed per pop): while k > 1 the stack
            // stays below WALK_STACK - 160; past that, nodes are popped one at
            // a time and a descent adds at most 7 entries per level (21
            // levels, 147 < 160), so the stack never overflows
            int k = (WALK_STACK - 160 - size) / 8;
            k = max(1, min(64, min(k, size)));
            const bool act = lane < k;
            int c = 0;
            if (act) c = stk[size - 1 - lane];
            __syncthreads();
            long long me = 0, mc = 0, ms = 0;
            GNode nd{0, 0, 0, 0, 0};
            if (act) {
                nd = a.nodes[c];
                bool open;
                if (nd.begin <= gb && ge <= nd.end) open = true;
                else if (!a.theta_pos) open = true;
                else {
                    const double L = a.root_len / (double)(1LL << nd.level);
                    double s = 0;
                    for (int d = 0; d < 3; d++) {
                        const double p = a.ncom[3 * (long long)c + d];
                        double e = 0;
                        if (p < bin[d]) e = bin[d] - p;
                        else if (p > bin[3 + d]) e = p - bin[3 + d];
                        s += e * e;
                    }
                    open = s < L * L * a.inv_theta2 || overlap6(bout, a.box_in + 6 * (long long)c) ||
                           overlap6(bin, a.box_out + 6 * (long long)c);
                }
                if (open) {
                    if (nd.n_child == 0) me = nd.end - nd.begin;
                    else mc = nd.n_child;
                } else {
                    ms = 1;
                }
            }
            // the three counts scanned as one word: EPJ entries in bits 0-31
            // (a step's popped leaves are distinct, so their total is at most
            // n < 2^31), children in 32-47 (at most 8 per lane), SPJ entries
            // in 48-63 (at most 1 per lane)
            const long long pk = wave_scan(me | (mc << 32) | (ms << 48), lane);
            const long long tk = __shfl(pk, 63, 64);
            const long long pe = pk & 0xffffffffLL, pc = (pk >> 32) & 0xffff, ps = (pk >> 48) & 0xffff;
            const long long te = tk & 0xffffffffLL, tc = (tk >> 32) & 0xffff, ts = (tk >> 48) & 0xffff;
            if (WRITE) {
                for (long long t = 0; t < me; t++) {
                    const long long k = ne + pe - me + t;
                    if (k < cap_e) oe[k] = (int)(nd.begin + t);
                }
                if (ms && os && ns + ps - 1 < cap_s) os[ns + ps - 1] = c;
            }
            ne += te;
            ns += ts;
            size -= k;
            for (long long t = 0; t < mc; t++) stk[size + (pc - mc) + t] = (int)(nd.first_child + (mc - 1 - t));
            size += (int)tc;
            __syncthreads();
        }
        if ((!WRITE || a.padded) && lane == 0) {
            a.cnt_e[g] = ne;
            if (a.cnt_s) a.cnt_s[g] = ns;
        }
        __syncthreads();
    }
}

// Group g's list, written by a padded walk at poff[g], moved to off[g]
// (off[g + 1] - off[g] entries, checked on the host to fit the region).
__global__ __launch_bounds__(256) void k_compact(long long ng, const long long *__restrict__ poff,
                                                 const long long *__restrict__ off, const int *__restrict__ src,
                                                 int *__restrict__ dst)
{
    for (long long g = blockIdx.x; g < ng; g += gridDim.x) {
        const long long n = off[g + 1] - off[g];
        const int *s = src + poff[g];
        int *d = dst + off[g];
        for (long long k = threadIdx.x; k < n; k += 256) d[k] = s[k];
    }
}

// Regions of a one-pass neighbour filter: the last search's count of
// particle i + 1/4 + 4 (cap[n] = 0, so the scan's last entry is the total).
__global__ __launch_bounds__(256) void k_ngb_caps(long long n, const int *__restrict__ cnt, int *__restrict__ cap)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const int c = max(cnt[i], 0);
        cap[i] = c + c / 4 + 4;
    } else if (i == n) {
        cap[n] = 0;
    }
}

// Particle i's neighbours, written by a padded filter pass at poff[i], moved
// to off[i] (checked on the host to fit the region).
__global__ __launch_bounds__(256) void k_ngb_compact(long long n, const long long *__restrict__ poff,
                                                     const long long *__restrict__ off, const int *__restrict__ src,
                                                     int *__restrict__ dst)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long m = off[i + 1] - off[i];
    const int *s = src + poff[i];
    int *d = dst + off[i];
    for (long long k = 0; k < m; k++) d[k] = s[k];
}

// Boxes of the neighbour-search tiles (the largest cells with at most 64
// particles): the cell's particle box and its box widened by r_search,
// pushed outwards by (|v| + rs_max) 2^-40, far above the rounding of any
// coordinate, difference or box bound involved, so a tile walk with these
// boxes opens every leaf holding a neighbour of a tile particle under the
// exact FP64 test (the walk's lists are a superset; k_ngb filters exactly).
__global__ __launch_bounds__(256) void k_tile_boxes(const int *__restrict__ tiles, long long nt,
                                                    const double *__restrict__ box_in,
                                                    const double *__restrict__ box_out, double rs_max,
                                                    double *__restrict__ tbox)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    const long long c = tiles[t];
    const double m = 0x1p-40;
    for (int d = 0; d < 3; d++) {
        const double lo = box_in[6 * c + d], hi = box_in[6 * c + 3 + d];
        const double lo2 = box_out[6 * c + d], hi2 = box_out[6 * c + 3 + d];
        tbox[12 * t + d] = lo - (fabs(lo) + rs_max) * m;
        tbox[12 * t + 3 + d] = hi + (fabs(hi) + rs_max) * m;
        tbox[12 * t + 6 + d] = lo2 - (fabs(lo2) + rs_max) * m;
        tbox[12 * t + 9 + d] = hi2 + (fabs(hi2) + rs_max) * m;
    }
}

// per tile: first particle, size, candidate count (k_ngb's group record)
__global__ __launch_bounds__(256) void k_tile_info(const GNode *__restrict__ nodes, const int *__restrict__ tiles,
                                                   long long nt, const long long *__restrict__ off,
                                                   int4 *__restrict__ info)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    const GNode c = nodes[tiles[t]];
    info[t] = make_int4((int)c.begin, (int)(c.end - c.begin), (int)(off[t + 1] - off[t]), 0);
}

// Neighbour search (src/soft_force.hpp:11-34, SearchNeighborEpEpNoSimd) on
// candidate lists that hold every symmetric-search neighbour (the tile walks
// of pth_gtree_neighbors, or a group's EPJ list): for every i of a tile the
// list entries j with |x_i - x_j|^2 < max(rs_i, rs_j)^2 in FP64, self
// included.  One wave per tile at a time, up to NGB_T sub-tiles of 64 i per
// pass (lane l holds i = l + 64 t), j staged
// 64 at a time in LDS (read as broadcasts).  A (tile, chunk) pair is
// skipped when the boxes of the tile's and the chunk's positions are at
// least R = max(rs) apart: the gap g (per axis, >= 0) bounds every pair's
// |dx| from below, and since each rounded operation is monotone,
// gx*gx + gy*gy + gz*gz >= R*R implies dx*dx + dy*dy + dz*dz >= rs*rs in
// the same FP64 arithmetic, so the skip never drops a neighbour.
// Counting pass: cnt[input index of i]; writing pass: input indices of the
// neighbours at ngb[off[input index of i]..], in list order.
constexpr int NGB_T = 8;

struct NgbArgs {
    const double *spos, *srs;
    const long long *order;
    const int4 *info;  // per group: first particle, size, n_epj, n_spj
    const long long *off_e;
    const int *id_epj;
    long long n_group;
    int *cnt;
    const long long *off;
    int *ngb;
    // padded writing pass: particle i writes into [off[i], off[i + 1]) at
    // most, stores its exact count, and sets *flag when it did not fit
    int padded;
    int *flag;
};

struct NgbJ {
    double x, y, z, rs;
};

__device__ __forceinline__ double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

struct Box {
    double lo[3], hi[3], rs;
};

// box of the lane values (inactive lanes pass act = false)
__device__ __forceinline__ Box wave_box(bool act, double x, double y, double z, double rs)
{
    Box b;
    b.lo[0] = wave_min(act ? x : 1e300);
    b.lo[1] = wave_min(act ? y : 1e300);
    b.lo[2] = wave_min(act ? z : 1e300);
    b.hi[0] = wave_max(act ? x : -1e300);
    b.hi[1] = wave_max(act ? y : -1e300);
    b.hi[2] = wave_max(act ? z : -1e300);
    b.rs = wave_max(act ? rs : 0.0);
    return b;
}

// true when no pair between the two boxes can be neighbours
__device__ __forceinline__ bool boxes_apart(const Box &a, const Box &b)
{
    const double gx = fmax(0.0, fmax(b.lo[0] - a.hi[0], a.lo[0] - b.hi[0]));
    const double gy = fmax(0.0, fmax(b.lo[1] - a.hi[1], a.lo[1] - b.hi[1]));
    const double gz = fmax(0.0, fmax(b.lo[2] - a.hi[2], a.lo[2] - b.hi[2]));
    const double r = fmax(a.rs, b.rs);
    return gx * gx + gy * gy + gz * gz >= r * r;
}

template <bool WRITE>
__global__ __launch_bounds__(64) void k_ngb(const NgbArgs a)
{
    __shared__ NgbJ sj[64];
    __shared__ int sid[64];
    __shared__ Box tb[NGB_T];  // boxes of the tiles (wave-uniform)
    const int lane = threadIdx.x;
    for (long long g = blockIdx.x; g < a.n_group; g += gridDim.x) {
        const int4 G = a.info[g];
        const long long eb = a.off_e[g], ee = a.off_e[g + 1];
        for (int s0 = 0; s0 < G.y; s0 += 64 * NGB_T) {
            const int nt = min(NGB_T, (G.y - s0 + 63) / 64);
            double xi[NGB_T], yi[NGB_T], zi[NGB_T], ri[NGB_T];
            long long w[NGB_T];
            int c[NGB_T], cap[NGB_T];
            bool act[NGB_T];
#pragma unroll
            for (int t = 0; t < NGB_T; t++) {
                c[t] = 0;
                cap[t] = INT_MAX;
                act[t] = false;
                xi[t] = yi[t] = zi[t] = ri[t] = 0.0;
                w[t] = 0;
                if (t < nt) {
                    act[t] = s0 + 64 * t + lane < G.y;
                    if (act[t]) {
                        const long long i = (long long)G.x + s0 + 64 * t + lane;
                        xi[t] = a.spos[3 * i];
                        yi[t] = a.spos[3 * i + 1];
                        zi[t] = a.spos[3 * i + 2];
                        ri[t] = a.srs[i];
                        if (WRITE) {
                            const long long oi = a.order[i];
                            w[t] = a.off[oi];
                            if (a.padded) cap[t] = (int)(a.off[oi + 1] - w[t]);
                        }
                    }
                    const Box b = wave_box(act[t], xi[t], yi[t], zi[t], ri[t]);
                    if (lane == 0) tb[t] = b;
                }
            }
            for (long long k0 = eb; k0 < ee; k0 += 64) {
                const long long k = k0 + lane;
                const bool jact = k < ee;
                NgbJ J{0, 0, 0, 0};
                if (jact) {
                    const int j = a.id_epj[k];
                    J = NgbJ{a.spos[3 * (long long)j], a.spos[3 * (long long)j + 1], a.spos[3 * (long long)j + 2],
                             a.srs[j]};
                    sj[lane] = J;
                    if (WRITE) sid[lane] = (int)a.order[j];
                }
                const Box cb = wave_box(jact, J.x, J.y, J.z, J.rs);
                __syncthreads();
                const int m = (int)min((long long)64, ee - k0);
#pragma unroll
                for (int t = 0; t < NGB_T; t++) {
                    if (t >= nt || boxes_apart(tb[t], cb)) continue;  // wave-uniform (LDS broadcast)
                    if (act[t])
                        for (int q = 0; q < m; q++) {
                            const NgbJ Q = sj[q];
                            const double dx = xi[t] - Q.x, dy = yi[t] - Q.y, dz = zi[t] - Q.z;
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            const double rs = (ri[t] < Q.rs) ? Q.rs : ri[t];
                            if (r2 < rs * rs) {
                                if (WRITE && c[t] < cap[t]) a.ngb[w[t] + c[t]] = sid[q];
                                c[t]++;
                            }
                        }
                }
                __syncthreads();
            }
            if (!WRITE || a.padded) {
                bool over = false;
#pragma unroll
                for (int t = 0; t < NGB_T; t++)
                    if (t < nt && act[t]) {
                        a.cnt[a.order[(long long)G.x + s0 + 64 * t + lane]] = c[t];
                        over = over || c[t] > cap[t];
                    }
                if (WRITE && over) atomicOr(a.flag, 1);
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

struct pth_gtree {
    pth_ctx *ctx = nullptr;
    long long n = 0, n_node = 0, n_group = 0, n_le = 0, n_ls = 0;
    DevBuf order, key, skey, sidx, spos, smass, srs, nodes, nmass, ncom, nquad, box_in, box_out, groups, off_e, off_s,
        id_epj, id_spj, tmp;
    // build-time scratch, released when the build ends (or by pth_gtree_free)
    DevBuf dpos, dmass, drs, part, bnd, cnt, off, flag, gsel, nsel, gkey, gkey2, ce, cs, alt, lkey, lkey2, info;
    // neighbour lists (pth_gtree_neighbors): count per input index, offsets,
    // entries; tile boxes, candidate counts / offsets / lists, tile records
    DevBuf ngb_cnt, ngb_off, ngb, tiles, tbox, tcnt, toff, tl, tl_alt, tinfo;
    // one-pass walks: list regions sized from the previous build's lengths
    // (padded offsets, EPJ then SPJ), the lists written there
    DevBuf pad_off, pad_e, pad_s;
    std::vector<long long> pad_h;
    int walk_passes = 0;  // of the last build: 1 (one padded pass) or 2 (count, then write)
    void *pinned = nullptr;  // pinned host block for small read-backs and uploads (PIN_* layout)
    // one-pass tile walks of the neighbour search: the last search's
    // candidate counts per tile (nt + 1 entries), padded offsets, lists
    std::vector<long long> tcnt_h, tpad_h;
    DevBuf tpad_off, tl_pad;
    int tile_passes = 0;  // of the last search's tile walk
    // one-pass neighbour filters: regions from the last completed search's
    // per-particle counts (ngb_cnt), valid while ngb_prev_ok
    DevBuf ngb_cap, ngb_poff, ngb_pad, ngb_flag;
    bool ngb_prev_ok = false;
    long long ngb_prev_n = -1, ngb_prev_total = 0;
    int ngb_passes = 0;  // of the last search's filter
    long long n_ngb = -1;  // total entries, -1 before the first search
    long long n_cand = 0;  // candidate pairs of the last search
    float ms_ngb = 0.f;
    void release_scratch()
    {
        DevBuf *s[] = {&dpos, &dmass, &drs, &part, &bnd, &cnt, &off, &flag, &gsel, &nsel, &gkey, &gkey2, &ce, &cs,
                       &alt, &lkey, &lkey2, &info};
        for (DevBuf *b : s) b->release();
    }
    std::vector<long long> level_start;
    std::vector<int4> groups_h;  // per group: first particle, size, n_epj, n_spj
    // the plan of a resident evaluation of the last build's lists, made on
    // the host while the device writes and orders them, and the parameters
    // it was made for
    mutable Plan pre_plan;  // taken over (swapped out) by a refill
    PlanKey pre_key;
    bool pre_ok = false;
    mutable bool pre_valid = false;
    int n_leaf_limit = 20, n_group_limit = 512;
    double theta = 0.3;
    double rs_max = 0.0, root_len = 1.0;  // largest r_search, root cell side
    bool canonical = false;
    bool built = false;
    float ms_build = 0.f, ms_walk = 0.f, ms_order = 0.f;
};

namespace {

#define GT_CHK(call, what)                               \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return t->ctx->hip_fail(e_, what); \
    } while (0)

inline unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

// Pinned host block of a tree: small counts read back at byte 0, the root
// node's upload at byte 64, the bounding-box partials (8 doubles for each of
// at most 1024 blocks) from byte 128.
constexpr size_t PIN_COUNT = 0, PIN_ROOT = 64, PIN_BBOX = 128, PIN_BYTES = PIN_BBOX + 8 * 8 * 1024;

hipError_t ensure_pinned(pth_gtree *t)
{
    if (t->pinned) return hipSuccess;
    const hipError_t e = hipHostMalloc(&t->pinned, PIN_BYTES, hipHostMallocDefault);
    if (e != hipSuccess) t->pinned = nullptr;
    return e;
}

template <class T> T *pinned_at(pth_gtree *t, size_t off) { return reinterpret_cast<T *>(static_cast<char *>(t->pinned) + off); }

// temp storage for a hipcub call, grown on demand
hipError_t cub_temp(pth_gtree *t, size_t bytes) { return t->tmp.ensure(std::max<size_t>(bytes, 256)); }

}  // namespace

namespace {

// The largest cells with at most `limit` particles (or leaves), as node
// indices in Morton order (by first particle): the i-groups for limit =
// n_group_limit, the neighbour-search tiles for limit = 64.
int select_cells(pth_gtree *t, int limit, DevBuf &out, long long &n_out)
{
    pth_ctx *ctx = t->ctx;
    hipStream_t st = ctx->stream;
    const long long nn = t->n_node;
    size_t tb = 0;
    hipError_t e;
    DevBuf &flag = t->flag, &gsel = t->gsel, &nsel = t->nsel, &gkey = t->gkey, &gkey2 = t->gkey2;
    if ((e = flag.ensure(nn)) || (e = gsel.ensure(sizeof(int) * nn)) || (e = nsel.ensure(sizeof(int))))
        return ctx->hip_fail(e, "cell selection allocation");
    hipLaunchKernelGGL(k_group_flags, dim3(blocks(nn)), dim3(256), 0, st, t->nodes.as<GNode>(), nn, limit,
                       flag.as<char>());
    hipLaunchKernelGGL(k_group_children, dim3(blocks(nn)), dim3(256), 0, st, t->nodes.as<GNode>(), nn, limit,
                       flag.as<char>());
    hipcub::CountingInputIterator<int> it(0);
    if ((e = hipcub::DeviceSelect::Flagged(nullptr, tb, it, flag.as<char>(), gsel.as<int>(), nsel.as<int>(), (int)nn,
                                           st)) ||
        (e = cub_temp(t, tb)) ||
        (e = hipcub::DeviceSelect::Flagged(t->tmp.p, tb, it, flag.as<char>(), gsel.as<int>(), nsel.as<int>(),
                                           (int)nn, st)))
        return ctx->hip_fail(e, "cell selection");
    if ((e = ensure_pinned(t)) ||
        (e = hipMemcpyAsync(pinned_at<int>(t, PIN_COUNT), nsel.p, sizeof(int), hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return ctx->hip_fail(e, "cell count");
    const long long ng = *static_cast<volatile int *>(pinned_at<int>(t, PIN_COUNT));
    if ((e = gkey.ensure(8 * ng)) || (e = gkey2.ensure(8 * ng)) || (e = out.ensure(sizeof(int) * std::max(ng, 1LL))))
        return ctx->hip_fail(e, "cell selection allocation");
    hipLaunchKernelGGL(k_group_keys, dim3(blocks(ng)), dim3(256), 0, st, t->nodes.as<GNode>(), gsel.as<int>(), ng,
                       gkey.as<long long>());
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, gkey.as<long long>(), gkey2.as<long long>(),
                                                gsel.as<int>(), out.as<int>(), (int)ng, 0, 32, st)) ||
        (e = cub_temp(t, tb)) ||
        (e = hipcub::DeviceRadixSort::SortPairs(t->tmp.p, tb, gkey.as<long long>(), gkey2.as<long long>(),
                                                gsel.as<int>(), out.as<int>(), (int)ng, 0, 32, st)))
        return ctx->hip_fail(e, "cell sort");
    n_out = ng;
    return PTH_OK;
}

// Builds (or rebuilds, reusing t's device memory) the tree of n particles.
// pos / mass / rs are host arrays, or device arrays when on_device.
int gtree_run(pth_gtree *t, long long n, const double *pos, const double *mass, const double *r_search,
              bool on_device)
{
    pth_ctx *ctx = t->ctx;
    const int n_leaf_limit = t->n_leaf_limit, n_group_limit = t->n_group_limit;
    const double theta = t->theta;
    const bool was_built = t->built;  // groups_h then holds the last build's list lengths
    t->built = false;
    t->pre_valid = false;
    t->n_ngb = -1;
    t->n = n;
    t->n_node = t->n_group = t->n_le = t->n_ls = 0;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = ctx->stream;
    hipError_t e;
    // particles in (unsorted), then sorted copies
    const double *dpos = pos, *dmass = mass, *drs = r_search;
    if (on_device) {
        if (!device_range(pos, 24 * (size_t)n) || !device_range(mass, 8 * (size_t)n) ||
            (r_search && !device_range(r_search, 8 * (size_t)n)))
            return ctx->fail(PTH_ERR_ARG, "pth_gtree: the particle arrays are not device allocations of n entries");
    } else {
        auto up = [&](DevBuf &b, const double *h, size_t cnt) -> hipError_t {
            hipError_t r = b.ensure(sizeof(double) * cnt);
            if (r == hipSuccess) r = hipMemcpyAsync(b.p, h, sizeof(double) * cnt, hipMemcpyHostToDevice, st);
            return r;
        };
        if ((e = up(t->dpos, pos, 3 * (size_t)n)) || (e = up(t->dmass, mass, (size_t)n)) ||
            (r_search && (e = up(t->drs, r_search, (size_t)n))))
            return ctx->hip_fail(e, "pth_gtree: upload");
        dpos = t->dpos.as<double>();
        dmass = t->dmass.as<double>();
        drs = r_search ? t->drs.as<double>() : nullptr;
    }
    // 1. bounding cube (per-block partials reduced on the host)
    const int nb = (int)std::min<long long>(1024, (n + 255) / 256);
    DevBuf &part = t->part;
    if ((e = ensure_pinned(t))) return ctx->hip_fail(e, "pinned buffer");
    const double *hp = pinned_at<double>(t, PIN_BBOX);
    if ((e = part.ensure(sizeof(double) * 8 * nb))) return ctx->hip_fail(e, "bbox");
    hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, st, dpos, drs, n, part.as<double>());
    if ((e = hipMemcpyAsync(pinned_at<double>(t, PIN_BBOX), part.p, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost,
                            st)) ||
        (e = hipStreamSynchronize(st)))
        return ctx->hip_fail(e, "bbox");
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    t->rs_max = 0.0;
    bool bad = false;
    for (int b = 0; b < nb; b++) {
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], hp[8 * b + d]);
            hi[d] = std::max(hi[d], hp[8 * b + 3 + d]);
        }
        t->rs_max = std::max(t->rs_max, hp[8 * b + 6]);
        bad = bad || hp[8 * b + 7] != 0.0;
    }
    if (bad) return ctx->fail(PTH_ERR_ARG, "pth_gtree: non-finite particle positions or r_search");
    double len = 0;
    for (int d = 0; d < 3; d++) len = std::max(len, hi[d] - lo[d]);
    len = (len > 0) ? len * (1.0 + 1e-9) : 1.0;
    double root[3];
    for (int d = 0; d < 3; d++) root[d] = 0.5 * (lo[d] + hi[d]) - 0.5 * len;
    const double scale = (double)(1u << GT_MAX_LEVEL) / len;
    t->root_len = len;
    // 2. keys, stable sort of (key, index), sorted particles
    if ((e = t->key.ensure(8 * (size_t)n)) || (e = t->skey.ensure(8 * (size_t)n)) ||
        (e = t->sidx.ensure(8 * (size_t)n)) || (e = t->order.ensure(8 * (size_t)n)))
        return ctx->hip_fail(e, "key allocation");
    hipLaunchKernelGGL(k_keys, dim3(blocks(n)), dim3(256), 0, st, dpos, n, root[0], root[1], root[2],
                       scale, t->key.as<unsigned long long>(), t->sidx.as<long long>());
    size_t tb = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, t->key.as<unsigned long long>(),
                                                t->skey.as<unsigned long long>(), t->sidx.as<long long>(),
                                                t->order.as<long long>(), n, 0, 3 * GT_MAX_LEVEL, st)) ||
        (e = cub_temp(t, tb)) ||
        (e = hipcub::DeviceRadixSort::SortPairs(t->tmp.p, tb, t->key.as<unsigned long long>(),
                                                t->skey.as<unsigned long long>(), t->sidx.as<long long>(),
                                                t->order.as<long long>(), n, 0, 3 * GT_MAX_LEVEL, st)))
        return ctx->hip_fail(e, "key sort");
    if ((e = t->spos.ensure(24 * (size_t)n)) || (e = t->smass.ensure(8 * (size_t)n)) ||
        (e = t->srs.ensure(8 * (size_t)n)))
        return ctx->hip_fail(e, "particle allocation");
    hipLaunchKernelGGL(k_gather, dim3(blocks(n)), dim3(256), 0, st, n, t->order.as<long long>(), dpos,
                       dmass, drs, t->spos.as<double>(),
                       t->smass.as<double>(), t->srs.as<double>());
    // 3. nodes level by level
    if ((e = t->nodes.ensure(sizeof(GNode) * std::max<long long>(2 * n / std::max(n_leaf_limit, 1) + 64, 1024))))
        return ctx->hip_fail(e, "node allocation");
    long long cap = (long long)(t->nodes.cap / sizeof(GNode));
    // (from pinned memory: the copy is ordered before the level loop's
    // launches and needs no wait)
    *pinned_at<GNode>(t, PIN_ROOT) = GNode{0, n, 0, 0, -1};
    if ((e = hipMemcpyAsync(t->nodes.p, pinned_at<GNode>(t, PIN_ROOT), sizeof(GNode), hipMemcpyHostToDevice, st)))
        return ctx->hip_fail(e, "node upload");
    t->level_start.assign(1, 0);
    long long lb = 0, le = 1;
    DevBuf &bnd = t->bnd, &cnt = t->cnt, &off = t->off;
    for (;;) {
        t->level_start.push_back(le);
        const long long nl = le - lb;
        if ((e = bnd.ensure(sizeof(long long) * 9 * nl)) || (e = cnt.ensure(sizeof(int) * (nl + 1))) ||
            (e = off.ensure(sizeof(int) * (nl + 1))))
            return ctx->hip_fail(e, "level buffers");
        hipLaunchKernelGGL(k_split_count, dim3(blocks(nl)), dim3(256), 0, st, t->nodes.as<GNode>(), lb, nl,
                           t->skey.as<unsigned long long>(), n_leaf_limit, bnd.as<long long>(), cnt.as<int>());
        if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<int>(), off.as<int>(), (int)nl + 1, st)) ||
            (e = cub_temp(t, tb)) ||
            (e = hipcub::DeviceScan::ExclusiveSum(t->tmp.p, tb, cnt.as<int>(), off.as<int>(), (int)nl + 1, st)))
            return ctx->hip_fail(e, "level scan");
        // one small copy into pinned memory per level
        if ((e = hipMemcpyAsync(pinned_at<int>(t, PIN_COUNT), off.as<int>() + nl, sizeof(int), hipMemcpyDeviceToHost,
                                st)) ||
            (e = hipStreamSynchronize(st)))
            return ctx->hip_fail(e, "level size");
        const long long total = *static_cast<volatile int *>(pinned_at<int>(t, PIN_COUNT));
        if (total == 0) break;
        if (le + total > cap) {
            // grow the node array, keeping its contents
            const long long ncap = std::max(2 * cap, le + total);
            DevBuf nb2;
            if ((e = nb2.ensure(sizeof(GNode) * ncap)) ||
                (e = hipMemcpyAsync(nb2.p, t->nodes.p, sizeof(GNode) * le, hipMemcpyDeviceToDevice, st)) ||
                (e = hipStreamSynchronize(st)))
                return ctx->hip_fail(e, "node growth");
            t->nodes.release();
            t->nodes = nb2;
            cap = ncap;
        }
        hipLaunchKernelGGL(k_split_write, dim3(blocks(nl)), dim3(256), 0, st, t->nodes.as<GNode>(), lb, nl, le,
                           bnd.as<long long>(), cnt.as<int>(), off.as<int>());
        lb = le;
        le += total;
    }
    t->n_node = le;
    if (le > 0x7fffffffLL)  // the walks keep node indices as int32
        return ctx->fail(PTH_ERR_CAPACITY, "pth_gtree: node count exceeds the int32 range");
    const long long nn = le;
    // 4. moments bottom-up
    if ((e = t->nmass.ensure(8 * nn)) || (e = t->ncom.ensure(24 * nn)) || (e = t->nquad.ensure(48 * nn)) ||
        (e = t->box_in.ensure(48 * nn)) || (e = t->box_out.ensure(48 * nn)))
        return ctx->hip_fail(e, "moment allocation");
    for (int l = (int)t->level_start.size() - 2; l >= 0; l--) {
        const long long a = t->level_start[l], b = t->level_start[l + 1];
        if (b <= a) continue;
        hipLaunchKernelGGL(k_moments, dim3(blocks(b - a)), dim3(256), 0, st, t->nodes.as<GNode>(), a, b,
                           t->spos.as<double>(), t->smass.as<double>(), t->srs.as<double>(), t->nmass.as<double>(),
                           t->ncom.as<double>(), t->nquad.as<double>(), t->box_in.as<double>(),
                           t->box_out.as<double>());
    }
    // 5. groups in Morton order
    {
        long long ng_sel = 0;
        const int rc = select_cells(t, n_group_limit, t->groups, ng_sel);
        if (rc != PTH_OK) return rc;
        t->n_group = ng_sel;
    }
    const long long ng = t->n_group;
    if ((e = hipStreamSynchronize(st))) return ctx->hip_fail(e, "tree build");
    const auto t1 = std::chrono::steady_clock::now();
    t->ms_build = std::chrono::duration<float, std::milli>(t1 - t0).count();
    // 6. walk. With the previous build's list lengths per group at hand (same
    //    number of groups: a rebuild of slowly moving particles), one writing
    //    pass into regions of those lengths + 1/4 + 64 that also counts;
    //    otherwise a counting pass first.  Either way the exact counts give
    //    the list offsets; a region that turned out too small (checked on the
    //    host) means a second, exact writing pass.
    WalkArgs wa{};
    wa.nodes = t->nodes.as<GNode>();
    wa.ncom = t->ncom.as<double>();
    wa.box_in = t->box_in.as<double>();
    wa.box_out = t->box_out.as<double>();
    wa.groups = t->groups.as<int>();
    wa.n_group = ng;
    wa.root_len = len;
    wa.inv_theta2 = (theta > 0) ? 1.0 / (theta * theta) : 0.0;
    wa.theta_pos = theta > 0;
    DevBuf &ce = t->ce, &cs = t->cs;
    if ((e = ce.ensure(8 * (ng + 1))) || (e = cs.ensure(8 * (ng + 1))) || (e = t->off_e.ensure(8 * (ng + 1))) ||
        (e = t->off_s.ensure(8 * (ng + 1))))
        return ctx->hip_fail(e, "walk allocation");
    wa.cnt_e = ce.as<long long>();
    wa.cnt_s = cs.as<long long>();
    const unsigned wg = (unsigned)std::min<long long>(ng, 65536);
    const bool one_pass = ng > 0 && was_built && (long long)t->groups_h.size() == ng;
    if (one_pass) {
        std::vector<long long> &ph = t->pad_h;  // padded offsets: EPJ [0, ng], SPJ [ng + 1, 2 ng + 1]
        ph.assign(2 * (ng + 1), 0);
        long long pe = 0, ps = 0;
        for (long long g = 0; g < ng; g++) {
            const long long le = std::max(t->groups_h[g].z, 0), ls = std::max(t->groups_h[g].w, 0);
            ph[g] = pe;
            ph[ng + 1 + g] = ps;
            pe += le + le / 4 + 64;
            ps += ls + ls / 4 + 64;
        }
        ph[ng] = pe;
        ph[2 * ng + 1] = ps;
        if ((e = t->pad_off.ensure(8 * ph.size())) || (e = t->pad_e.ensure(4 * (size_t)pe)) ||
            (e = t->pad_s.ensure(4 * (size_t)ps)) ||
            (e = hipMemcpyAsync(t->pad_off.p, ph.data(), 8 * ph.size(), hipMemcpyHostToDevice, st)))
            return ctx->hip_fail(e, "padded list allocation");
        wa.off_e = t->pad_off.as<long long>();
        wa.off_s = t->pad_off.as<long long>() + ng + 1;
        wa.id_epj = t->pad_e.as<int>();
        wa.id_spj = t->pad_s.as<int>();
        wa.padded = 1;
        hipLaunchKernelGGL(k_walk<true>, dim3(wg), dim3(64), 0, st, wa);
    } else if (ng > 0) {
        hipLaunchKernelGGL(k_walk<false>, dim3(wg), dim3(64), 0, st, wa);
    }
    long long zero = 0;
    if ((e = hipMemcpyAsync(ce.as<long long>() + ng, &zero, 8, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(cs.as<long long>() + ng, &zero, 8, hipMemcpyHostToDevice, st)))
        return ctx->hip_fail(e, "walk counts");
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ce.as<long long>(), t->off_e.as<long long>(), (int)ng + 1,
                                              st)) ||
        (e = cub_temp(t, tb)) ||
        (e = hipcub::DeviceScan::ExclusiveSum(t->tmp.p, tb, ce.as<long long>(), t->off_e.as<long long>(), (int)ng + 1,
                                              st)) ||
        (e = hipcub::DeviceScan::ExclusiveSum(t->tmp.p, tb, cs.as<long long>(), t->off_s.as<long long>(), (int)ng + 1,
                                              st)))
        return ctx->hip_fail(e, "walk scan");
    // per-group sizes (first particle, size, list lengths) and the totals
    t->groups_h.assign(ng, make_int4(0, 0, 0, 0));
    if ((e = t->info.ensure(sizeof(int4) * std::max<long long>(ng, 1)))) return ctx->hip_fail(e, "group sizes");
    if (ng > 0)
        hipLaunchKernelGGL(k_group_info, dim3(blocks(ng)), dim3(256), 0, st, t->nodes.as<GNode>(), t->groups.as<int>(),
                           ng, t->off_e.as<long long>(), t->off_s.as<long long>(), t->info.as<int4>());
    long long tot[2] = {0, 0};
    if ((ng > 0 &&
         (e = hipMemcpyAsync(t->groups_h.data(), t->info.p, sizeof(int4) * ng, hipMemcpyDeviceToHost, st))) ||
        (e = hipMemcpyAsync(&tot[0], t->off_e.as<long long>() + ng, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(&tot[1], t->off_s.as<long long>() + ng, 8, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return ctx->hip_fail(e, "walk totals");
    if (tot[0] > 0x7fffffffLL || tot[1] > 0x7fffffffLL)
        return ctx->fail(PTH_ERR_CAPACITY, "pth_gtree_build: lists exceed the int32 range");
    t->n_le = tot[0];
    t->n_ls = tot[1];
    // did every group's lists fit the padded regions?
    bool rewalk = !one_pass;
    for (long long g = 0; g < ng && !rewalk; g++)
        rewalk = t->groups_h[g].z > t->pad_h[g + 1] - t->pad_h[g] ||
                 t->groups_h[g].w > t->pad_h[ng + 2 + g] - t->pad_h[ng + 1 + g];
    t->walk_passes = ng > 0 ? (one_pass && !rewalk ? 1 : 2) : 0;
    // every buffer of the write pass and the ordering sorts, before anything
    // is enqueued (a growing buffer's free would wait for the device)
    const bool order = t->canonical && ng > 0;
    int bits = 1;
    while (bits < 31 && (1LL << bits) < n) bits++;  // keys are particle positions < n
    const long long *ob = t->off_e.as<long long>(), *osp = t->off_s.as<long long>();
    size_t tb_e = 0, tb_s = 0;
    // the ordering swaps id_epj / id_spj with alt, so each of the three may
    // hold either list next: all three are sized for the longer one
    const size_t lbytes = 4 * (size_t)std::max<long long>(std::max(tot[0], tot[1]), 1);
    if ((e = t->id_epj.ensure(lbytes)) || (e = t->id_spj.ensure(lbytes)))
        return ctx->hip_fail(e, "list allocation");
    if (order) {
        if ((e = t->alt.ensure(lbytes)) || (e = t->lkey.ensure(4 * std::max<long long>(tot[1], 1))) ||
            (e = t->lkey2.ensure(4 * std::max<long long>(tot[1], 1))) ||
            (e = hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb_e, t->id_epj.as<int>(), t->alt.as<int>(),
                                                            (int)tot[0], (int)ng, ob, ob + 1, 0, bits, st)) ||
            (e = hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tb_s, t->lkey.as<int>(), t->lkey2.as<int>(),
                                                             t->id_spj.as<int>(), t->alt.as<int>(), (int)tot[1],
                                                             (int)ng, osp, osp + 1, 0, bits, st)) ||
            (e = cub_temp(t, std::max(tb_e, tb_s))))
            return ctx->hip_fail(e, "list ordering allocation");
    }
    // write pass (or the move of the padded lists), then the ordering, all
    // enqueued without waits
    wa.off_e = t->off_e.as<long long>();
    wa.off_s = t->off_s.as<long long>();
    wa.id_epj = t->id_epj.as<int>();
    wa.id_spj = t->id_spj.as<int>();
    wa.padded = 0;
    const auto t2 = std::chrono::steady_clock::now();
    t->ms_walk = std::chrono::duration<float, std::milli>(t2 - t1).count();
    if (t->id_epj.cap < 4 * (size_t)tot[0] || t->id_spj.cap < 4 * (size_t)tot[1])
        return ctx->fail(PTH_ERR_STATE, "pth_gtree: list buffers smaller than the lists");
    if (ng > 0 && rewalk) {
        hipLaunchKernelGGL(k_walk<true>, dim3(wg), dim3(64), 0, st, wa);
    } else if (ng > 0) {
        // every group fits its region (checked above): n <= region size
        hipLaunchKernelGGL(k_compact, dim3(wg), dim3(256), 0, st, ng, t->pad_off.as<long long>(),
                           t->off_e.as<long long>(), t->pad_e.as<int>(), t->id_epj.as<int>());
        hipLaunchKernelGGL(k_compact, dim3(wg), dim3(256), 0, st, ng, t->pad_off.as<long long>() + ng + 1,
                           t->off_s.as<long long>(), t->pad_s.as<int>(), t->id_spj.as<int>());
    }
    // 7. canonical order: segmented sorts of the lists (one segment per group)
    if (order) {
        if (t->alt.cap < lbytes || t->id_epj.cap < lbytes || t->id_spj.cap < lbytes ||
            t->lkey.cap < 4 * (size_t)tot[1] || t->lkey2.cap < 4 * (size_t)tot[1] ||
            t->tmp.cap < std::max(tb_e, tb_s))
            return ctx->fail(PTH_ERR_STATE, "pth_gtree: ordering buffers smaller than the lists");
        if (tot[0] > 0) {
            if ((e = hipcub::DeviceSegmentedRadixSort::SortKeys(t->tmp.p, tb_e, t->id_epj.as<int>(), t->alt.as<int>(),
                                                                (int)tot[0], (int)ng, ob, ob + 1, 0, bits, st)))
                return ctx->hip_fail(e, "EPJ list ordering");
            std::swap(t->id_epj, t->alt);
        }
        if (tot[1] > 0) {
            hipLaunchKernelGGL(k_spj_keys, dim3(blocks(tot[1])), dim3(256), 0, st, t->nodes.as<GNode>(),
                               t->id_spj.as<int>(), tot[1], t->lkey.as<int>());
            if ((e = hipcub::DeviceSegmentedRadixSort::SortPairs(t->tmp.p, tb_s, t->lkey.as<int>(), t->lkey2.as<int>(),
                                                                 t->id_spj.as<int>(), t->alt.as<int>(), (int)tot[1],
                                                                 (int)ng, osp, osp + 1, 0, bits, st)))
                return ctx->hip_fail(e, "SPJ list ordering");
            std::swap(t->id_spj, t->alt);
        }
    }
    // meanwhile, the plan of a resident evaluation of these lists under the
    // context's parameters (pth_resident_from_gtree / _refill take it over)
    {
        std::vector<int> ni(ng), nje(ng), njs(ng);
        for (long long g = 0; g < ng; g++) {
            ni[g] = t->groups_h[g].y;
            nje[g] = t->groups_h[g].z;
            njs[g] = t->groups_h[g].w;
        }
        const pth_params &par = ctx->par;
        t->pre_key = PlanKey{par.exec_mode, par.chunk_j, par.max_ipl,
                             par.exec_mode == PTH_EXEC_FUSED ? par.segments : 1, par.spj_monopole,
                             seg_pairs(par)};
        t->pre_ok = make_plan((int)ng, ni.data(), nje.data(), njs.data(), t->pre_key.exec_mode, t->pre_key.chunk_j,
                              t->pre_key.max_ipl, t->pre_plan, t->pre_key.n_seg, t->pre_key.mono != 0,
                              t->pre_key.min_seg_pairs);
        t->pre_valid = true;
    }
    if ((e = hipGetLastError()) || (e = hipStreamSynchronize(st))) return ctx->hip_fail(e, "tree walk");
    const auto t3 = std::chrono::steady_clock::now();
    t->ms_order = std::chrono::duration<float, std::milli>(t3 - t2).count();
    t->built = true;
    return PTH_OK;
}

// (Re)fills resident r from tree t: plan, buffers, records and lists.
int resident_fill_gtree(pth_resident *r, const pth_gtree *t, const char *who)
{
    pth_ctx *c = r->ctx;
    const std::string w(who);
    if (!t) return c->fail(PTH_ERR_ARG, w + ": null tree");
    if (!t->built) return c->fail(PTH_ERR_STATE, w + ": the last build of the tree failed");
    if (t->ctx->device != c->device) return c->fail(PTH_ERR_ARG, w + ": the tree lives on another device");
    if (t->n_node > 0x7fffffffLL) return c->fail(PTH_ERR_CAPACITY, w + ": node count exceeds the int32 range");
    const long long ng = t->n_group;
    std::vector<int> ni(ng), ne(ng), ns(ng);
    long long next = 0;
    for (long long g = 0; g < ng; g++) {
        if (t->groups_h[g].x != next)
            return c->fail(PTH_ERR_STATE, w + ": the groups do not tile the particles");
        next += t->groups_h[g].y;
        ni[g] = t->groups_h[g].y;
        ne[g] = t->groups_h[g].z;
        ns[g] = t->groups_h[g].w;
    }
    // the plan made during the build, when it was made for this context's
    // parameters (then taken over: a second refill plans again)
    const pth_params &par = c->par;
    const PlanKey key{par.exec_mode, par.chunk_j, par.max_ipl, par.exec_mode == PTH_EXEC_FUSED ? par.segments : 1,
                      par.spj_monopole, seg_pairs(par)};
    const bool take = t->pre_valid && t->pre_key == key;
    const int rc = resident_setup(r, (int)ng, ni.data(), ne.data(), ns.data(), t->n_le, t->n_ls, (int)t->n,
                                  (int)t->n_node, who, take ? &t->pre_plan : nullptr, t->pre_ok);
    if (take) t->pre_valid = false;
    if (rc != PTH_OK) return rc;
    const Plan &P = r->plan;
    const Arena &A = r->A;
    auto bail = [&](hipError_t e, const char *what) { return c->hip_fail(e, what); };
    if (P.ni_tot != t->n) return c->fail(PTH_ERR_STATE, w + ": the groups do not cover the particles");
    hipStream_t st = c->stream;
    hipError_t e;
    char *d = r->arena.as<char>();
    // plan tables from the host, lists and records on the device
    auto h2d = [&](size_t off, const void *h, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpyAsync(d + off, h, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
    };
    if ((e = h2d(A.walks, P.walks.data(), sizeof(WalkDev) * P.walks.size())) ||
        (e = h2d(A.es, P.es_off.data(), sizeof(long long) * P.es_off.size())) ||
        (e = h2d(A.ss, P.ss_off.data(), sizeof(long long) * P.ss_off.size())) ||
        (e = h2d(A.items, P.items.data(), sizeof(Item) * P.items.size())) ||
        (e = h2d(A.ritems, P.ritems.data(), sizeof(RItem) * P.ritems.size())))
        return bail(e, "resident upload");
    if ((t->n_le && (e = hipMemcpyAsync(d + A.ide, t->id_epj.p, 4 * (size_t)t->n_le, hipMemcpyDeviceToDevice, st))) ||
        (t->n_ls && (e = hipMemcpyAsync(d + A.ids, t->id_spj.p, 4 * (size_t)t->n_ls, hipMemcpyDeviceToDevice, st))))
        return bail(e, "resident lists");
    hipLaunchKernelGGL(k_res_particles, dim3(blocks(t->n)), dim3(256), 0, st, t->n, t->spos.as<double>(),
                       t->smass.as<double>(), t->srs.as<double>(), reinterpret_cast<EpiDev *>(d + A.epi),
                       r->epj.as<EpjDev>());
    if (P.mono)
        hipLaunchKernelGGL(k_res_nodes_mono, dim3(blocks(t->n_node)), dim3(256), 0, st, t->n_node,
                           t->nmass.as<double>(), t->ncom.as<double>(), r->spj.as<EpjDev>());
    else
        hipLaunchKernelGGL(k_res_nodes, dim3(blocks(t->n_node)), dim3(256), 0, st, t->n_node, t->nmass.as<double>(),
                           t->ncom.as<double>(), t->nquad.as<double>(), r->spj.as<SpjDev>());
    if ((e = hipGetLastError()) || (e = hipStreamSynchronize(st))) return bail(e, "resident records");
    if ((e = resident_finish(r))) return bail(e, "resident SPJ table");
    return PTH_OK;
}

}  // namespace

extern "C" {

int pth_gtree_build(pth_ctx *ctx, pth_gtree **out, long long n, const double *pos, const double *mass,
                    const double *r_search, int n_leaf_limit, int n_group_limit, double theta, int canonical)
{
    if (!ctx) return PTH_ERR_ARG;
    if (!out) return ctx->fail(PTH_ERR_ARG, "pth_gtree_build: null output");
    *out = nullptr;
    if (n <= 0 || n > 0x7fffffffLL) return ctx->fail(PTH_ERR_ARG, "pth_gtree_build: n outside [1, 2^31)");
    if (!pos || !mass) return ctx->fail(PTH_ERR_ARG, "pth_gtree_build: null particle arrays");
    if (n_leaf_limit < 1 || n_group_limit < 1) return ctx->fail(PTH_ERR_ARG, "pth_gtree_build: limits must be >= 1");
    HIPCHK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    pth_gtree *t = new pth_gtree;
    t->ctx = ctx;
    t->n_leaf_limit = n_leaf_limit;
    t->n_group_limit = n_group_limit;
    t->theta = theta;
    t->canonical = canonical != 0;
    const int rc = gtree_run(t, n, pos, mass, r_search, false);
    if (rc != PTH_OK) {
        pth_gtree_free(t);
        return rc;
    }
    *out = t;
    return PTH_OK;
}

int pth_gtree_update(pth_gtree *t, long long n, const double *pos, const double *mass, const double *r_search,
                     int on_device)
{
    if (!t) return PTH_ERR_ARG;
    pth_ctx *ctx = t->ctx;
    if (n <= 0 || n > 0x7fffffffLL) return ctx->fail(PTH_ERR_ARG, "pth_gtree_update: n outside [1, 2^31)");
    if (!pos || !mass) return ctx->fail(PTH_ERR_ARG, "pth_gtree_update: null particle arrays");
    HIPCHK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return gtree_run(t, n, pos, mass, r_search, on_device != 0);
}

int pth_gtree_sizes(const pth_gtree *t, long long *n_node, long long *n_group, long long *n_epj_list,
                    long long *n_spj_list)
{
    if (!t) return PTH_ERR_ARG;
    if (!t->built) return t->ctx->fail(PTH_ERR_STATE, "pth_gtree_sizes: the last build failed");
    if (n_node) *n_node = t->n_node;
    if (n_group) *n_group = t->n_group;
    if (n_epj_list) *n_epj_list = t->n_le;
    if (n_spj_list) *n_spj_list = t->n_ls;
    return PTH_OK;
}

int pth_gtree_download(const pth_gtree *t, long long *order, long long *node_begin, long long *node_end,
                       int *node_level, double *node_mass, double *node_com, double *node_quad, long long *epi_begin,
                       int *n_epi, int *n_epj, int *n_spj, int *id_epj, int *id_spj)
{
    if (!t) return PTH_ERR_ARG;
    pth_ctx *c = t->ctx;
    if (!t->built) return c->fail(PTH_ERR_STATE, "pth_gtree_download: the last build failed");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    auto d2h = [&](void *h, const DevBuf &b, size_t bytes) -> hipError_t {
        return (h && bytes) ? hipMemcpy(h, b.p, bytes, hipMemcpyDeviceToHost) : hipSuccess;
    };
    const long long nn = t->n_node, ng = t->n_group;
    HIPCHK(c, d2h(order, t->order, 8 * t->n), "download");
    HIPCHK(c, d2h(node_mass, t->nmass, 8 * nn), "download");
    HIPCHK(c, d2h(node_com, t->ncom, 24 * nn), "download");
    HIPCHK(c, d2h(node_quad, t->nquad, 48 * nn), "download");
    HIPCHK(c, d2h(id_epj, t->id_epj, 4 * t->n_le), "download");
    HIPCHK(c, d2h(id_spj, t->id_spj, 4 * t->n_ls), "download");
    std::vector<GNode> nodes(nn);
    HIPCHK(c, d2h(nodes.data(), t->nodes, sizeof(GNode) * nn), "download");
    for (long long k = 0; k < nn; k++) {
        if (node_begin) node_begin[k] = nodes[k].begin;
        if (node_end) node_end[k] = nodes[k].end;
        if (node_level) node_level[k] = nodes[k].level;
    }
    for (long long g = 0; g < ng; g++) {
        const int4 G = t->groups_h[g];
        if (epi_begin) epi_begin[g] = G.x;
        if (n_epi) n_epi[g] = G.y;
        if (n_epj) n_epj[g] = G.z;
        if (n_spj) n_spj[g] = G.w;
    }
    return PTH_OK;
}

int pth_gtree_timing(const pth_gtree *t, float *ms_build, float *ms_walk, float *ms_order)
{
    if (!t) return PTH_ERR_ARG;
    if (ms_build) *ms_build = t->ms_build;
    if (ms_walk) *ms_walk = t->ms_walk;
    if (ms_order) *ms_order = t->ms_order;
    return PTH_OK;
}

int pth_gtree_stats(const pth_gtree *t, int *walk_passes, int *tile_walk_passes, int *filter_passes)
{
    if (!t) return PTH_ERR_ARG;
    if (walk_passes) *walk_passes = t->walk_passes;
    if (tile_walk_passes) *tile_walk_passes = t->tile_passes;
    if (filter_passes) *filter_passes = t->ngb_passes;
    return PTH_OK;
}

int pth_resident_from_gtree(pth_ctx *c, pth_resident **out, const pth_gtree *t)
{
    if (!c) return PTH_ERR_ARG;
    if (!out) return c->fail(PTH_ERR_ARG, "pth_resident_from_gtree: null output");
    *out = nullptr;
    pth_resident *r = new pth_resident;
    r->ctx = c;
    const int rc = resident_fill_gtree(r, t, "pth_resident_from_gtree");
    if (rc != PTH_OK) {
        pth_resident_free(r);
        return rc;
    }
    *out = r;
    return PTH_OK;
}

int pth_resident_refill(pth_resident *r, const pth_gtree *t)
{
    if (!r) return PTH_ERR_ARG;
    HIPCHK(r->ctx, hipSetDevice(r->ctx->device), "hipSetDevice");
    HIPCHK(r->ctx, hipStreamSynchronize(r->ctx->stream), "hipStreamSynchronize");
    return resident_fill_gtree(r, t, "pth_resident_refill");
}

int pth_gtree_neighbors(pth_gtree *t, long long *n_total)
{
    if (!t) return PTH_ERR_ARG;
    pth_ctx *c = t->ctx;
    if (!t->built) return c->fail(PTH_ERR_STATE, "pth_gtree_neighbors: the last build failed");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    hipStream_t st = c->stream;
    const auto t0 = std::chrono::steady_clock::now();
    const long long n = t->n;
    size_t tb = 0;
    hipError_t e;
    // 1. candidate lists: one overlap-only walk per tile (the largest cells
    //    with at most 64 particles; theta opening off: inv_theta2 = 0), the
    //    group walk's symmetric r_search tests on widened tile boxes
    long long nt = 0;
    {
        const int rc = select_cells(t, 64, t->tiles, nt);
        if (rc != PTH_OK) return rc;
    }
    HIPCHK(c, t->tbox.ensure(sizeof(double) * 12 * nt), "neighbour buffers");
    HIPCHK(c, t->tcnt.ensure(sizeof(long long) * (nt + 1)), "neighbour buffers");
    HIPCHK(c, t->toff.ensure(sizeof(long long) * (nt + 1)), "neighbour buffers");
    HIPCHK(c, t->tinfo.ensure(sizeof(int4) * nt), "neighbour buffers");
    hipLaunchKernelGGL(k_tile_boxes, dim3(blocks(nt)), dim3(256), 0, st, t->tiles.as<int>(), nt,
                       t->box_in.as<double>(), t->box_out.as<double>(), t->rs_max, t->tbox.as<double>());
    WalkArgs wa{};
    wa.nodes = t->nodes.as<GNode>();
    wa.ncom = t->ncom.as<double>();
    wa.box_in = t->box_in.as<double>();
    wa.box_out = t->box_out.as<double>();
    wa.groups = t->tiles.as<int>();
    wa.n_group = nt;
    wa.tbox = t->tbox.as<double>();
    wa.root_len = t->root_len;
    wa.inv_theta2 = 0.0;
    wa.theta_pos = 1;
    wa.cnt_e = t->tcnt.as<long long>();
    const unsigned wgt = (unsigned)std::min<long long>(nt, 65536);
    // as many tiles as the last search (the same tree, or one rebuilt from
    // particles that moved a little): one writing pass into regions of the
    // last search's candidate counts + 1/4 + 64 that also counts; otherwise
    // a counting pass first (a region that turns out too small means an
    // exact writing pass, as for the group walk)
    const bool one_pass = nt > 0 && (long long)t->tcnt_h.size() == nt + 1;
    if (one_pass) {
        std::vector<long long> &ph = t->tpad_h;
        ph.assign(nt + 1, 0);
        long long pe = 0;
        for (long long g = 0; g < nt; g++) {
            const long long k = std::max(t->tcnt_h[g], 0LL);
            ph[g] = pe;
            pe += k + k / 4 + 64;
        }
        ph[nt] = pe;
        HIPCHK(c, t->tpad_off.ensure(8 * (size_t)(nt + 1)), "neighbour buffers");
        HIPCHK(c, t->tl_pad.ensure(4 * (size_t)pe), "neighbour buffers");
        HIPCHK(c, hipMemcpyAsync(t->tpad_off.p, ph.data(), 8 * (size_t)(nt + 1), hipMemcpyHostToDevice, st),
               "neighbour regions");
        wa.off_e = t->tpad_off.as<long long>();
        wa.id_epj = t->tl_pad.as<int>();
        wa.padded = 1;
        hipLaunchKernelGGL(k_walk<true>, dim3(wgt), dim3(64), 0, st, wa);
    } else if (nt > 0) {
        hipLaunchKernelGGL(k_walk<false>, dim3(wgt), dim3(64), 0, st, wa);
    }
    const long long zero = 0;
    HIPCHK(c, hipMemcpyAsync(t->tcnt.as<long long>() + nt, &zero, 8, hipMemcpyHostToDevice, st), "neighbour counts");
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, t->tcnt.as<long long>(), t->toff.as<long long>(),
                                              (int)(nt + 1), st)) ||
        (e = cub_temp(t, tb)) ||
        (e = hipcub::DeviceScan::ExclusiveSum(t->tmp.p, tb, t->tcnt.as<long long>(), t->toff.as<long long>(),
                                              (int)(nt + 1), st)))
        return c->hip_fail(e, "candidate scan");
    long long ncand = 0;
    t->tcnt_h.resize(nt + 1);
    HIPCHK(c, hipMemcpyAsync(&ncand, t->toff.as<long long>() + nt, 8, hipMemcpyDeviceToHost, st), "candidates");
    HIPCHK(c, hipMemcpyAsync(t->tcnt_h.data(), t->tcnt.p, 8 * (size_t)(nt + 1), hipMemcpyDeviceToHost, st),
           "candidates");
    HIPCHK(c, hipStreamSynchronize(st), "candidates");
    if (ncand > 0x7fffffffLL) {
        t->tcnt_h.clear();
        return c->fail(PTH_ERR_CAPACITY, "pth_gtree_neighbors: candidates exceed the int32 range");
    }
    // did every tile's candidates fit its region?
    bool rewalk = !one_pass;
    for (long long g = 0; g < nt && !rewalk; g++) rewalk = t->tcnt_h[g] > t->tpad_h[g + 1] - t->tpad_h[g];
    t->tile_passes = nt > 0 ? (rewalk ? 2 : 1) : 0;
    HIPCHK(c, t->tl.ensure(sizeof(int) * std::max<long long>(ncand, 1)), "neighbour buffers");
    HIPCHK(c, t->tl_alt.ensure(sizeof(int) * std::max<long long>(ncand, 1)), "neighbour buffers");
    wa.off_e = t->toff.as<long long>();
    wa.id_epj = t->tl.as<int>();
    wa.padded = 0;
    if (nt > 0 && rewalk) {
        hipLaunchKernelGGL(k_walk<true>, dim3(wgt), dim3(64), 0, st, wa);
    } else if (nt > 0) {
        // every tile fits its region (checked above)
        hipLaunchKernelGGL(k_compact, dim3(wgt), dim3(256), 0, st, nt, t->tpad_off.as<long long>(),
                           t->toff.as<long long>(), t->tl_pad.as<int>(), t->tl.as<int>());
    }
    // candidates ascending in tree order, so rows come out in that order
    if (ncand > 0) {
        int bits = 1;
        while (bits < 31 && (1LL << bits) < n) bits++;
        const long long *ob = t->toff.as<long long>();
        if ((e = hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, t->tl.as<int>(), t->tl_alt.as<int>(),
                                                            (int)ncand, (int)nt, ob, ob + 1, 0, bits, st)) ||
            (e = cub_temp(t, tb)) ||
            (e = hipcub::DeviceSegmentedRadixSort::SortKeys(t->tmp.p, tb, t->tl.as<int>(), t->tl_alt.as<int>(),
                                                            (int)ncand, (int)nt, ob, ob + 1, 0, bits, st)))
            return c->hip_fail(e, "candidate ordering");
        std::swap(t->tl, t->tl_alt);
    }
    hipLaunchKernelGGL(k_tile_info, dim3(blocks(nt)), dim3(256), 0, st, t->nodes.as<GNode>(), t->tiles.as<int>(), nt,
                       t->toff.as<long long>(), t->tinfo.as<int4>());
    // 2. exact FP64 filter of the candidates.  After a completed search of
    //    as many particles: one writing pass into regions of that search's
    //    counts + 1/4 + 4 that also counts (their total is at most 1.25 times
    //    its total + 4 n, since ngb_cnt holds exactly its counts), then the
    //    move into place, or an exact writing pass if some particle did not
    //    fit.  Otherwise: count, scan, write.
    const long long prev_total = t->ngb_prev_total;
    // (the region scan accumulates in int: its total must stay below 2^31)
    const bool one_pass_ngb = t->ngb_prev_ok && t->ngb_prev_n == n && n > 0 &&
                              prev_total + prev_total / 4 + 4 * n + 64 < 0x7fffffffLL;
    t->ngb_prev_ok = false;  // ngb_cnt is rewritten below
    HIPCHK(c, t->ngb_cnt.ensure(sizeof(int) * (n + 1)), "neighbour buffers");
    HIPCHK(c, t->ngb_off.ensure(sizeof(long long) * (n + 1)), "neighbour buffers");
    NgbArgs na{};
    na.spos = t->spos.as<double>();
    na.srs = t->srs.as<double>();
    na.order = t->order.as<long long>();
    na.info = t->tinfo.as<int4>();
    na.off_e = t->toff.as<long long>();
    na.id_epj = t->tl.as<int>();
    na.n_group = nt;
    na.cnt = t->ngb_cnt.as<int>();
    if (one_pass_ngb) {
        const long long bound = prev_total + prev_total / 4 + 4 * n + 64;
        HIPCHK(c, t->ngb_cap.ensure(sizeof(int) * (n + 1)), "neighbour buffers");
        HIPCHK(c, t->ngb_poff.ensure(sizeof(long long) * (n + 1)), "neighbour buffers");
        HIPCHK(c, t->ngb_pad.ensure(sizeof(int) * (size_t)bound), "neighbour buffers");
        HIPCHK(c, t->ngb_flag.ensure(sizeof(int)), "neighbour buffers");
        hipLaunchKernelGGL(k_ngb_caps, dim3(blocks(n + 1)), dim3(256), 0, st, n, t->ngb_cnt.as<int>(),
                           t->ngb_cap.as<int>());
        if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, t->ngb_cap.as<int>(), t->ngb_poff.as<long long>(),
                                                  (int)(n + 1), st)) ||
            (e = cub_temp(t, tb)) ||
            (e = hipcub::DeviceScan::ExclusiveSum(t->tmp.p, tb, t->ngb_cap.as<int>(), t->ngb_poff.as<long long>(),
                                                  (int)(n + 1), st)))
            return c->hip_fail(e, "neighbour regions");
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, t->ngb_flag.as<int>(), 1, 0);
        na.off = t->ngb_poff.as<long long>();
        na.ngb = t->ngb_pad.as<int>();
        na.padded = 1;
        na.flag = t->ngb_flag.as<int>();
    }
    // the count of entry n stays 0 so the scan's last entry is the total
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, t->ngb_cnt.as<int>() + n, 1, 0);
    if (one_pass_ngb) hipLaunchKernelGGL(k_ngb<true>, dim3(wgt), dim3(64), 0, st, na);
    else hipLaunchKernelGGL(k_ngb<false>, dim3(wgt), dim3(64), 0, st, na);
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, t->ngb_cnt.as<int>(), t->ngb_off.as<long long>(),
                                              (int)(n + 1), st)) ||
        (e = cub_temp(t, tb)) ||
        (e = hipcub::DeviceScan::ExclusiveSum(t->tmp.p, tb, t->ngb_cnt.as<int>(), t->ngb_off.as<long long>(),
                                              (int)(n + 1), st)))
        return c->hip_fail(e, "neighbour scan");
    long long total = 0;
    int over = 0;
    HIPCHK(c, hipMemcpyAsync(&total, t->ngb_off.as<long long>() + n, 8, hipMemcpyDeviceToHost, st), "neighbour total");
    if (one_pass_ngb)
        HIPCHK(c, hipMemcpyAsync(&over, t->ngb_flag.p, sizeof(int), hipMemcpyDeviceToHost, st), "neighbour total");
    HIPCHK(c, hipStreamSynchronize(st), "neighbour count");
    if (total > 0x7fffffffLL) return c->fail(PTH_ERR_CAPACITY, "pth_gtree_neighbors: lists exceed the int32 range");
    HIPCHK(c, t->ngb.ensure(sizeof(int) * std::max<long long>(total, 1)), "neighbour buffers");
    const bool rewrite = !one_pass_ngb || over != 0;
    if (total > 0 && rewrite) {
        na.off = t->ngb_off.as<long long>();
        na.ngb = t->ngb.as<int>();
        na.padded = 0;
        hipLaunchKernelGGL(k_ngb<true>, dim3(wgt), dim3(64), 0, st, na);
    } else if (total > 0) {
        // every particle fits its region (no flag raised)
        hipLaunchKernelGGL(k_ngb_compact, dim3(blocks(n)), dim3(256), 0, st, n, t->ngb_poff.as<long long>(),
                           t->ngb_off.as<long long>(), t->ngb_pad.as<int>(), t->ngb.as<int>());
    }
    HIPCHK(c, hipGetLastError(), "neighbour search");
    HIPCHK(c, hipStreamSynchronize(st), "neighbour search");
    t->ngb_passes = rewrite ? 2 : 1;
    t->ngb_prev_ok = true;
    t->ngb_prev_n = n;
    t->ngb_prev_total = total;
    t->n_ngb = total;
    t->n_cand = ncand;
    t->ms_ngb = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (n_total) *n_total = total;
    return PTH_OK;
}

int pth_gtree_neighbors_download(const pth_gtree *t, int *count, long long *offset, int *ngb, float *ms)
{
    if (!t) return PTH_ERR_ARG;
    pth_ctx *c = t->ctx;
    if (!t->built || t->n_ngb < 0)
        return c->fail(PTH_ERR_STATE, "pth_gtree_neighbors_download: no neighbour search since the last build");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    if (count) HIPCHK(c, hipMemcpy(count, t->ngb_cnt.p, sizeof(int) * t->n, hipMemcpyDeviceToHost), "download");
    if (offset)
        HIPCHK(c, hipMemcpy(offset, t->ngb_off.p, sizeof(long long) * (t->n + 1), hipMemcpyDeviceToHost), "download");
    if (ngb && t->n_ngb > 0)
        HIPCHK(c, hipMemcpy(ngb, t->ngb.p, sizeof(int) * t->n_ngb, hipMemcpyDeviceToHost), "download");
    if (ms) *ms = t->ms_ngb;
    return PTH_OK;
}

void pth_gtree_free(pth_gtree *t)
{
    if (!t) return;
    if (t->ctx && t->ctx->device >= 0) {
        (void)hipSetDevice(t->ctx->device);
        if (t->ctx->stream) (void)hipStreamSynchronize(t->ctx->stream);
    }
    DevBuf *all[] = {&t->order, &t->key, &t->skey, &t->sidx, &t->spos, &t->smass, &t->srs, &t->nodes, &t->nmass,
                     &t->ncom, &t->nquad, &t->box_in, &t->box_out, &t->groups, &t->off_e, &t->off_s, &t->id_epj,
                     &t->id_spj, &t->tmp, &t->ngb_cnt, &t->ngb_off, &t->ngb, &t->tiles, &t->tbox, &t->tcnt,
                     &t->toff, &t->tl, &t->tl_alt, &t->tinfo, &t->pad_off, &t->pad_e, &t->pad_s,
                     &t->tpad_off, &t->tl_pad, &t->ngb_cap, &t->ngb_poff, &t->ngb_pad, &t->ngb_flag};
    for (DevBuf *b : all) b->release();
    t->release_scratch();
    if (t->pinned) (void)hipHostFree(t->pinned);
    delete t;
}

}  // extern "C"
