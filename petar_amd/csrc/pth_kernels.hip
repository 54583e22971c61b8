// pth_kernels.hip -- gfx950 kernels of the soft-force path.
//
// Pipeline for one batch of walks (i-groups with their interaction lists):
//   k_spj_table   per superparticle, once per send: {m, q'xx q'yy q'zz q'xy
//                 q'yz q'zx, -eps^2 tr}, q' = 3Q - tr I
//   k_stage_walk  one workgroup per walk: FP64 bounding box of the walk's i
//                 (centre c, half extents, largest r_search), every i as FP32
//                 (x-c, y-c, z-c, rs^2), and the walk's EPJ list gathered
//                 into FP32 records relative to c, split stably into a "near"
//                 region (entries that may lie within max(rs_i, rs_j, r_out)
//                 of some i of the walk) at the front and a "far" region at
//                 the back; records are SoA blocks of 8 {x[8] y[8] z[8] m[8]
//                 rs2[8]}, gaps padded with far massless entries; beside it,
//                 one or more workgroups per walk gather the walk's SPJ list
//                 into 4-float records {x-c y-c z-c table index}
//   k_force       one wave per work item (walk, tile of 64/128/256 i with
//                 1/2/4 i per lane, j-range), one launch per i-per-lane
//                 class: j records are read through the constant address
//                 space (uniform addresses, scalar loads into SGPRs) and used
//                 as SGPR operands by every i of the lane (the VC variant
//                 copies them into VGPRs first; on gfx950 that copy issues at
//                 single rate and costs more than it saves).  Far EPJ blocks
//                 skip the neighbour count and the cutoff (both are known to
//                 be no-ops there).  FP32 pair arithmetic, FP32 partial sums
//                 per chunk, FP64 running sums.
//   k_reduce      fixed-order FP64 sum of the partial results of split walks
//   k_force_plain the plain per-walk form (first kernel of this build): one
//                 workgroup per walk, 64-i tiles, SGPR operands, whole list
//
// Pair formulas (the physics of src/force_gpu_cuda.cu:50-77 and 276-323):
//   EPJ  dx = xj - xi, r2 = |dx|^2 + eps^2, nnb += r2 < max(rs_i,rs_j)^2,
//        rinv = rsqrt(max(r2, r_out^2)), acc += m rinv^3 dx, pot -= m rinv
//   SPJ  dx = xi - xs, r2 = |dx|^2 + eps^2, traceless form
//        rqr = dx.q'.dx - eps^2 tr, meff = m + rqr/(2 r^4),
//        meff3 = (m + 5 rqr/(2 r^4))/r^3,
//        acc += q'.dx/r^5 - meff3 dx, pot -= meff/r
//   (the same quadrupole expansion as the raw-moment form of the CUDA
//   kernel, rearranged as in src/phantomquad_for_p3t_x86.hpp:144-163)
// G multiplies the FP64 sums once at the end.
#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>

#include "pth_device.h"

namespace pth {

typedef const __attribute__((address_space(4))) float *cfp;
typedef const __attribute__((address_space(4))) int *cip;

__device__ __forceinline__ cfp as_const(const float *p) { return (cfp)(p); }
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// With VC, copies a wave-uniform (SGPR) value into a VGPR, so that the
// arithmetic that uses it has VGPR operands only; otherwise the SGPR value
// is used as an operand directly.
template <bool VC> __device__ __forceinline__ float jv(float s)
{
    if (!VC) return s;
    float v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return v;
}
__device__ __forceinline__ float vcopy(float s) { return jv<true>(s); }

// ---------------------------------------------------------------------------
// staging

__device__ __forceinline__ void put_epj(float *se, long long slot, float x, float y, float z, float m,
                                        float r2)
{
    float *blk = se + (slot >> 3) * 40;
    const int l = (int)(slot & 7);
    blk[l] = x;
    blk[8 + l] = y;
    blk[16 + l] = z;
    blk[24 + l] = m;
    blk[32 + l] = r2;
}

// Staged SPJ record {x-c, y-c, z-c, index of the superparticle's SpjTab
// entry}: only the position depends on the walk.
__device__ __forceinline__ void put_spj(float4 *o, double x, double y, double z, const double (&c)[3], int id)
{
    *o = make_float4((float)(x - c[0]), (float)(y - c[1]), (float)(z - c[2]), __int_as_float(id));
}

// The SPJ table: m, q' = 3Q - tr I (the traceless rearrangement of
// src/phantomquad_for_p3t_x86.hpp:144-163) and -eps^2 tr, from the FP32
// moments in FP64 and rounded once.
__global__ __launch_bounds__(256) void k_spj_table(const SpjDev *__restrict__ spj, long long n, double eps2,
                                                   SpjTab *__restrict__ tab)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SpjDev p = spj[i];
    const double qxx = p.qxx, qyy = p.qyy, qzz = p.qzz;
    const double tr = qxx + qyy + qzz;
    SpjTab t;
    t.m = p.m;
    t.qxx = (float)(3.0 * qxx - tr);
    t.qyy = (float)(3.0 * qyy - tr);
    t.qzz = (float)(3.0 * qzz - tr);
    t.qxy = (float)(3.0 * (double)p.qxy);
    t.qyz = (float)(3.0 * (double)p.qyz);
    t.qzx = (float)(3.0 * (double)p.qxz);
    t.mtr = (float)(-eps2 * tr);
    tab[i] = t;
}

// bucket of a particle id in the per-walk LDS table (PTH_CO_HASH buckets)
__device__ __forceinline__ unsigned co_hash(long long id)
{
    const unsigned long long u = (unsigned long long)id;
    return ((unsigned)(u ^ (u >> 32)) * 2654435761u) >> 21;
}

struct StageArgs {
    int walk_begin;  // block b < n_walk stages walk walk_begin + b
    // blocks n_walk + p n_walk + v (p < spj_parts) stage part p of the SPJ
    // list of walk walk_begin + v
    int n_walk, spj_parts;
    const EpiDev *epi;
    const WalkDev *walks;
    const long long *es_off, *ss_off;
    const int *id_epj, *id_spj;
    const EpjDev *epj;
    const SpjDev *spj;
    double r_out, eps2;
    double4 *centre;
    int *near_blk;
    float4 *staged_i;
    float *staged_e, *staged_s;
    // changeover mode
    const long long *epi_id;
    const EpjCoDev *epj_co;
    float4 *staged_ic;
    float4 *staged_il;
    float *staged_ec;
    int *walk_err;
    // linear-cutoff mode with double-float near pairs
    int df;
    float *staged_el;
    // monopole superparticles (centre and mass), staged as far EPJ entries
    int mono;
    const EpjDev *spj_mono;
    // persistent force launches: the segment's item counters (zeroed here,
    // before the segment's force launches), or null
    int *counters;
};

// SPJ list entries [k_begin, k_end) of walk w: positions relative to the
// centre and table indices, U per thread and pass so that U gathers are in
// flight
template <int U>
__device__ __forceinline__ void stage_spj(const StageArgs &a, int w, const WalkDev &wk, const double (&c)[3],
                                          int k_begin, int k_end)
{
    float4 *out_s = reinterpret_cast<float4 *>(a.staged_s) + a.ss_off[w];
    const int *ids = a.id_spj + wk.s_off;
    const SpjDev *__restrict__ spj = a.spj;
    const int tid = (int)threadIdx.x;
    for (int k0 = k_begin; k0 < k_end; k0 += 256 * U) {
        int id[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + 256 * u + tid;
            id[u] = k < k_end ? ids[k] : -1;
        }
        double px[U], py[U], pz[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (id[u] >= 0) {
                px[u] = spj[id[u]].x;
                py[u] = spj[id[u]].y;
                pz[u] = spj[id[u]].z;
            }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (id[u] >= 0) put_spj(out_s + (k0 + 256 * u + tid), px[u], py[u], pz[u], c, id[u]);
    }
}

// CO (changeover mode): every near EPJ also gets its changeover radii and the
// low-order parts of its FP32 relative coordinates staged (per block of 8:
// r_in[8], r_out[8], xlo[8], ylo[8], zlo[8]; x - c = x_hi + x_lo as a
// double-float), every i the low-order parts of its coordinates, and each i
// finds its own EPJ entry
// (same id) among the walk's near entries through an LDS hash table that
// maps the hash of an id to the i's index (the ids themselves are compared
// in place, in global memory); the i's staged slot of that entry and its radii go to
// staged_ic (the force kernel drops the self pair by slot and takes the i's
// changeover radii from there).  walk_err counts i with no such entry.
// EU: EPJ entries per thread and pass of the near/far split, all gathers in
// flight at once: 4 where the launch has few walks (its latency is exposed),
// 2 where it has many (fewer registers, so that staging waves fit beside the
// force kernels of other segments) and in changeover mode
template <bool CO, int EU>
__global__ __launch_bounds__(256) void k_stage_walk(const StageArgs a)
{
    // the SPJ blocks need only the walk's centre: they run beside the EPJ
    // split, which is sequential within a walk
    const int nw = a.n_walk;
    const bool spj_wg = (int)blockIdx.x >= nw;
    const int w = a.walk_begin + (spj_wg ? ((int)blockIdx.x - nw) % nw : (int)blockIdx.x);
    const WalkDev wk = a.walks[w];
    if (blockIdx.x == 0 && a.counters && threadIdx.x < PLAN_CLASSES * 8) a.counters[threadIdx.x] = 0;
    if (spj_wg && (a.mono || wk.n_s == 0)) return;
    const EpiDev *__restrict__ epi = a.epi;
    const EpjDev *__restrict__ epj = a.epj;
    const int *__restrict__ id_epj = a.id_epj;
    const double r_out = a.r_out;
    float4 *__restrict__ staged_i = a.staged_i;
    float *__restrict__ staged_e = a.staged_e;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __shared__ double red[4][7];
    __shared__ int cnt[2 * EU][4];
    __shared__ int co_idx[CO ? PTH_CO_HASH : 1];
    __shared__ int co_self[CO ? PTH_CO_MAX_NI : 1];
    __shared__ float co_rin[CO ? PTH_CO_MAX_NI : 1], co_rout[CO ? PTH_CO_MAX_NI : 1], co_m[CO ? PTH_CO_MAX_NI : 1];
    const int n_co = (CO && !spj_wg) ? min(wk.n_i, PTH_CO_MAX_NI) : 0;
    if constexpr (CO) {
        if (!spj_wg) {
            for (int k = tid; k < PTH_CO_HASH; k += 256) co_idx[k] = -1;
            __syncthreads();
            for (int k = tid; k < n_co; k += 256) {
                const long long id = a.epi_id[wk.i_off + k];
                unsigned h = co_hash(id);
                // at most PTH_CO_MAX_NI keys in PTH_CO_HASH buckets: a free one exists
                while (atomicCAS(&co_idx[h], -1, k) != -1) h = (h + 1) & (PTH_CO_HASH - 1);
                co_self[k] = -1;
                co_rin[k] = 0.f;
                co_rout[k] = 0.f;
                co_m[k] = 0.f;
            }
            __syncthreads();
        }
    }
    // -- bounding box and largest r_search of the walk's i
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rsm = 0.0;
    for (int k = tid; k < wk.n_i; k += 256) {
        const EpiDev e = epi[wk.i_off + k];
        lo[0] = fmin(lo[0], e.x); hi[0] = fmax(hi[0], e.x);
        lo[1] = fmin(lo[1], e.y); hi[1] = fmax(hi[1], e.y);
        lo[2] = fmin(lo[2], e.z); hi[2] = fmax(hi[2], e.z);
        rsm = fmax(rsm, (double)e.rs);
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
        }
        rsm = fmax(rsm, __shfl_xor(rsm, off, 64));
    }
    if (lane == 0) {
        for (int d = 0; d < 3; d++) { red[wave][d] = lo[d]; red[wave][3 + d] = hi[d]; }
        red[wave][6] = rsm;
    }
    __syncthreads();
    double c[3], h[3];
    for (int d = 0; d < 3; d++) {
        const double l = fmin(fmin(red[0][d], red[1][d]), fmin(red[2][d], red[3][d]));
        const double u = fmax(fmax(red[0][3 + d], red[1][3 + d]), fmax(red[2][3 + d], red[3][3 + d]));
        c[d] = (wk.n_i > 0) ? 0.5 * (l + u) : 0.0;
        h[d] = (wk.n_i > 0) ? 0.5 * (u - l) : 0.0;
    }
    rsm = fmax(fmax(red[0][6], red[1][6]), fmax(red[2][6], red[3][6]));
    if (spj_wg) {
        const int part = ((int)blockIdx.x - nw) / nw;
        const int k_begin = (int)((long long)wk.n_s * part / a.spj_parts);
        const int k_end = (int)((long long)wk.n_s * (part + 1) / a.spj_parts);
        // (two in changeover mode: fewer registers, so the staging of the
        // next segment fits beside the changeover force kernels, which are
        // held to 7 waves per SIMD)
        stage_spj<CO ? 2 : 4>(a, w, wk, c, k_begin, k_end);
        return;
    }
    if (tid == 0) a.centre[w] = make_double4(c[0], c[1], c[2], 0.0);
    for (int k = tid; k < wk.n_i; k += 256) {
        const EpiDev e = epi[wk.i_off + k];
        const double rs = (double)e.rs;
        const double dx = e.x - c[0], dy = e.y - c[1], dz = e.z - c[2];
        const float hx = (float)dx, hy = (float)dy, hz = (float)dz;
        staged_i[wk.i_off + k] = make_float4(hx, hy, hz, (float)(rs * rs));
        if (CO || a.df)
            a.staged_il[wk.i_off + k] =
                make_float4((float)(dx - (double)hx), (float)(dy - (double)hy), (float)(dz - (double)hz), 0.f);
    }
    // -- EPJ list: stable split into near (front) and far (back, reversed);
    //    EU entries per thread and pass (entries t0 + 256 u + tid) so that EU
    //    gathers are in flight per thread between the barriers
    const long long base = a.es_off[w];
    const long long cap = a.es_off[w + 1] - base;
    const double hmax = fmax(h[0], fmax(h[1], h[2]));
    int n_near = 0, n_far = 0;
    struct Ent {
        bool valid, nearj;
        float x, y, z, m, r2, xl, yl, zl;
        int jx;
    };
    auto load = [&](int t) {
        Ent q{t < wk.n_e, false, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0};
        if (q.valid) q.jx = id_epj[(long long)wk.e_off + t];
        return q;
    };
    auto classify = [&](Ent &q) {
        if (!q.valid) return;
        const EpjDev e = epj[q.jx];
        const double dx = e.x - c[0], dy = e.y - c[1], dz = e.z - c[2];
        // near: inside the walk's box widened by max(rs_i, rs_j, r_out),
        // with a margin that covers the FP32 rounding of the pair arithmetic
        const double R = fmax(fmax(rsm, (double)e.rs), r_out);
        const double dm = fmax(fabs(dx), fmax(fabs(dy), fabs(dz)));
        const double Rm = R * 1.00001 + 1e-5 * (dm + hmax);
        q.nearj = fabs(dx) <= h[0] + Rm && fabs(dy) <= h[1] + Rm && fabs(dz) <= h[2] + Rm;
        q.x = (float)dx;
        q.y = (float)dy;
        q.z = (float)dz;
        q.xl = (float)(dx - (double)q.x);
        q.yl = (float)(dy - (double)q.y);
        q.zl = (float)(dz - (double)q.z);
        q.m = e.m;
        const double rs = (double)e.rs;
        q.r2 = (float)(rs * rs);
    };
    auto place = [&](const Ent &q, long long slot) {
        put_epj(staged_e, slot, q.x, q.y, q.z, q.m, q.r2);
        if constexpr (!CO) {
            if (a.df && q.nearj) {
                float *blk = a.staged_el + (slot >> 3) * PTH_DF_EXTRA;
                const int l = (int)(slot & 7);
                blk[l] = q.xl;
                blk[8 + l] = q.yl;
                blk[16 + l] = q.zl;
            }
        }
        if constexpr (CO) {
            if (q.nearj) {
                const EpjCoDev co = a.epj_co[q.jx];
                float *blk = a.staged_ec + (slot >> 3) * PTH_CO_EXTRA;
                const int l = (int)(slot & 7);
                blk[l] = co.rin;
                blk[8 + l] = co.rout;
                blk[16 + l] = q.xl;
                blk[24 + l] = q.yl;
                blk[32 + l] = q.zl;
                blk[40 + l] = co_pot_mass(co, q.m);
                // is this the EPJ of one of the walk's i?
                unsigned hb = co_hash(co.id);
                for (;;) {
                    const int k = co_idx[hb];
                    if (k < 0) break;
                    // (the walk's ids are compared in place, not copied to LDS)
                    if (a.epi_id[wk.i_off + k] == co.id) {
                        co_self[k] = (int)slot;
                        co_rin[k] = co.rin;
                        co_rout[k] = co.rout;
                        co_m[k] = co_self_mass(co, q.m);
                        break;
                    }
                    hb = (hb + 1) & (PTH_CO_HASH - 1);
                }
            }
        }
    };
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int t0 = 0; t0 < wk.n_e; t0 += 256 * EU) {
        Ent q[EU];
#pragma unroll
        for (int u = 0; u < EU; u++) q[u] = load(t0 + 256 * u + tid);
#pragma unroll
        for (int u = 0; u < EU; u++) classify(q[u]);
        unsigned long long bn[EU], bf[EU];
#pragma unroll
        for (int u = 0; u < EU; u++) {
            bn[u] = __ballot(q[u].valid && q[u].nearj);
            bf[u] = __ballot(q[u].valid && !q[u].nearj);
        }
        if (lane == 0) {
#pragma unroll
            for (int u = 0; u < EU; u++) {
                cnt[2 * u][wave] = __popcll(bn[u]);
                cnt[2 * u + 1][wave] = __popcll(bf[u]);
            }
        }
        __syncthreads();
        // the entries of part u (t0 + 256 u + tid) follow all of part u - 1's
#pragma unroll
        for (int u = 0; u < EU; u++) {
            int pn = 0, pf = 0, tn = 0, tf = 0;
            for (int v = 0; v < 4; v++) {
                const int cn = cnt[2 * u][v], cf = cnt[2 * u + 1][v];
                if (v < wave) {
                    pn += cn;
                    pf += cf;
                }
                tn += cn;
                tf += cf;
            }
            if (q[u].valid)
                place(q[u], q[u].nearj ? base + n_near + pn + __popcll(bn[u] & below)
                                       : base + cap - 1 - (n_far + pf + __popcll(bf[u] & below)));
            n_near += tn;
            n_far += tf;
        }
        __syncthreads();
    }
    // monopole superparticles: far entries after the EPJ list's far ones (a
    // monopole pair is a far pair: no cutoff, no neighbour count), in list
    // order from the back
    if (a.mono) {
        const int *__restrict__ mid = a.id_spj + wk.s_off;
        for (int k = tid; k < wk.n_s; k += 256) {
            const EpjDev e = a.spj_mono[mid[k]];
            put_epj(staged_e, base + cap - 1 - (n_far + k), (float)(e.x - c[0]), (float)(e.y - c[1]),
                    (float)(e.z - c[2]), e.m, 0.f);
        }
        n_far += wk.n_s;
    }
    // pad the gap between the two regions (far, massless); the padding of the
    // last near block is read by the double-float near loop: zero low parts
    for (long long s = base + n_near + tid; s < base + cap - n_far; s += 256) {
        put_epj(staged_e, s, 1e18f, 1e18f, 1e18f, 0.f, 0.f);
        if (!CO && a.df) {
            float *blk = a.staged_el + (s >> 3) * PTH_DF_EXTRA;
            const int l = (int)(s & 7);
            blk[l] = blk[8 + l] = blk[16 + l] = 0.f;
        }
    }
    if (tid == 0) a.near_blk[w] = (int)((base + ((n_near + 7) & ~7)) >> 3);
    if constexpr (CO) {
        // (the loop above ended with a barrier: co_self is complete)
        int miss = 0;
        for (int k = tid; k < wk.n_i; k += 256) {
            const bool in = k < n_co;
            const int s = in ? co_self[k] : -1;
            a.staged_ic[wk.i_off + k] =
                make_float4(__int_as_float(s), in ? co_rin[k] : 0.f, in ? co_rout[k] : 0.f, in ? co_m[k] : 0.f);
            miss |= (s < 0);
        }
        const int n_miss = __syncthreads_count(miss);
        if (tid == 0) a.walk_err[w] = n_miss;
    }
}

// ---------------------------------------------------------------------------
// force, fused form

struct Consts {
    float eps2, rout2, half, two_half;
};

// EPJ blocks [b_begin, b_end) of a walk's staged list against the lane's i.
// NEAR blocks also count neighbours and apply the cutoff; with DF (near
// blocks in linear-cutoff mode) each separation is taken from double-float
// walk-relative coordinates, (x_j - x_i) + (xlo_j - xlo_i): the FP32 high
// parts of two close particles differ exactly, so the separation keeps FP32
// relative precision however far the pair lies from the walk centre.
template <int IPL, int NNB_LE, bool NEAR, bool VC, bool DF = false>
__device__ __forceinline__ void epj_blocks(cfp se, int b_begin, int b_end, const float (&xi)[IPL],
                                           const float (&yi)[IPL], const float (&zi)[IPL],
                                           const float (&rsi2)[IPL], const Consts &K, double (&A)[IPL][4],
                                           int (&nnb)[IPL], cfp sel = nullptr, const float *xil = nullptr,
                                           const float *yil = nullptr, const float *zil = nullptr)
{
    for (int b0 = b_begin; b0 < b_end; b0 += PTH_FLUSH_BLOCKS) {
        const int b1 = min(b0 + PTH_FLUSH_BLOCKS, b_end);
        float ax[IPL], ay[IPL], az[IPL], p[IPL];
        int nb[IPL];
#pragma unroll
        for (int q = 0; q < IPL; q++) { ax[q] = ay[q] = az[q] = p[q] = 0.f; nb[q] = 0; }
        for (int b = b0; b < b1; b++) {
            const cfp B = se + (size_t)b * 40;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float xj = jv<VC>(B[k]), yj = jv<VC>(B[8 + k]), zj = jv<VC>(B[16 + k]);
                const float mj = jv<VC>(B[24 + k]);
                float rj = 0.f;
                if (NEAR) rj = jv<VC>(B[32 + k]);
                float xjl = 0.f, yjl = 0.f, zjl = 0.f;
                if (DF) {
                    const cfp E = sel + (size_t)b * PTH_DF_EXTRA;
                    xjl = E[k];
                    yjl = E[8 + k];
                    zjl = E[16 + k];
                }
#pragma unroll
                for (int q = 0; q < IPL; q++) {
                    float dx = xj - xi[q];
                    float dy = yj - yi[q];
                    float dz = zj - zi[q];
                    if (DF) {
                        dx += xjl - xil[q];
                        dy += yjl - yil[q];
                        dz += zjl - zil[q];
                    }
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, K.eps2)));
                    float ri;
                    if (NEAR) {
                        // rs^2 >= 0 (or -1 on idle lanes): integer max of the bit patterns
                        const float rs2 = __int_as_float(max(__float_as_int(rsi2[q]), __float_as_int(rj)));
                        if (NNB_LE) nb[q] += (r2 <= rs2) ? 1 : 0;
                        else nb[q] += (r2 < rs2) ? 1 : 0;
                        ri = __builtin_amdgcn_rsqf(__builtin_fmaxf(r2, K.rout2));
                    } else {
                        ri = __builtin_amdgcn_rsqf(r2);
                    }
                    const float mri = mj * ri;
                    const float ri2 = ri * ri;
                    const float mri3 = mri * ri2;
                    ax[q] = __builtin_fmaf(mri3, dx, ax[q]);
                    ay[q] = __builtin_fmaf(mri3, dy, ay[q]);
                    az[q] = __builtin_fmaf(mri3, dz, az[q]);
                    p[q] -= mri;
                }
                // keep each j's VGPR copies next to their use (no hoisting of
                // the whole block's copies, which costs occupancy)
                if (VC) __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int q = 0; q < IPL; q++) {
            A[q][0] += (double)ax[q];
            A[q][1] += (double)ay[q];
            A[q][2] += (double)az[q];
            A[q][3] += (double)p[q];
            nnb[q] += nb[q];
        }
    }
}

// Changeover weight of one neighbour pair (src/changeover.hpp:294-334 with
// the pair rule of :366-375, i being the first changeover): the Newtonian
// terms weighted by the changeover, f = m (1 - W_acc) / r^3 (acc factor,
// with the j's tree mass m), pg = mp (1 - W_pot) / r (pot_soft) and pt = -mp
// W_pot / r (pot_tot - pot_soft) with its potential mass mp, i.e. what
// src/hard.hpp:1404-1476 leaves for j once the linear-cutoff terms are
// removed (mp = m for a single, the backup mass for a member, 0 for an
// artificial particle).
__device__ __forceinline__ void co_pair(float r2, float mj, float mp, float rin_i, float rout_i, float rin_j,
                                        float rout_j, float &f, float &pg, float &pt)
{
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;  // sqrt(r^2 + eps^2), as dr_eps of src/hard.hpp:1420
    const bool use_i = rout_i > rout_j;
    const float rin = use_i ? rin_i : rin_j;
    const float rout = use_i ? rout_i : rout_j;
    const float xr = (r - rin) * __builtin_amdgcn_rcpf(rout - rin);
    const float coff = (rout - rin) * __builtin_amdgcn_rcpf(rout + rin);
    const float x = __builtin_fminf(__builtin_fmaxf(xr, 0.f), 1.f);
    const float x_1 = x - 1.f, x_2 = x_1 * x_1, x_4 = x_2 * x_2;
    const float wacc = x_4 * __builtin_fmaf(
                                 x, __builtin_fmaf(x, __builtin_fmaf(x, __builtin_fmaf(x, 35.f * coff, 20.f), 10.f), 4.f),
                                 1.f);
    const float x2 = x * x, x5 = x2 * x2 * x;
    float wpot = __builtin_fmaf(-coff * x5,
                                __builtin_fmaf(x, __builtin_fmaf(x, __builtin_fmaf(x, 5.f, -20.f), 28.f), -14.f), 1.f);
    if (xr >= 1.f) wpot = (1.f + coff) * r * __builtin_amdgcn_rcpf(rout);
    f = mj * rinv * rinv * rinv * (1.f - wacc);
    const float mpri = mp * rinv;
    pg = mpri * (1.f - wpot);
    pt = -mpri * wpot;
}

// Near EPJ blocks in changeover mode.  Every pair first gets the
// linear-cutoff terms of src/force_gpu_cuda.cu:50-77 (the same instructions
// as epj_blocks), and is tested for being a neighbour pair: r^2 + eps^2 <
// ngb2 max(rs_i, rs_j)^2 + eps^2 with ngb2 = 0.99^2 (getRSearch() = 0.99
// r_search of src/soft_ptcl.hpp:97-101,349-353 on both sides), self excluded
// (by staged slot).  The j of a block of 8 for which some lane of the wave
// has a neighbour pair are revisited (in order) and, on the neighbour pairs
// only, the linear-cutoff terms are replaced by co_pair's; the other j would
// only add zeros there.
// The self pair's -m_i/r_out is removed by the caller (src/hard.hpp:1669-1673
// adds the same term back on the CPU).
template <int IPL, int NNB_LE>
__device__ __forceinline__ void epj_near_co(cfp se, cfp sec, int b_begin, int b_end, const float (&xi)[IPL],
                                            const float (&yi)[IPL], const float (&zi)[IPL],
                                            const float (&xil)[IPL], const float (&yil)[IPL],
                                            const float (&zil)[IPL], const float (&rsi2)[IPL],
                                            const int (&self)[IPL], const float (&rin_i)[IPL],
                                            const float (&rout_i)[IPL], const Consts &K, float ngb2,
                                            double (&A)[IPL][4], double (&PT)[IPL], int (&nnb)[IPL])
{
    for (int b0 = b_begin; b0 < b_end; b0 += PTH_FLUSH_BLOCKS) {
        const int b1 = min(b0 + PTH_FLUSH_BLOCKS, b_end);
        float ax[IPL], ay[IPL], az[IPL], p[IPL], pt[IPL];
        int nb[IPL];
#pragma unroll
        for (int q = 0; q < IPL; q++) { ax[q] = ay[q] = az[q] = p[q] = pt[q] = 0.f; nb[q] = 0; }
        for (int b = b0; b < b1; b++) {
            const cfp B = se + (size_t)b * 40;
            unsigned hits = 0;  // bit k: some lane has a neighbour pair with j k
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float xj = B[k], yj = B[8 + k], zj = B[16 + k], mj = B[24 + k], rj = B[32 + k];
                const int sj = b * 8 + k;
                bool hit = false;
#pragma unroll
                for (int q = 0; q < IPL; q++) {
                    const float dx = xj - xi[q];
                    const float dy = yj - yi[q];
                    const float dz = zj - zi[q];
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, K.eps2)));
                    const float rs2 = __int_as_float(max(__float_as_int(rsi2[q]), __float_as_int(rj)));
                    if (NNB_LE) nb[q] += (r2 <= rs2) ? 1 : 0;
                    else nb[q] += (r2 < rs2) ? 1 : 0;
                    const float ri = __builtin_amdgcn_rsqf(__builtin_fmaxf(r2, K.rout2));
                    const float mri = mj * ri;
                    const float ri2 = ri * ri;
                    const float mri3 = mri * ri2;
                    ax[q] = __builtin_fmaf(mri3, dx, ax[q]);
                    ay[q] = __builtin_fmaf(mri3, dy, ay[q]);
                    az[q] = __builtin_fmaf(mri3, dz, az[q]);
                    p[q] -= mri;
                    hit = ((r2 < __builtin_fmaf(ngb2, rs2, K.eps2)) && sj != self[q]) ? true : hit;
                }
                if (__ballot(hit)) hits |= 1u << k;
            }
            if (hits) {
                // revisit the j with neighbour pairs
                const cfp E = sec + (size_t)b * PTH_CO_EXTRA;
                for (unsigned mk = (unsigned)uniform((int)hits); mk; mk &= mk - 1) {
                    const int k = uniform(__builtin_ctz(mk));
                    const float xj = B[k], yj = B[8 + k], zj = B[16 + k], mj = B[24 + k], rj = B[32 + k];
                    const float rinj = E[k], routj = E[8 + k];
                    const float xjl = E[16 + k], yjl = E[24 + k], zjl = E[32 + k];
                    const float mpj = E[40 + k];
                    const int sj = b * 8 + k;
#pragma unroll
                    for (int q = 0; q < IPL; q++) {
                        const float dx = xj - xi[q];
                        const float dy = yj - yi[q];
                        const float dz = zj - zi[q];
                        const float r2 =
                            __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, K.eps2)));
                        const float rs2 = __int_as_float(max(__float_as_int(rsi2[q]), __float_as_int(rj)));
                        const bool nbr = (r2 < __builtin_fmaf(ngb2, rs2, K.eps2)) && sj != self[q];
                        // the linear-cutoff terms added above
                        const float ri = __builtin_amdgcn_rsqf(__builtin_fmaxf(r2, K.rout2));
                        const float mri = mj * ri;
                        const float mri3 = mri * (ri * ri);
                        // the pair in double-float relative coordinates: the
                        // high parts are close, so their difference is exact
                        // and the separation keeps full FP32 relative
                        // precision however small it is next to |x - c|
                        const float dxp = dx + (xjl - xil[q]);
                        const float dyp = dy + (yjl - yil[q]);
                        const float dzp = dz + (zjl - zil[q]);
                        const float r2p =
                            __builtin_fmaf(dzp, dzp, __builtin_fmaf(dyp, dyp, __builtin_fmaf(dxp, dxp, K.eps2)));
                        float f, pg, ptj;
                        co_pair(r2p, mj, mpj, rin_i[q], rout_i[q], rinj, routj, f, pg, ptj);
                        ax[q] += nbr ? __builtin_fmaf(f, dxp, -mri3 * dx) : 0.f;
                        ay[q] += nbr ? __builtin_fmaf(f, dyp, -mri3 * dy) : 0.f;
                        az[q] += nbr ? __builtin_fmaf(f, dzp, -mri3 * dz) : 0.f;
                        p[q] -= nbr ? pg - mri : 0.f;
                        pt[q] += nbr ? ptj : 0.f;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < IPL; q++) {
            A[q][0] += (double)ax[q];
            A[q][1] += (double)ay[q];
            A[q][2] += (double)az[q];
            A[q][3] += (double)p[q];
            PT[q] += (double)pt[q];
            nnb[q] += nb[q];
        }
    }
}

// One staged SPJ record held in registers (SGPRs unless VC).
struct SpjRec {
    float x, y, z, m, qxx, qyy, qzz, qxy, qyz, qzx, mtr;
};

// A staged SPJ position and its table index (scalar loads).
struct SpjPos {
    float x, y, z;
    int id;
};
__device__ __forceinline__ SpjPos spj_pos(cfp ss, int s)
{
    const cfp R = ss + (size_t)s * PTH_SPJ_STAGED;
    return SpjPos{R[0], R[1], R[2], __float_as_int(R[3])};
}

// One SPJ record from wave-uniform addresses: its staged position and its
// table entry, scalar loads used as SGPR operands, or with VC copied into
// VGPRs first.
template <bool VC> __device__ __forceinline__ SpjRec spj_rec(const SpjPos &q, cfp tab)
{
    const cfp T = tab + (size_t)q.id * 8;
    return SpjRec{jv<VC>(q.x), jv<VC>(q.y), jv<VC>(q.z), jv<VC>(T[0]), jv<VC>(T[1]), jv<VC>(T[2]),
                  jv<VC>(T[3]),  jv<VC>(T[4]), jv<VC>(T[5]), jv<VC>(T[6]), jv<VC>(T[7])};
}

// One SPJ record against the lane's i (src/force_gpu_cuda.cu:276-323 in the
// traceless form of src/phantomquad_for_p3t_x86.hpp:144-163)
template <int IPL>
__device__ __forceinline__ void spj_pair(const SpjRec &c, const float (&xi)[IPL], const float (&yi)[IPL],
                                         const float (&zi)[IPL], const Consts &K, float (&ax)[IPL],
                                         float (&ay)[IPL], float (&az)[IPL], float (&p)[IPL])
{
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const float dx = xi[q] - c.x;
        const float dy = yi[q] - c.y;
        const float dz = zi[q] - c.z;
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, K.eps2)));
        const float ri = __builtin_amdgcn_rsqf(r2);
        const float ri2 = ri * ri;
        const float ri3 = ri * ri2;
        const float ri4 = ri2 * ri2;
        const float ri5 = ri2 * ri3;
        const float qrx = __builtin_fmaf(c.qzx, dz, __builtin_fmaf(c.qxy, dy, c.qxx * dx));
        const float qry = __builtin_fmaf(c.qyz, dz, __builtin_fmaf(c.qxy, dx, c.qyy * dy));
        const float qrz = __builtin_fmaf(c.qyz, dy, __builtin_fmaf(c.qzx, dx, c.qzz * dz));
        const float rqr = __builtin_fmaf(qrz, dz, __builtin_fmaf(qry, dy, __builtin_fmaf(qrx, dx, c.mtr)));
        const float rqr_ri4 = rqr * ri4;
        const float meff = __builtin_fmaf(rqr_ri4, K.half, c.m);
        const float meff3 = __builtin_fmaf(rqr_ri4, K.two_half, c.m) * ri3;
        p[q] = __builtin_fmaf(-meff, ri, p[q]);
        ax[q] = __builtin_fmaf(-meff3, dx, __builtin_fmaf(ri5, qrx, ax[q]));
        ay[q] = __builtin_fmaf(-meff3, dy, __builtin_fmaf(ri5, qry, ay[q]));
        az[q] = __builtin_fmaf(-meff3, dz, __builtin_fmaf(ri5, qrz, az[q]));
    }
}

template <int IPL>
__device__ __forceinline__ void spj_flush(const float (&ax)[IPL], const float (&ay)[IPL], const float (&az)[IPL],
                                          const float (&p)[IPL], double (&A)[IPL][4])
{
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        A[q][0] += (double)ax[q];
        A[q][1] += (double)ay[q];
        A[q][2] += (double)az[q];
        A[q][3] += (double)p[q];
    }
}

// SPJ records [s_begin, s_end): the next record's table entry and the
// position after it are loaded before the current record is used (scalar
// loads in flight during the pair arithmetic instead of a wait at the top of
// every record; the table index comes from the position loaded one record
// earlier).
template <int IPL, bool VC>
__device__ __forceinline__ void spj_records(cfp ss, cfp tab, int s_begin, int s_end, const float (&xi)[IPL],
                                            const float (&yi)[IPL], const float (&zi)[IPL], const Consts &K,
                                            double (&A)[IPL][4])
{
    if (s_begin >= s_end) return;
    SpjRec c = spj_rec<VC>(spj_pos(ss, s_begin), tab);
    SpjPos pn = spj_pos(ss, min(s_begin + 1, s_end - 1));
    for (int s0 = s_begin; s0 < s_end; s0 += PTH_FLUSH_SPJ) {
        const int s1 = min(s0 + PTH_FLUSH_SPJ, s_end);
        float ax[IPL], ay[IPL], az[IPL], p[IPL];
#pragma unroll
        for (int q = 0; q < IPL; q++) ax[q] = ay[q] = az[q] = p[q] = 0.f;
        for (int s = s0; s < s1; s++) {
            // the next record and the position after it (the last ones are
            // re-read: stays in bounds)
            const SpjPos pnn = spj_pos(ss, min(s + 2, s_end - 1));
            const SpjRec n = spj_rec<VC>(pn, tab);
            spj_pair<IPL>(c, xi, yi, zi, K, ax, ay, az, p);
            c = n;
            pn = pnn;
        }
        spj_flush<IPL>(ax, ay, az, p, A);
    }
}

// Arguments of the fused force launches.
struct ForceArgs {
    const Item *items;
    int n_items;
    // persistent launches: the class's 8 item counters (one per XCD) and the
    // slots of each XCD's run of items
    int *counter;
    int n_run;
    const int *near_blk;
    const float4 *staged_i;
    const float *staged_e, *staged_s;
    const SpjTab *spj_tab;
    const float4 *staged_ic;  // changeover mode
    const float4 *staged_il;  // changeover mode
    const float *staged_ec;   // changeover mode
    const float *staged_el;   // linear mode, double-float near pairs
    float eps2, rout2, ngb2;
    double G;
    double4 *out_accp;
    int *out_nnb;
    double *out_pott;  // changeover mode
    double4 *part_accp;
    int *part_nnb;
    double *part_pott;  // changeover mode
};

template <int IPL, int NNB_LE, bool VC, bool CO, bool DF>
__device__ __forceinline__ void run_tile(const ForceArgs &a, const Item &it, int near_end, int lane,
                                         const Consts &K)
{
    const int n_i = item_n_i(it.n_i_ipl);
    float xi[IPL], yi[IPL], zi[IPL], rsi2[IPL];
    double A[IPL][4], PT[IPL];
    int nnb[IPL];
    int self[IPL];
    float rin_i[IPL], rout_i[IPL], m_i[IPL], xil[IPL], yil[IPL], zil[IPL];
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const bool in = lane + 64 * q < n_i;
        float4 v = make_float4(0.f, 0.f, 0.f, -1.f);
        if (in) v = a.staged_i[it.i_begin + lane + 64 * q];
        xi[q] = v.x;
        yi[q] = v.y;
        zi[q] = v.z;
        rsi2[q] = v.w;
        A[q][0] = A[q][1] = A[q][2] = A[q][3] = 0.0;
        PT[q] = 0.0;
        nnb[q] = 0;
        if constexpr (DF) {
            float4 l = make_float4(0.f, 0.f, 0.f, 0.f);
            if (in) l = a.staged_il[it.i_begin + lane + 64 * q];
            xil[q] = l.x;
            yil[q] = l.y;
            zil[q] = l.z;
        }
        if constexpr (CO) {
            float4 c = make_float4(__int_as_float(-2), 0.f, 0.f, 0.f);
            if (in) c = a.staged_ic[it.i_begin + lane + 64 * q];
            self[q] = __float_as_int(c.x);
            rin_i[q] = c.y;
            rout_i[q] = c.z;
            m_i[q] = c.w;
            float4 l = make_float4(0.f, 0.f, 0.f, 0.f);
            if (in) l = a.staged_il[it.i_begin + lane + 64 * q];
            xil[q] = l.x;
            yil[q] = l.y;
            zil[q] = l.z;
        }
    }
    const cfp se = as_const(a.staged_e);
    const int e_mid = min(max(near_end, it.e_blk_begin), it.e_blk_end);
    if constexpr (CO) {
        epj_near_co<IPL, NNB_LE>(se, as_const(a.staged_ec), it.e_blk_begin, e_mid, xi, yi, zi, xil, yil, zil, rsi2,
                                 self, rin_i, rout_i, K, a.ngb2, A, PT, nnb);
        // remove the self pair's linear-cutoff potential -m_i / max(eps, r_out)
        // (added above with these same FP32 operations) in the item whose
        // j-range holds the i's own entry
#pragma unroll
        for (int q = 0; q < IPL; q++) {
            const int sb = self[q] >> 3;
            if (self[q] >= 0 && sb >= it.e_blk_begin && sb < e_mid)
                A[q][3] += (double)(m_i[q] * __builtin_amdgcn_rsqf(__builtin_fmaxf(K.eps2, K.rout2)));
        }
    } else if constexpr (DF) {
        epj_blocks<IPL, NNB_LE, true, VC, true>(se, it.e_blk_begin, e_mid, xi, yi, zi, rsi2, K, A, nnb,
                                                as_const(a.staged_el), xil, yil, zil);
    } else
        epj_blocks<IPL, NNB_LE, true, VC>(se, it.e_blk_begin, e_mid, xi, yi, zi, rsi2, K, A, nnb);
    epj_blocks<IPL, NNB_LE, false, VC>(se, e_mid, it.e_blk_end, xi, yi, zi, rsi2, K, A, nnb);
    spj_records<IPL, VC>(as_const(a.staged_s), as_const(reinterpret_cast<const float *>(a.spj_tab)), it.s_begin,
                         it.s_end, xi, yi, zi, K, A);
    const double G = a.G;
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = lane + 64 * q;
        if (i >= n_i) continue;
        if (it.part < 0) {
            a.out_accp[it.i_begin + i] = make_double4(G * A[q][0], G * A[q][1], G * A[q][2], G * A[q][3]);
            a.out_nnb[it.i_begin + i] = nnb[q];
            if constexpr (CO) a.out_pott[it.i_begin + i] = G * (A[q][3] + PT[q]);
        } else {
            const size_t slot = (size_t)(it.part + q) * 64 + lane;
            a.part_accp[slot] = make_double4(A[q][0], A[q][1], A[q][2], A[q][3]);
            a.part_nnb[slot] = nnb[q];
            if constexpr (CO) a.part_pott[slot] = A[q][3] + PT[q];
        }
    }
}

// Work item idx of this class's launch (empty slots have walk < 0).
template <int NNB_LE, int IPL, bool VC, bool CO, bool DF>
__device__ __forceinline__ void force_item(const ForceArgs &a, int idx)
{
    const cip ci = (cip)(a.items) + (size_t)idx * 8;
    Item it;
    it.walk = ci[0];
    it.i_begin = ci[1];
    it.n_i_ipl = ci[2];
    it.e_blk_begin = ci[3];
    it.e_blk_end = ci[4];
    it.s_begin = ci[5];
    it.s_end = ci[6];
    it.part = ci[7];
    if (it.walk < 0) return;
    const int near_end = ((cip)(a.near_blk))[it.walk];
    const Consts K{jv<VC>(a.eps2), jv<VC>(a.rout2), jv<VC>(0.5f), jv<VC>(2.5f)};
    run_tile<IPL, NNB_LE, VC, CO, DF>(a, it, near_end, threadIdx.x & 63, K);
}

// One launch per i-per-lane class: items [0, n_items) of this class, one
// per wave; or, persistent (P), one wave per free wave slot: the waves of
// block b (on XCD b % 8, workgroups being handed to the XCDs round robin)
// take the items of that XCD's run from its counter, in plan order (most
// expensive first), until the run is done.
template <int NNB_LE, int IPL, bool VC, bool CO, bool DF, bool P>
__device__ __forceinline__ void force_items(const ForceArgs &a)
{
    if constexpr (P) {
        const int x = (int)blockIdx.x & 7;
        int *cnt = a.counter + x;
        for (;;) {
            int k = 0;
            if ((threadIdx.x & 63) == 0) k = atomicAdd(cnt, 1);
            k = uniform(k);
            if (k >= a.n_run) break;
            force_item<NNB_LE, IPL, VC, CO, DF>(a, ((k >> 2) * 8 + x) * 4 + (k & 3));
        }
    } else {
        const int idx = uniform((int)blockIdx.x * 4 + ((int)threadIdx.x >> 6));
        if (idx >= a.n_items) return;
        force_item<NNB_LE, IPL, VC, CO, DF>(a, idx);
    }
}

template <int NNB_LE, int IPL, bool VC, bool CO, bool DF, bool P = false>
__global__ __launch_bounds__(256) void k_force(const ForceArgs a)
{
    force_items<NNB_LE, IPL, VC, CO, DF, P>(a);
}

// Changeover mode: its extra per-i state and the revisit of neighbour pairs
// leave room for 5 waves per SIMD at 2 i per lane; held to 7 (a few spilled
// registers, on the revisit path) the force phase runs 10% faster.  With 3
// or 4 i per lane the spills would be heavy: no limit there.
template <int NNB_LE, int IPL, bool VC, bool P = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(IPL <= 2 ? 7 : 1))) void k_force_co(const ForceArgs a)
{
    force_items<NNB_LE, IPL, VC, true, false, P>(a);
}

// ---------------------------------------------------------------------------
// force, plain per-walk form: SGPR operands, one i per lane, whole list

// NEAR blocks count neighbours and apply the cutoff; far blocks (beyond
// max(r_search, r_out) of every i of the walk, and monopole superparticles)
// have neither.
template <int NNB_LE, bool NEAR>
__device__ __forceinline__ void plain_epj(cfp se, int b_begin, int b_end, float xi, float yi, float zi,
                                          float rsi2, float eps2, float rout2, double (&A)[4], int &nnb)
{
    for (int b0 = b_begin; b0 < b_end; b0 += PTH_FLUSH_BLOCKS) {
        const int b1 = min(b0 + PTH_FLUSH_BLOCKS, b_end);
        float ax = 0.f, ay = 0.f, az = 0.f, p = 0.f;
        int nb = 0;
        for (int b = b0; b < b1; b++) {
            const cfp B = se + (size_t)b * 40;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float dx = B[k] - xi;
                const float dy = B[8 + k] - yi;
                const float dz = B[16 + k] - zi;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
                float ri;
                if (NEAR) {
                    const float rs2 = __int_as_float(max(__float_as_int(rsi2), __float_as_int(B[32 + k])));
                    if (NNB_LE) nb += (r2 <= rs2) ? 1 : 0;
                    else nb += (r2 < rs2) ? 1 : 0;
                    ri = __builtin_amdgcn_rsqf(__builtin_fmaxf(r2, rout2));
                } else {
                    ri = __builtin_amdgcn_rsqf(r2);
                }
                const float mri = B[24 + k] * ri;
                const float mri3 = mri * (ri * ri);
                ax = __builtin_fmaf(mri3, dx, ax);
                ay = __builtin_fmaf(mri3, dy, ay);
                az = __builtin_fmaf(mri3, dz, az);
                p -= mri;
            }
        }
        A[0] += (double)ax;
        A[1] += (double)ay;
        A[2] += (double)az;
        A[3] += (double)p;
        nnb += nb;
    }
}

__device__ __forceinline__ void plain_spj(cfp ss, cfp tab, int s_begin, int s_end, float xi, float yi, float zi,
                                          float eps2, double (&A)[4])
{
    for (int s0 = s_begin; s0 < s_end; s0 += PTH_FLUSH_SPJ) {
        const int s1 = min(s0 + PTH_FLUSH_SPJ, s_end);
        float ax = 0.f, ay = 0.f, az = 0.f, p = 0.f;
        for (int s = s0; s < s1; s++) {
            const cfp R = ss + (size_t)s * PTH_SPJ_STAGED;
            // table entry: m, q'xx, q'yy, q'zz, q'xy, q'yz, q'zx, -eps^2 tr
            const cfp T = tab + (size_t)__float_as_int(R[3]) * 8;
            const float dx = xi - R[0], dy = yi - R[1], dz = zi - R[2];
            const float m = T[0];
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float ri2 = ri * ri, ri3 = ri * ri2, ri4 = ri2 * ri2, ri5 = ri2 * ri3;
            const float qrx = __builtin_fmaf(T[6], dz, __builtin_fmaf(T[4], dy, T[1] * dx));
            const float qry = __builtin_fmaf(T[5], dz, __builtin_fmaf(T[4], dx, T[2] * dy));
            const float qrz = __builtin_fmaf(T[5], dy, __builtin_fmaf(T[6], dx, T[3] * dz));
            const float rqr = __builtin_fmaf(qrz, dz, __builtin_fmaf(qry, dy, __builtin_fmaf(qrx, dx, T[7])));
            const float rqr_ri4 = rqr * ri4;
            const float meff = __builtin_fmaf(rqr_ri4, 0.5f, m);
            const float meff3 = __builtin_fmaf(rqr_ri4, 2.5f, m) * ri3;
            p = __builtin_fmaf(-meff, ri, p);
            ax = __builtin_fmaf(-meff3, dx, __builtin_fmaf(ri5, qrx, ax));
            ay = __builtin_fmaf(-meff3, dy, __builtin_fmaf(ri5, qry, ay));
            az = __builtin_fmaf(-meff3, dz, __builtin_fmaf(ri5, qrz, az));
        }
        A[0] += (double)ax;
        A[1] += (double)ay;
        A[2] += (double)az;
        A[3] += (double)p;
    }
}

template <int NNB_LE>
__global__ __launch_bounds__(256) void k_force_plain(const WalkDev *__restrict__ walks,
                                                     const int *__restrict__ near_blk,
                                                     const float4 *__restrict__ staged_i,
                                                     const float *__restrict__ staged_e,
                                                     const float *__restrict__ staged_s,
                                                     const SpjTab *__restrict__ spj_tab,
                                                     const long long *__restrict__ es_off,
                                                     const long long *__restrict__ ss_off,
                                                     float eps2, float rout2, double G,
                                                     double4 *__restrict__ out_accp,
                                                     int *__restrict__ out_nnb)
{
    const int w = blockIdx.x;
    const WalkDev wk = walks[w];
    const int wave = uniform((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int n_tiles = (wk.n_i + 63) >> 6;
    const int eb = uniform((int)(es_off[w] >> 3)), ee = uniform((int)(es_off[w + 1] >> 3));
    const int em = min(max(uniform(near_blk[w]), eb), ee);
    // staged SPJ records of the walk (none with monopole superparticles,
    // which are staged as far EPJ entries)
    const int sb = uniform((int)ss_off[w]), se_ = uniform((int)ss_off[w + 1]);
    for (int t = wave; t < n_tiles; t += 4) {
        const int i = 64 * t + lane;
        float4 v = make_float4(0.f, 0.f, 0.f, -1.f);
        if (i < wk.n_i) v = staged_i[wk.i_off + i];
        double A[4] = {0.0, 0.0, 0.0, 0.0};
        int nnb = 0;
        plain_epj<NNB_LE, true>(as_const(staged_e), eb, em, v.x, v.y, v.z, v.w, eps2, rout2, A, nnb);
        plain_epj<NNB_LE, false>(as_const(staged_e), em, ee, v.x, v.y, v.z, v.w, eps2, rout2, A, nnb);
        plain_spj(as_const(staged_s), as_const(reinterpret_cast<const float *>(spj_tab)), sb, se_, v.x, v.y, v.z,
                  eps2, A);
        if (i < wk.n_i) {
            out_accp[wk.i_off + i] = make_double4(G * A[0], G * A[1], G * A[2], G * A[3]);
            out_nnb[wk.i_off + i] = nnb;
        }
    }
}

// One 256-thread workgroup per split tile.
template <bool CO>
__global__ __launch_bounds__(256) void k_reduce(const RItem *__restrict__ ritems,
                                                const double4 *__restrict__ part_accp,
                                                const int *__restrict__ part_nnb,
                                                const double *__restrict__ part_pott, double G,
                                                double4 *__restrict__ out_accp, int *__restrict__ out_nnb,
                                                double *__restrict__ out_pott)
{
    const RItem r = ritems[blockIdx.x];
    const int t = threadIdx.x;
    if (t >= r.n_i) return;
    const int sub = t >> 6, lane = t & 63;
    double ax = 0, ay = 0, az = 0, p = 0, pt = 0;
    int nb = 0;
    for (int k = 0; k < r.n_part; k++) {
        const size_t slot = (size_t)(r.part_first + k * r.ipl + sub) * 64 + lane;
        const double4 a = part_accp[slot];
        ax += a.x;
        ay += a.y;
        az += a.z;
        p += a.w;
        nb += part_nnb[slot];
        if constexpr (CO) pt += part_pott[slot];
    }
    out_accp[r.out + t] = make_double4(G * ax, G * ay, G * az, G * p);
    out_nnb[r.out + t] = nb;
    if constexpr (CO) out_pott[r.out + t] = G * pt;
}

// ---------------------------------------------------------------------------
// index lists sent as runs (pth_dispatch): workgroup g expands part
// g / (2 n_walk) of n_part of the EPJ list (g % (2 n_walk) < n_walk) or the
// SPJ list of one walk.  The part's first run is found by a binary search
// over the walk's run positions; the runs are then taken EXPAND_RUNS at a
// time into LDS (first index, position); each entry of the part that those
// runs cover finds its run by a binary search in LDS and is written
// coalesced.  The first run of a non-empty list has position 0 and
// positions ascend.
#define EXPAND_RUNS 1024
__global__ __launch_bounds__(256) void k_expand_runs(const WalkDev *__restrict__ walks, int w_begin, int n_walk,
                                                     int n_part, const int *__restrict__ rid_e,
                                                     const int *__restrict__ rpos_e, const int *__restrict__ run_eo,
                                                     const int *__restrict__ rid_s, const int *__restrict__ rpos_s,
                                                     const int *__restrict__ run_so, int *__restrict__ out_e,
                                                     int *__restrict__ out_s)
{
    __shared__ int l_id[EXPAND_RUNS], l_pos[EXPAND_RUNS];
    const int part = (int)blockIdx.x / (2 * n_walk), g = (int)blockIdx.x % (2 * n_walk);
    const bool sp = g >= n_walk;
    const int w = w_begin + (sp ? g - n_walk : g);
    const WalkDev wk = walks[w];
    const int n = sp ? wk.n_s : wk.n_e;
    const int *__restrict__ rid = sp ? rid_s : rid_e;
    const int *__restrict__ rpos = sp ? rpos_s : rpos_e;
    const int *__restrict__ ro = sp ? run_so : run_eo;
    const int r0 = ro[w], r1 = ro[w + 1];
    // this part's entries [q0, q1)
    const int q0 = (int)((long long)n * part / n_part), q1 = (int)((long long)n * (part + 1) / n_part);
    if (q0 >= q1) return;
    // its first run: the last one whose position is <= q0 (the first run of a
    // non-empty list has position 0)
    int f = r0;
    for (int hi = r1 - 1; f < hi;) {
        const int mid = (f + hi + 1) >> 1;
        if (rpos[mid] <= q0) f = mid;
        else hi = mid - 1;
    }
    int *__restrict__ out = sp ? out_s + wk.s_off : out_e + wk.e_off;
    const int tid = (int)threadIdx.x;
    for (int c0 = f; c0 < r1; c0 += EXPAND_RUNS) {
        const int nc = min(EXPAND_RUNS, r1 - c0);
        for (int k = tid; k < nc; k += 256) {
            l_id[k] = rid[c0 + k];
            l_pos[k] = rpos[c0 + k];
        }
        __syncthreads();
        // the part's entries these runs cover: up to the next chunk's first run
        const int next = (c0 + nc < r1) ? rpos[c0 + nc] : n;
        const int p0 = max(l_pos[0], q0), p1 = min(next, q1);
        for (int k = p0 + tid; k < p1; k += 256) {
            int lo = 0, hi = nc - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (l_pos[mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            out[k] = l_id[lo] + (k - l_pos[lo]);
        }
        __syncthreads();
        if (next >= q1) break;
    }
}

// ---------------------------------------------------------------------------
// launchers (host)

hipError_t launch_spj_table(const SpjDev *spj, long long n, float eps2, SpjTab *tab, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spj_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, spj, n, (double)eps2, tab);
    return hipGetLastError();
}

hipError_t launch_expand_runs(const WalkDev *walks, int w_begin, int n_walk, const int *rid_e, const int *rpos_e,
                              const int *run_eo, const int *rid_s, const int *rpos_s, const int *run_so, int *out_e,
                              int *out_s, hipStream_t st)
{
    if (n_walk <= 0) return hipSuccess;
    // parts per list: a few blocks per CU for a batch of a few hundred walks
    const int n_part = std::max(1, std::min(8, 4096 / (2 * n_walk)));
    hipLaunchKernelGGL(k_expand_runs, dim3(2u * (unsigned)n_walk * (unsigned)n_part), dim3(256), 0, st, walks,
                       w_begin, n_walk, n_part, rid_e, rpos_e, run_eo, rid_s, rpos_s, run_so, out_e, out_s);
    return hipGetLastError();
}

hipError_t launch_stage(const BatchDev &b, int k, hipStream_t st)
{
    const SegDev &g = b.seg[k];
    if (g.n_walk <= 0) return hipSuccess;
    StageArgs a;
    a.walk_begin = g.walk_begin;
    a.epi = b.epi;
    a.walks = b.walks;
    a.es_off = b.es_off;
    a.ss_off = b.ss_off;
    a.id_epj = b.id_epj;
    a.id_spj = b.id_spj;
    a.epj = b.epj;
    a.spj = b.spj;
    a.r_out = std::sqrt((double)b.rout2);
    a.eps2 = (double)b.eps2;
    a.centre = b.centre;
    a.near_blk = b.near_blk;
    a.staged_i = b.staged_i;
    a.staged_e = b.staged_e;
    a.staged_s = b.staged_s;
    a.epi_id = b.epi_id;
    a.epj_co = b.epj_co;
    a.staged_ic = b.staged_ic;
    a.staged_il = b.staged_il;
    a.staged_ec = b.staged_ec;
    a.walk_err = b.walk_err;
    a.df = b.df;
    a.staged_el = b.staged_el;
    a.mono = b.mono;
    a.spj_mono = reinterpret_cast<const EpjDev *>(b.spj);
    a.counters = b.persistent ? b.counters + k * PLAN_CLASSES * 8 : nullptr;
    // SPJ parts per walk: enough blocks for a few per CU when the launch
    // has few walks (a dispatched batch), one beside each EPJ block otherwise
    a.n_walk = g.n_walk;
    a.spj_parts = b.mono ? 0 : std::max(1, std::min(8, 2048 / g.n_walk));
    const unsigned nblk = (unsigned)g.n_walk * (1u + (unsigned)a.spj_parts);
    if (b.co) hipLaunchKernelGGL((k_stage_walk<true, 2>), dim3(nblk), dim3(256), 0, st, a);
    else if (g.n_walk <= 2048) hipLaunchKernelGGL((k_stage_walk<false, 4>), dim3(nblk), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_stage_walk<false, 2>), dim3(nblk), dim3(256), 0, st, a);
    return hipGetLastError();
}

// 256-thread blocks of kernel f that the current device holds at once
static unsigned fit_blocks(const void *f)
{
    static std::mutex m;
    static std::unordered_map<const void *, int> per_cu;
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    std::lock_guard<std::mutex> lk(m);
    if (!cus[dev] && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    auto it = per_cu.find(f);
    if (it == per_cu.end()) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, 0) != hipSuccess) return 0;
        it = per_cu.emplace(f, nb).first;
    }
    return (unsigned)(it->second * cus[dev]);
}

// nblk blocks of one item per wave; persistent (P): as many blocks as the
// device holds, at most nblk, a multiple of 8 so that every XCD's run of
// items has blocks that drain it
template <int NNB, bool VC, bool CO, bool DF, bool P>
static void launch_class(int ipl, unsigned nblk, const ForceArgs &a, hipStream_t st)
{
    auto go = [&](void (*kern)(ForceArgs)) {
        unsigned n = nblk;
        if constexpr (P) n = std::max(8u, std::min(nblk, fit_blocks((const void *)kern)) & ~7u);
        hipLaunchKernelGGL(kern, dim3(n), dim3(256), 0, st, a);
    };
    if constexpr (CO) {
        if (ipl == 1) go(k_force_co<NNB, 1, VC, P>);
        else if (ipl == 2) go(k_force_co<NNB, 2, VC, P>);
        else if (ipl == 3) go(k_force_co<NNB, 3, VC, P>);
        else go(k_force_co<NNB, 4, VC, P>);
    } else {
        if (ipl == 1) go(k_force<NNB, 1, VC, false, DF, P>);
        else if (ipl == 2) go(k_force<NNB, 2, VC, false, DF, P>);
        else if (ipl == 3) go(k_force<NNB, 3, VC, false, DF, P>);
        else go(k_force<NNB, 4, VC, false, DF, P>);
    }
}

template <int NNB, bool VC, bool CO, bool DF>
static void launch_class(bool persistent, int ipl, unsigned nblk, const ForceArgs &a, hipStream_t st)
{
    if (persistent) launch_class<NNB, VC, CO, DF, true>(ipl, nblk, a, st);
    else launch_class<NNB, VC, CO, DF, false>(ipl, nblk, a, st);
}

hipError_t launch_force(const BatchDev &b, int k, hipStream_t st, unsigned classes)
{
    const SegDev &g = b.seg[k];
    if (b.plain) {
        if (b.n_walk <= 0) return hipSuccess;
        if (b.nnb_le)
            hipLaunchKernelGGL(k_force_plain<1>, dim3(b.n_walk), dim3(256), 0, st, b.walks, b.near_blk, b.staged_i,
                               b.staged_e, b.staged_s, b.spj_tab, b.es_off, b.ss_off, b.eps2, b.rout2, b.G,
                               b.out_accp, b.out_nnb);
        else
            hipLaunchKernelGGL(k_force_plain<0>, dim3(b.n_walk), dim3(256), 0, st, b.walks, b.near_blk, b.staged_i,
                               b.staged_e, b.staged_s, b.spj_tab, b.es_off, b.ss_off, b.eps2, b.rout2, b.G,
                               b.out_accp, b.out_nnb);
        return hipGetLastError();
    }
    static const int ipls[PLAN_CLASSES] = {1, 2, 3, 4};
    ForceArgs a;
    a.near_blk = b.near_blk;
    a.staged_i = b.staged_i;
    a.staged_e = b.staged_e;
    a.staged_s = b.staged_s;
    a.spj_tab = b.spj_tab;
    a.staged_ic = b.staged_ic;
    a.staged_il = b.staged_il;
    a.staged_ec = b.staged_ec;
    a.staged_el = b.staged_el;
    a.eps2 = b.eps2;
    a.rout2 = b.rout2;
    a.ngb2 = b.ngb2;
    a.G = b.G;
    a.out_accp = b.out_accp;
    a.out_nnb = b.out_nnb;
    a.out_pott = b.out_pott;
    a.part_accp = b.part_accp;
    a.part_nnb = b.part_nnb;
    a.part_pott = b.part_pott;
    // the longer classes first: the 1-i launch is about one round of waves,
    // and its run-down tail then overlaps the other force stream's work
    // instead of holding back a longer launch behind it on this stream
    for (int c = PLAN_CLASSES - 1; c >= 0; c--) {
        const int n = g.seg_count[c];
        if (n <= 0 || !((classes >> c) & 1u)) continue;
        a.items = b.items + g.seg_begin[c];
        a.n_items = n;
        // (the plan gives every class whole rows of 8 x 4 items, XCD-major)
        a.counter = b.persistent ? b.counters + (k * PLAN_CLASSES + c) * 8 : nullptr;
        a.n_run = n / 8;
        const bool pers = b.persistent != 0;
        const unsigned nblk = (unsigned)((n + 3) / 4);
        // the changeover and double-float near modes always use SGPR operands
        if (b.co) {
            if (b.nnb_le) launch_class<1, false, true, false>(pers, ipls[c], nblk, a, st);
            else launch_class<0, false, true, false>(pers, ipls[c], nblk, a, st);
        } else if (b.df) {
            if (b.nnb_le) launch_class<1, false, false, true>(pers, ipls[c], nblk, a, st);
            else launch_class<0, false, false, true>(pers, ipls[c], nblk, a, st);
        } else if (b.j_vgpr) {
            if (b.nnb_le) launch_class<1, true, false, false, false>(ipls[c], nblk, a, st);
            else launch_class<0, true, false, false, false>(ipls[c], nblk, a, st);
        } else {
            if (b.nnb_le) launch_class<1, false, false, false>(pers, ipls[c], nblk, a, st);
            else launch_class<0, false, false, false>(pers, ipls[c], nblk, a, st);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_reduce(const BatchDev &b, int k, hipStream_t st)
{
    const SegDev &g = b.seg[k];
    if (b.plain || g.n_ritems <= 0) return hipSuccess;
    const RItem *ri = b.ritems + g.ritem_begin;
    if (b.co)
        hipLaunchKernelGGL(k_reduce<true>, dim3((unsigned)g.n_ritems), dim3(256), 0, st, ri, b.part_accp,
                           b.part_nnb, b.part_pott, b.G, b.out_accp, b.out_nnb, b.out_pott);
    else
        hipLaunchKernelGGL(k_reduce<false>, dim3((unsigned)g.n_ritems), dim3(256), 0, st, ri, b.part_accp,
                           b.part_nnb, b.part_pott, b.G, b.out_accp, b.out_nnb, b.out_pott);
    return hipGetLastError();
}

hipError_t launch_all(const BatchDev &b, hipStream_t st)
{
    for (int k = 0; k < b.n_seg; k++) {
        hipError_t e;
        if ((e = launch_stage(b, k, st)) || (e = launch_force(b, k, st)) || (e = launch_reduce(b, k, st))) return e;
    }
    return hipSuccess;
}

}  // namespace pth
