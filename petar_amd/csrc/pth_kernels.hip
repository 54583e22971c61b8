// pth_kernels.hip -- gfx950 kernels of the soft-force path.
//
// Pipeline for one batch of walks (i-groups with their interaction lists):
//   k_stage_walk  one workgroup per walk: FP64 bounding box of the walk's i
//                 (centre c, half extents, largest r_search), every i as FP32
//                 (x-c, y-c, z-c, rs^2), and the walk's EPJ list gathered
//                 into FP32 records relative to c, split stably into a "near"
//                 region (entries that may lie within max(rs_i, rs_j, r_out)
//                 of some i of the walk) at the front and a "far" region at
//                 the back; records are SoA blocks of 8 {x[8] y[8] z[8] m[8]
//                 rs2[8]}, gaps padded with far massless entries; then the
//                 walk's SPJ list gathered into 12-float records relative to
//                 the centre {x y z m q'xx q'yy q'zz q'xy q'yz q'zx
//                 -eps^2 tr pad}, q' = 3Q - tr I
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

// SPJ record {x-c, y-c, z-c, m, q'xx, q'yy, q'zz, q'xy, q'yz, q'zx, -eps^2 tr, 0},
// q' = 3Q - tr I (the traceless rearrangement of
// src/phantomquad_for_p3t_x86.hpp:144-163)
__device__ __forceinline__ void put_spj(float4 *o, const SpjDev &p, const double (&c)[3], double eps2)
{
    const double qxx = p.qxx, qyy = p.qyy, qzz = p.qzz;
    const double tr = qxx + qyy + qzz;
    o[0] = make_float4((float)(p.x - c[0]), (float)(p.y - c[1]), (float)(p.z - c[2]), p.m);
    o[1] = make_float4((float)(3.0 * qxx - tr), (float)(3.0 * qyy - tr), (float)(3.0 * qzz - tr),
                       (float)(3.0 * (double)p.qxy));
    o[2] = make_float4((float)(3.0 * (double)p.qyz), (float)(3.0 * (double)p.qxz), (float)(-eps2 * tr), 0.f);
}

__global__ __launch_bounds__(256) void k_stage_walk(
    const EpiDev *__restrict__ epi, const WalkDev *__restrict__ walks,
    const long long *__restrict__ es_off, const int *__restrict__ id_epj,
    const EpjDev *__restrict__ epj, const long long *__restrict__ ss_off,
    const int *__restrict__ id_spj, const SpjDev *__restrict__ spj, double r_out, double eps2,
    double4 *__restrict__ centre, int *__restrict__ near_blk, float4 *__restrict__ staged_i,
    float *__restrict__ staged_e, float *__restrict__ staged_s)
{
    const int w = blockIdx.x;
    const WalkDev wk = walks[w];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __shared__ double red[4][7];
    __shared__ int cnt[2][4];
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
    if (tid == 0) centre[w] = make_double4(c[0], c[1], c[2], 0.0);
    for (int k = tid; k < wk.n_i; k += 256) {
        const EpiDev e = epi[wk.i_off + k];
        const double rs = (double)e.rs;
        staged_i[wk.i_off + k] = make_float4((float)(e.x - c[0]), (float)(e.y - c[1]),
                                             (float)(e.z - c[2]), (float)(rs * rs));
    }
    // -- EPJ list: stable split into near (front) and far (back, reversed)
    const long long base = es_off[w];
    const long long cap = es_off[w + 1] - base;
    const double hmax = fmax(h[0], fmax(h[1], h[2]));
    int n_near = 0, n_far = 0;
    for (int t0 = 0; t0 < wk.n_e; t0 += 256) {
        const int t = t0 + tid;
        const bool valid = t < wk.n_e;
        bool nearj = false;
        float x = 0.f, y = 0.f, z = 0.f, m = 0.f, r2 = 0.f;
        if (valid) {
            const EpjDev e = epj[id_epj[(long long)wk.e_off + t]];
            const double dx = e.x - c[0], dy = e.y - c[1], dz = e.z - c[2];
            // near: inside the walk's box widened by max(rs_i, rs_j, r_out),
            // with a margin that covers the FP32 rounding of the pair arithmetic
            const double R = fmax(fmax(rsm, (double)e.rs), r_out);
            const double dm = fmax(fabs(dx), fmax(fabs(dy), fabs(dz)));
            const double Rm = R * 1.00001 + 1e-5 * (dm + hmax);
            nearj = fabs(dx) <= h[0] + Rm && fabs(dy) <= h[1] + Rm && fabs(dz) <= h[2] + Rm;
            x = (float)dx;
            y = (float)dy;
            z = (float)dz;
            m = e.m;
            const double rs = (double)e.rs;
            r2 = (float)(rs * rs);
        }
        const unsigned long long bn = __ballot(valid && nearj);
        const unsigned long long bf = __ballot(valid && !nearj);
        const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (lane == 0) {
            cnt[0][wave] = __popcll(bn);
            cnt[1][wave] = __popcll(bf);
        }
        __syncthreads();
        int on = n_near, of = n_far;
        for (int q = 0; q < wave; q++) {
            on += cnt[0][q];
            of += cnt[1][q];
        }
        const int tn = cnt[0][0] + cnt[0][1] + cnt[0][2] + cnt[0][3];
        const int tf = cnt[1][0] + cnt[1][1] + cnt[1][2] + cnt[1][3];
        if (valid) {
            const long long slot = nearj ? base + on + __popcll(bn & below)
                                         : base + cap - 1 - (of + __popcll(bf & below));
            put_epj(staged_e, slot, x, y, z, m, r2);
        }
        n_near += tn;
        n_far += tf;
        __syncthreads();
    }
    // pad the gap between the two regions (far, massless)
    for (long long s = base + n_near + tid; s < base + cap - n_far; s += 256)
        put_epj(staged_e, s, 1e18f, 1e18f, 1e18f, 0.f, 0.f);
    if (tid == 0) near_blk[w] = (int)((base + ((n_near + 7) & ~7)) >> 3);
    // -- SPJ list: 12-float records relative to the centre, two per thread
    //    and pass so that two gathers are in flight
    float4 *out_s = reinterpret_cast<float4 *>(staged_s) + ss_off[w] * 3;
    const int *ids = id_spj + wk.s_off;
    for (int k0 = 0; k0 < wk.n_s; k0 += 512) {
        const int ka = k0 + tid, kb = ka + 256;
        const bool va = ka < wk.n_s, vb = kb < wk.n_s;
        SpjDev pa, pb;
        if (va) pa = spj[ids[ka]];
        if (vb) pb = spj[ids[kb]];
        if (va) put_spj(out_s + 3 * ka, pa, c, eps2);
        if (vb) put_spj(out_s + 3 * kb, pb, c, eps2);
    }
}

// ---------------------------------------------------------------------------
// force, fused form

struct Consts {
    float eps2, rout2, half, two_half;
};

template <int IPL, int NNB_LE, bool NEAR, bool VC>
__device__ __forceinline__ void epj_blocks(cfp se, int b_begin, int b_end, const float (&xi)[IPL],
                                           const float (&yi)[IPL], const float (&zi)[IPL],
                                           const float (&rsi2)[IPL], const Consts &K, double (&A)[IPL][4],
                                           int (&nnb)[IPL])
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
#pragma unroll
                for (int q = 0; q < IPL; q++) {
                    const float dx = xj - xi[q];
                    const float dy = yj - yi[q];
                    const float dz = zj - zi[q];
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

template <int IPL, bool VC>
__device__ __forceinline__ void spj_records(cfp ss, int s_begin, int s_end, const float (&xi)[IPL],
                                            const float (&yi)[IPL], const float (&zi)[IPL], const Consts &K,
                                            double (&A)[IPL][4])
{
    for (int s0 = s_begin; s0 < s_end; s0 += PTH_FLUSH_SPJ) {
        const int s1 = min(s0 + PTH_FLUSH_SPJ, s_end);
        float ax[IPL], ay[IPL], az[IPL], p[IPL];
#pragma unroll
        for (int q = 0; q < IPL; q++) ax[q] = ay[q] = az[q] = p[q] = 0.f;
        for (int s = s0; s < s1; s++) {
            const cfp R = ss + (size_t)s * 12;
            const float xs = jv<VC>(R[0]), ys = jv<VC>(R[1]), zs = jv<VC>(R[2]), m = jv<VC>(R[3]);
            const float qxx = jv<VC>(R[4]), qyy = jv<VC>(R[5]), qzz = jv<VC>(R[6]);
            const float qxy = jv<VC>(R[7]), qyz = jv<VC>(R[8]), qzx = jv<VC>(R[9]);
            const float mtr = jv<VC>(R[10]);
#pragma unroll
            for (int q = 0; q < IPL; q++) {
                const float dx = xi[q] - xs;
                const float dy = yi[q] - ys;
                const float dz = zi[q] - zs;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, K.eps2)));
                const float ri = __builtin_amdgcn_rsqf(r2);
                const float ri2 = ri * ri;
                const float ri3 = ri * ri2;
                const float ri4 = ri2 * ri2;
                const float ri5 = ri2 * ri3;
                const float qrx = __builtin_fmaf(qzx, dz, __builtin_fmaf(qxy, dy, qxx * dx));
                const float qry = __builtin_fmaf(qyz, dz, __builtin_fmaf(qxy, dx, qyy * dy));
                const float qrz = __builtin_fmaf(qyz, dy, __builtin_fmaf(qzx, dx, qzz * dz));
                const float rqr = __builtin_fmaf(qrz, dz, __builtin_fmaf(qry, dy, __builtin_fmaf(qrx, dx, mtr)));
                const float rqr_ri4 = rqr * ri4;
                const float meff = __builtin_fmaf(rqr_ri4, K.half, m);
                const float meff3 = __builtin_fmaf(rqr_ri4, K.two_half, m) * ri3;
                p[q] = __builtin_fmaf(-meff, ri, p[q]);
                ax[q] = __builtin_fmaf(-meff3, dx, __builtin_fmaf(ri5, qrx, ax[q]));
                ay[q] = __builtin_fmaf(-meff3, dy, __builtin_fmaf(ri5, qry, ay[q]));
                az[q] = __builtin_fmaf(-meff3, dz, __builtin_fmaf(ri5, qrz, az[q]));
            }
        }
#pragma unroll
        for (int q = 0; q < IPL; q++) {
            A[q][0] += (double)ax[q];
            A[q][1] += (double)ay[q];
            A[q][2] += (double)az[q];
            A[q][3] += (double)p[q];
        }
    }
}

template <int IPL, int NNB_LE, bool VC>
__device__ __forceinline__ void run_tile(const Item &it, int near_end, int lane,
                                         const float4 *__restrict__ staged_i, cfp se, cfp ss, const Consts &K,
                                         double G, double4 *__restrict__ out_accp, int *__restrict__ out_nnb,
                                         double4 *__restrict__ part_accp, int *__restrict__ part_nnb)
{
    const int n_i = item_n_i(it.n_i_ipl);
    float xi[IPL], yi[IPL], zi[IPL], rsi2[IPL];
    double A[IPL][4];
    int nnb[IPL];
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        float4 v = make_float4(0.f, 0.f, 0.f, -1.f);
        if (lane + 64 * q < n_i) v = staged_i[it.i_begin + lane + 64 * q];
        xi[q] = v.x;
        yi[q] = v.y;
        zi[q] = v.z;
        rsi2[q] = v.w;
        A[q][0] = A[q][1] = A[q][2] = A[q][3] = 0.0;
        nnb[q] = 0;
    }
    const int e_mid = min(max(near_end, it.e_blk_begin), it.e_blk_end);
    epj_blocks<IPL, NNB_LE, true, VC>(se, it.e_blk_begin, e_mid, xi, yi, zi, rsi2, K, A, nnb);
    epj_blocks<IPL, NNB_LE, false, VC>(se, e_mid, it.e_blk_end, xi, yi, zi, rsi2, K, A, nnb);
    spj_records<IPL, VC>(ss, it.s_begin, it.s_end, xi, yi, zi, K, A);
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = lane + 64 * q;
        if (i >= n_i) continue;
        if (it.part < 0) {
            out_accp[it.i_begin + i] = make_double4(G * A[q][0], G * A[q][1], G * A[q][2], G * A[q][3]);
            out_nnb[it.i_begin + i] = nnb[q];
        } else {
            const size_t slot = (size_t)(it.part + q) * 64 + lane;
            part_accp[slot] = make_double4(A[q][0], A[q][1], A[q][2], A[q][3]);
            part_nnb[slot] = nnb[q];
        }
    }
}

// One launch per i-per-lane class: items [0, n_items) of this class.
template <int NNB_LE, int IPL, bool VC>
__global__ __launch_bounds__(256) void k_force(const Item *__restrict__ items, int n_items,
                                               const int *__restrict__ near_blk,
                                               const float4 *__restrict__ staged_i,
                                               const float *__restrict__ staged_e,
                                               const float *__restrict__ staged_s, float eps2,
                                               float rout2, double G, double4 *__restrict__ out_accp,
                                               int *__restrict__ out_nnb,
                                               double4 *__restrict__ part_accp,
                                               int *__restrict__ part_nnb)
{
    const int idx = uniform((int)blockIdx.x * 4 + ((int)threadIdx.x >> 6));
    if (idx >= n_items) return;
    const cip ci = (cip)(items) + (size_t)idx * 8;
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
    const int near_end = ((cip)(near_blk))[it.walk];
    const Consts K{jv<VC>(eps2), jv<VC>(rout2), jv<VC>(0.5f), jv<VC>(2.5f)};
    run_tile<IPL, NNB_LE, VC>(it, near_end, threadIdx.x & 63, staged_i, as_const(staged_e), as_const(staged_s),
                              K, G, out_accp, out_nnb, part_accp, part_nnb);
}

// ---------------------------------------------------------------------------
// force, plain per-walk form: SGPR operands, one i per lane, whole list

template <int NNB_LE>
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
                const float rs2 = __int_as_float(max(__float_as_int(rsi2), __float_as_int(B[32 + k])));
                if (NNB_LE) nb += (r2 <= rs2) ? 1 : 0;
                else nb += (r2 < rs2) ? 1 : 0;
                const float ri = __builtin_amdgcn_rsqf(__builtin_fmaxf(r2, rout2));
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

__device__ __forceinline__ void plain_spj(cfp ss, int s_begin, int s_end, float xi, float yi, float zi,
                                          float eps2, double (&A)[4])
{
    for (int s0 = s_begin; s0 < s_end; s0 += PTH_FLUSH_SPJ) {
        const int s1 = min(s0 + PTH_FLUSH_SPJ, s_end);
        float ax = 0.f, ay = 0.f, az = 0.f, p = 0.f;
        for (int s = s0; s < s1; s++) {
            const cfp R = ss + (size_t)s * 12;
            const float dx = xi - R[0], dy = yi - R[1], dz = zi - R[2];
            const float m = R[3];
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float ri2 = ri * ri, ri3 = ri * ri2, ri4 = ri2 * ri2, ri5 = ri2 * ri3;
            const float qrx = __builtin_fmaf(R[9], dz, __builtin_fmaf(R[7], dy, R[4] * dx));
            const float qry = __builtin_fmaf(R[8], dz, __builtin_fmaf(R[7], dx, R[5] * dy));
            const float qrz = __builtin_fmaf(R[8], dy, __builtin_fmaf(R[9], dx, R[6] * dz));
            const float rqr = __builtin_fmaf(qrz, dz, __builtin_fmaf(qry, dy, __builtin_fmaf(qrx, dx, R[10])));
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
                                                     const float4 *__restrict__ staged_i,
                                                     const float *__restrict__ staged_e,
                                                     const float *__restrict__ staged_s,
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
    const int sb = uniform((int)ss_off[w]), se_ = uniform((int)ss_off[w] + wk.n_s);
    for (int t = wave; t < n_tiles; t += 4) {
        const int i = 64 * t + lane;
        float4 v = make_float4(0.f, 0.f, 0.f, -1.f);
        if (i < wk.n_i) v = staged_i[wk.i_off + i];
        double A[4] = {0.0, 0.0, 0.0, 0.0};
        int nnb = 0;
        plain_epj<NNB_LE>(as_const(staged_e), eb, ee, v.x, v.y, v.z, v.w, eps2, rout2, A, nnb);
        plain_spj(as_const(staged_s), sb, se_, v.x, v.y, v.z, eps2, A);
        if (i < wk.n_i) {
            out_accp[wk.i_off + i] = make_double4(G * A[0], G * A[1], G * A[2], G * A[3]);
            out_nnb[wk.i_off + i] = nnb;
        }
    }
}

// One 256-thread workgroup per split tile.
__global__ __launch_bounds__(256) void k_reduce(const RItem *__restrict__ ritems,
                                                const double4 *__restrict__ part_accp,
                                                const int *__restrict__ part_nnb, double G,
                                                double4 *__restrict__ out_accp,
                                                int *__restrict__ out_nnb)
{
    const RItem r = ritems[blockIdx.x];
    const int t = threadIdx.x;
    if (t >= r.n_i) return;
    const int sub = t >> 6, lane = t & 63;
    double ax = 0, ay = 0, az = 0, p = 0;
    int nb = 0;
    for (int k = 0; k < r.n_part; k++) {
        const size_t slot = (size_t)(r.part_first + k * r.ipl + sub) * 64 + lane;
        const double4 a = part_accp[slot];
        ax += a.x;
        ay += a.y;
        az += a.z;
        p += a.w;
        nb += part_nnb[slot];
    }
    out_accp[r.out + t] = make_double4(G * ax, G * ay, G * az, G * p);
    out_nnb[r.out + t] = nb;
}

// ---------------------------------------------------------------------------
// launchers (host)

hipError_t launch_stage(const BatchDev &b, hipStream_t st)
{
    if (b.n_walk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_walk, dim3(b.n_walk), dim3(256), 0, st, b.epi, b.walks, b.es_off, b.id_epj,
                       b.epj, b.ss_off, b.id_spj, b.spj, std::sqrt((double)b.rout2), (double)b.eps2, b.centre,
                       b.near_blk, b.staged_i, b.staged_e, b.staged_s);
    return hipGetLastError();
}

hipError_t launch_force(const BatchDev &b, hipStream_t st)
{
    if (b.plain) {
        if (b.n_walk <= 0) return hipSuccess;
        if (b.nnb_le)
            hipLaunchKernelGGL(k_force_plain<1>, dim3(b.n_walk), dim3(256), 0, st, b.walks, b.staged_i,
                               b.staged_e, b.staged_s, b.es_off, b.ss_off, b.eps2, b.rout2, b.G,
                               b.out_accp, b.out_nnb);
        else
            hipLaunchKernelGGL(k_force_plain<0>, dim3(b.n_walk), dim3(256), 0, st, b.walks, b.staged_i,
                               b.staged_e, b.staged_s, b.es_off, b.ss_off, b.eps2, b.rout2, b.G,
                               b.out_accp, b.out_nnb);
        return hipGetLastError();
    }
    static const int ipls[3] = {1, 2, 4};
    for (int c = 0; c < 3; c++) {
        const int n = b.seg_count[c];
        if (n <= 0) continue;
        const Item *it = b.items + b.seg_begin[c];
        const unsigned nblk = (unsigned)((n + 3) / 4);
#define PTH_LAUNCH(NNB, IPL, VC)                                                                           \
    hipLaunchKernelGGL((k_force<NNB, IPL, VC>), dim3(nblk), dim3(256), 0, st, it, n, b.near_blk, b.staged_i, \
                       b.staged_e, b.staged_s, b.eps2, b.rout2, b.G, b.out_accp, b.out_nnb, b.part_accp,     \
                       b.part_nnb)
#define PTH_LAUNCH_IPL(NNB, VC)                  \
    do {                                         \
        if (ipls[c] == 1) PTH_LAUNCH(NNB, 1, VC); \
        else if (ipls[c] == 2) PTH_LAUNCH(NNB, 2, VC); \
        else PTH_LAUNCH(NNB, 4, VC);             \
    } while (0)
        if (b.nnb_le) {
            if (b.j_vgpr) PTH_LAUNCH_IPL(1, true);
            else PTH_LAUNCH_IPL(1, false);
        } else {
            if (b.j_vgpr) PTH_LAUNCH_IPL(0, true);
            else PTH_LAUNCH_IPL(0, false);
        }
#undef PTH_LAUNCH_IPL
#undef PTH_LAUNCH
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_reduce(const BatchDev &b, hipStream_t st)
{
    if (b.plain || b.n_ritems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)b.n_ritems), dim3(256), 0, st, b.ritems, b.part_accp,
                       b.part_nnb, b.G, b.out_accp, b.out_nnb);
    return hipGetLastError();
}

}  // namespace pth
