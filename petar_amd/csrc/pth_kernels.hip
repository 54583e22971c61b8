// pth_kernels.hip -- gfx950 kernels of the soft-force path.
//
// Pipeline for one batch of walks (i-groups with their interaction lists):
//   k_stage_i   one workgroup per walk: FP64 bounding-box centre of the
//               walk's i-particles, then every i as FP32 (x-c, y-c, z-c, rs^2)
//   k_stage_j   gather of the index lists (id_epj / id_spj) into per-walk
//               contiguous FP32 records relative to the walk centre:
//               EPJ as SoA blocks of 8 {x[8] y[8] z[8] m[8] rs2[8]},
//               SPJ as 12-float records {x y z m q'xx q'yy q'zz q'xy q'yz
//               q'zx -eps^2 tr pad} with q' = 3Q - tr I
//   k_force     one wave per work item (walk, 64-i tile, j-range): i in
//               registers (one per lane), j records read through the
//               constant address space so that they arrive in SGPRs with
//               one scalar fetch per wave; FP32 pair arithmetic, FP32
//               partial sums per chunk, FP64 running sums
//   k_reduce    fixed-order FP64 sum of the partial results of walks whose
//               j-range was split over several items
//   k_force_plain  the plain per-walk form: one workgroup per walk, every
//               wave takes 64-i tiles of the walk over the whole list
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

// ---------------------------------------------------------------------------
// staging

__global__ __launch_bounds__(256) void k_stage_i(const EpiDev *__restrict__ epi,
                                                 const WalkDev *__restrict__ walks,
                                                 double4 *__restrict__ centre,
                                                 float4 *__restrict__ staged_i)
{
    const int w = blockIdx.x;
    const WalkDev wk = walks[w];
    const int tid = threadIdx.x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = tid; k < wk.n_i; k += 256) {
        const EpiDev e = epi[wk.i_off + k];
        lo[0] = fmin(lo[0], e.x); hi[0] = fmax(hi[0], e.x);
        lo[1] = fmin(lo[1], e.y); hi[1] = fmax(hi[1], e.y);
        lo[2] = fmin(lo[2], e.z); hi[2] = fmax(hi[2], e.z);
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
        }
    }
    __shared__ double red[4][6];
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0)
        for (int d = 0; d < 3; d++) { red[wave][d] = lo[d]; red[wave][3 + d] = hi[d]; }
    __syncthreads();
    double c[3];
    for (int d = 0; d < 3; d++) {
        double l = fmin(fmin(red[0][d], red[1][d]), fmin(red[2][d], red[3][d]));
        double h = fmax(fmax(red[0][3 + d], red[1][3 + d]), fmax(red[2][3 + d], red[3][3 + d]));
        c[d] = (wk.n_i > 0) ? 0.5 * (l + h) : 0.0;
    }
    if (tid == 0) centre[w] = make_double4(c[0], c[1], c[2], 0.0);
    for (int k = tid; k < wk.n_i; k += 256) {
        const EpiDev e = epi[wk.i_off + k];
        const double rs = (double)e.rs;
        staged_i[wk.i_off + k] = make_float4((float)(e.x - c[0]), (float)(e.y - c[1]),
                                             (float)(e.z - c[2]), (float)(rs * rs));
    }
}

__device__ __forceinline__ int walk_of(const long long *__restrict__ off, int n_walk, long long s)
{
    // largest w with off[w] <= s (off is non-decreasing, off[0] = 0)
    int lo = 0, hi = n_walk;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= s) lo = mid; else hi = mid;
    }
    return lo;
}

// Blocks [0, nb_e) gather EPJ slots, blocks [nb_e, nb_e + nb_s) SPJ records.
__global__ __launch_bounds__(256) void k_stage_j(
    int n_walk, const WalkDev *__restrict__ walks, const long long *__restrict__ es_off,
    const long long *__restrict__ ss_off, const double4 *__restrict__ centre,
    const int *__restrict__ id_epj, const int *__restrict__ id_spj,
    const EpjDev *__restrict__ epj, const SpjDev *__restrict__ spj,
    long long n_e_slots, long long n_s_rec, int nb_e, float eps2,
    float *__restrict__ staged_e, float *__restrict__ staged_s)
{
    if ((int)blockIdx.x < nb_e) {
        const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
        if (s >= n_e_slots) return;
        const int w = walk_of(es_off, n_walk, s);
        const WalkDev wk = walks[w];
        const long long k = s - es_off[w];
        float x = 1e18f, y = 1e18f, z = 1e18f, m = 0.f, r2 = 0.f;  // padding: far, massless
        if (k < wk.n_e) {
            const EpjDev e = epj[id_epj[(long long)wk.e_off + k]];
            const double4 c = centre[w];
            x = (float)(e.x - c.x);
            y = (float)(e.y - c.y);
            z = (float)(e.z - c.z);
            m = e.m;
            const double rs = (double)e.rs;
            r2 = (float)(rs * rs);
        }
        float *blk = staged_e + (s >> 3) * 40;
        const int l = (int)(s & 7);
        blk[l] = x;
        blk[8 + l] = y;
        blk[16 + l] = z;
        blk[24 + l] = m;
        blk[32 + l] = r2;
    } else {
        const long long s = (long long)(blockIdx.x - nb_e) * 256 + threadIdx.x;
        if (s >= n_s_rec) return;
        const int w = walk_of(ss_off, n_walk, s);
        const WalkDev wk = walks[w];
        const long long k = s - ss_off[w];
        if (k >= wk.n_s) return;
        const SpjDev p = spj[id_spj[(long long)wk.s_off + k]];
        const double4 c = centre[w];
        const double qxx = p.qxx, qyy = p.qyy, qzz = p.qzz;
        const double tr = qxx + qyy + qzz;
        float4 *o = reinterpret_cast<float4 *>(staged_s + s * 12);
        o[0] = make_float4((float)(p.x - c.x), (float)(p.y - c.y), (float)(p.z - c.z), p.m);
        o[1] = make_float4((float)(3.0 * qxx - tr), (float)(3.0 * qyy - tr), (float)(3.0 * qzz - tr),
                           (float)(3.0 * (double)p.qxy));
        o[2] = make_float4((float)(3.0 * (double)p.qyz), (float)(3.0 * (double)p.qxz),
                           (float)(-(double)eps2 * tr), 0.f);
    }
}

// ---------------------------------------------------------------------------
// force

struct Acc {
    double ax, ay, az, p;
    int nnb;
};

template <int NNB_LE>
__device__ __forceinline__ void epj_range(cfp se, int b_begin, int b_end, float xi, float yi,
                                          float zi, float rsi2, float eps2, float rout2, Acc &A)
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
                // rs^2 values are >= 0 (or -1 on idle lanes), so the signed
                // integer maximum of their bit patterns is the float maximum
                const float rs2 = __int_as_float(max(__float_as_int(rsi2), __float_as_int(B[32 + k])));
                if (NNB_LE) nb += (r2 <= rs2) ? 1 : 0;
                else nb += (r2 < rs2) ? 1 : 0;
                const float r2c = __builtin_fmaxf(r2, rout2);
                const float ri = __builtin_amdgcn_rsqf(r2c);
                const float mri = B[24 + k] * ri;
                const float ri2 = ri * ri;
                const float mri3 = mri * ri2;
                ax = __builtin_fmaf(mri3, dx, ax);
                ay = __builtin_fmaf(mri3, dy, ay);
                az = __builtin_fmaf(mri3, dz, az);
                p -= mri;
            }
        }
        A.ax += (double)ax;
        A.ay += (double)ay;
        A.az += (double)az;
        A.p += (double)p;
        A.nnb += nb;
    }
}

__device__ __forceinline__ void spj_range(cfp ss, int s_begin, int s_end, float xi, float yi,
                                          float zi, float eps2, Acc &A)
{
    for (int s0 = s_begin; s0 < s_end; s0 += PTH_FLUSH_SPJ) {
        const int s1 = min(s0 + PTH_FLUSH_SPJ, s_end);
        float ax = 0.f, ay = 0.f, az = 0.f, p = 0.f;
        for (int s = s0; s < s1; s++) {
            const cfp R = ss + (size_t)s * 12;
            const float dx = xi - R[0];
            const float dy = yi - R[1];
            const float dz = zi - R[2];
            const float m = R[3];
            const float qxx = R[4], qyy = R[5], qzz = R[6], qxy = R[7], qyz = R[8], qzx = R[9];
            const float mtr = R[10];
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
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
            const float meff = __builtin_fmaf(rqr_ri4, 0.5f, m);
            const float meff3 = __builtin_fmaf(rqr_ri4, 2.5f, m) * ri3;
            p = __builtin_fmaf(-meff, ri, p);
            ax = __builtin_fmaf(-meff3, dx, __builtin_fmaf(ri5, qrx, ax));
            ay = __builtin_fmaf(-meff3, dy, __builtin_fmaf(ri5, qry, ay));
            az = __builtin_fmaf(-meff3, dz, __builtin_fmaf(ri5, qrz, az));
        }
        A.ax += (double)ax;
        A.ay += (double)ay;
        A.az += (double)az;
        A.p += (double)p;
    }
}

template <int NNB_LE>
__device__ __forceinline__ void run_item(const Item &it, int lane, const float4 *__restrict__ staged_i,
                                         cfp se, cfp ss, float eps2, float rout2, double G,
                                         double4 *__restrict__ out_accp, int *__restrict__ out_nnb,
                                         double4 *__restrict__ part_accp, int *__restrict__ part_nnb)
{
    const bool valid = lane < it.n_i;
    float4 xi = make_float4(0.f, 0.f, 0.f, -1.f);
    if (valid) xi = staged_i[it.i_begin + lane];
    Acc A{0.0, 0.0, 0.0, 0.0, 0};
    epj_range<NNB_LE>(se, it.e_blk_begin, it.e_blk_end, xi.x, xi.y, xi.z, xi.w, eps2, rout2, A);
    spj_range(ss, it.s_begin, it.s_end, xi.x, xi.y, xi.z, eps2, A);
    if (!valid) return;
    if (it.part < 0) {
        out_accp[it.i_begin + lane] = make_double4(G * A.ax, G * A.ay, G * A.az, G * A.p);
        out_nnb[it.i_begin + lane] = A.nnb;
    } else {
        part_accp[(size_t)it.part * 64 + lane] = make_double4(A.ax, A.ay, A.az, A.p);
        part_nnb[(size_t)it.part * 64 + lane] = A.nnb;
    }
}

template <int NNB_LE>
__global__ __launch_bounds__(256) void k_force(const Item *__restrict__ items, int n_items,
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
    it.n_i = ci[2];
    it.e_blk_begin = ci[3];
    it.e_blk_end = ci[4];
    it.s_begin = ci[5];
    it.s_end = ci[6];
    it.part = ci[7];
    if (it.walk < 0) return;
    run_item<NNB_LE>(it, threadIdx.x & 63, staged_i, as_const(staged_e), as_const(staged_s), eps2,
                     rout2, G, out_accp, out_nnb, part_accp, part_nnb);
}

// One workgroup per walk, waves take the walk's 64-i tiles round robin.
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
    const int wave = uniform((int)threadIdx.x >> 6);
    const int n_tiles = (wk.n_i + 63) >> 6;
    Item it;
    it.walk = w;
    it.e_blk_begin = uniform((int)(es_off[w] >> 3));
    it.e_blk_end = uniform((int)(es_off[w + 1] >> 3));
    it.s_begin = uniform((int)ss_off[w]);
    it.s_end = uniform((int)ss_off[w] + wk.n_s);
    it.part = -1;
    for (int t = wave; t < n_tiles; t += 4) {
        it.i_begin = wk.i_off + 64 * t;
        it.n_i = min(64, wk.n_i - 64 * t);
        run_item<NNB_LE>(it, threadIdx.x & 63, staged_i, as_const(staged_e), as_const(staged_s), eps2,
                         rout2, G, out_accp, out_nnb, nullptr, nullptr);
    }
}

__global__ __launch_bounds__(256) void k_reduce(const RItem *__restrict__ ritems, int n_ritems,
                                                const double4 *__restrict__ part_accp,
                                                const int *__restrict__ part_nnb, double G,
                                                double4 *__restrict__ out_accp,
                                                int *__restrict__ out_nnb)
{
    const int idx = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (idx >= n_ritems) return;
    const RItem r = ritems[idx];
    const int lane = threadIdx.x & 63;
    if (lane >= r.n_i) return;
    double ax = 0, ay = 0, az = 0, p = 0;
    int nb = 0;
    for (int k = 0; k < r.n_part; k++) {
        const double4 a = part_accp[(size_t)(r.part_first + k) * 64 + lane];
        ax += a.x;
        ay += a.y;
        az += a.z;
        p += a.w;
        nb += part_nnb[(size_t)(r.part_first + k) * 64 + lane];
    }
    out_accp[r.out + lane] = make_double4(G * ax, G * ay, G * az, G * p);
    out_nnb[r.out + lane] = nb;
}

// ---------------------------------------------------------------------------
// launchers (host)

hipError_t launch_stage(const BatchDev &b, hipStream_t st)
{
    if (b.n_walk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_i, dim3(b.n_walk), dim3(256), 0, st, b.epi, b.walks, b.centre,
                       b.staged_i);
    const long long nbe = (b.n_e_slots + 255) / 256, nbs = (b.n_s_rec + 255) / 256;
    if (nbe + nbs > 0)
        hipLaunchKernelGGL(k_stage_j, dim3((unsigned)(nbe + nbs)), dim3(256), 0, st, b.n_walk,
                           b.walks, b.es_off, b.ss_off, b.centre, b.id_epj, b.id_spj, b.epj, b.spj,
                           b.n_e_slots, b.n_s_rec, (int)nbe, b.eps2, b.staged_e, b.staged_s);
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
    if (b.n_items <= 0) return hipSuccess;
    const unsigned nblk = (unsigned)((b.n_items + 3) / 4);
    if (b.nnb_le)
        hipLaunchKernelGGL(k_force<1>, dim3(nblk), dim3(256), 0, st, b.items, b.n_items, b.staged_i,
                           b.staged_e, b.staged_s, b.eps2, b.rout2, b.G, b.out_accp, b.out_nnb,
                           b.part_accp, b.part_nnb);
    else
        hipLaunchKernelGGL(k_force<0>, dim3(nblk), dim3(256), 0, st, b.items, b.n_items, b.staged_i,
                           b.staged_e, b.staged_s, b.eps2, b.rout2, b.G, b.out_accp, b.out_nnb,
                           b.part_accp, b.part_nnb);
    return hipGetLastError();
}

hipError_t launch_reduce(const BatchDev &b, hipStream_t st)
{
    if (b.plain || b.n_ritems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)((b.n_ritems + 3) / 4)), dim3(256), 0, st, b.ritems,
                       b.n_ritems, b.part_accp, b.part_nnb, b.G, b.out_accp, b.out_nnb);
    return hipGetLastError();
}

}  // namespace pth
