// SPJ pair throughput, two forms (a measurement, not part of the library):
//  - k_valu<IPL>: the library's form; records read as wave-uniform scalar
//    operands, IPL i per lane, 34 VALU instructions per pair;
//  - k_mfma<S>: everything linear in the i position from the matrix cores:
//    v_mfma_f32_16x16x4_f32 with features (x, y, z, 1) of 16 i as B and
//    weights of 4 records as A, two MFMAs giving dx, dy, dz, (Q.r - Q.c)xyz,
//    m, mtr per (i, record) pair, 22 VALU instructions per pair after that;
//    S sets of 16 i per wave, lane group lane/16 takes record 4b + lane/16.
// All i are active, every wave reads the same 4096 records, so both forms
// run at their issue limits.  k_layout first checks the MFMA operand and
// result lane layout on exact small integers.
// Build and run: make -C petar_amd/csrc probe && petar_amd/lib/probe_spj
// Bounds: xi[wave*tile + k], k < tile, waves*tile = total_i; W[b*64+lane] and
// rec16[b*64 + f], b < nblk, f < 64; rec[s*12 + f], s < nj (the prefetch index
// is clamped to nj-1); out[i*4 + k], i < total_i.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>
typedef const __attribute__((address_space(4))) float *cfp;
typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void k_layout(const float *a, const float *b, float *d)
{
    const int l = threadIdx.x;
    f4 z = {0.f, 0.f, 0.f, 0.f};
    f4 r = __builtin_amdgcn_mfma_f32_16x16x4f32(a[l], b[l], z, 0, 0, 0);
    d[l * 4 + 0] = r[0]; d[l * 4 + 1] = r[1]; d[l * 4 + 2] = r[2]; d[l * 4 + 3] = r[3];
}

struct Rec { float x, y, z, m, qxx, qyy, qzz, qxy, qyz, qzx, mtr; };
__device__ __forceinline__ Rec ld(cfp R) { return Rec{R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8], R[9], R[10]}; }

template <int IPL>
__global__ __launch_bounds__(256) void k_valu(const float4 *xi_all, const float *rec, int nj, float *out)
{
    const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    const int base = wave * 64 * IPL;
    float xi[IPL], yi[IPL], zi[IPL];
#pragma unroll
    for (int q = 0; q < IPL; q++) { const float4 p = xi_all[base + q * 64 + lane]; xi[q] = p.x; yi[q] = p.y; zi[q] = p.z; }
    double A[IPL][4] = {};
    cfp ss = (cfp)rec;
    Rec c = ld(ss);
    for (int s0 = 0; s0 < nj; s0 += 32) {
        float ax[IPL] = {}, ay[IPL] = {}, az[IPL] = {}, p[IPL] = {};
        for (int s = s0; s < s0 + 32; s++) {
            const Rec n = ld(ss + (size_t)min(s + 1, nj - 1) * 12);
#pragma unroll
            for (int q = 0; q < IPL; q++) {
                const float dx = xi[q] - c.x, dy = yi[q] - c.y, dz = zi[q] - c.z;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float ri = __builtin_amdgcn_rsqf(r2);
                const float ri2 = ri * ri, ri3 = ri * ri2, ri4 = ri2 * ri2, ri5 = ri2 * ri3;
                const float qrx = __builtin_fmaf(c.qzx, dz, __builtin_fmaf(c.qxy, dy, c.qxx * dx));
                const float qry = __builtin_fmaf(c.qyz, dz, __builtin_fmaf(c.qxy, dx, c.qyy * dy));
                const float qrz = __builtin_fmaf(c.qyz, dy, __builtin_fmaf(c.qzx, dx, c.qzz * dz));
                const float rqr = __builtin_fmaf(qrz, dz, __builtin_fmaf(qry, dy, __builtin_fmaf(qrx, dx, c.mtr)));
                const float t = rqr * ri4;
                const float meff = __builtin_fmaf(t, 0.5f, c.m);
                const float meff3 = __builtin_fmaf(t, 2.5f, c.m) * ri3;
                p[q] = __builtin_fmaf(-meff, ri, p[q]);
                ax[q] = __builtin_fmaf(-meff3, dx, __builtin_fmaf(ri5, qrx, ax[q]));
                ay[q] = __builtin_fmaf(-meff3, dy, __builtin_fmaf(ri5, qry, ay[q]));
                az[q] = __builtin_fmaf(-meff3, dz, __builtin_fmaf(ri5, qrz, az[q]));
            }
            c = n;
        }
#pragma unroll
        for (int q = 0; q < IPL; q++) { A[q][0] += ax[q]; A[q][1] += ay[q]; A[q][2] += az[q]; A[q][3] += p[q]; }
    }
#pragma unroll
    for (int q = 0; q < IPL; q++)
        for (int k = 0; k < 4; k++) out[(size_t)(base + q * 64 + lane) * 4 + k] = (float)A[q][k];
}

// S sets of 16 i per wave; lane group g = lane/16 takes record 4b+g of block b.
template <int S, bool GATHER>
__global__ __launch_bounds__(256) void k_mfma(const float4 *xi_all, const float2 *W, const float *rec16,
                                              const int *fidx, int nblk, float *out)
{
    const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, c16 = lane & 15;
    const int base = wave * 16 * S;
    float bf[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const float4 p = xi_all[base + s * 16 + c16];
        bf[s] = g == 0 ? p.x : g == 1 ? p.y : g == 2 ? p.z : 1.f;
    }
    int f1 = 0, f2 = 0;
    if (GATHER) { f1 = ((lane & 15) >> 2) * 16 + fidx[lane]; f2 = ((lane & 15) >> 2) * 16 + fidx[64 + lane]; }
    double A[S][4] = {};
    const f4 z = {0.f, 0.f, 0.f, 0.f};
    float n1, n2;
    if (GATHER) { n1 = rec16[f1]; n2 = rec16[f2]; }
    else { const float2 w = W[lane]; n1 = w.x; n2 = w.y; }
    for (int b0 = 0; b0 < nblk; b0 += 8) {
        float ax[S] = {}, ay[S] = {}, az[S] = {}, p[S] = {};
        for (int b = b0; b < b0 + 8; b++) {
            const float w1 = n1, w2 = n2;
            const int bn = min(b + 1, nblk - 1);
            if (GATHER) { n1 = rec16[bn * 64 + f1]; n2 = rec16[bn * 64 + f2]; }
            else { const float2 w = W[bn * 64 + lane]; n1 = w.x; n2 = w.y; }
#pragma unroll
            for (int s = 0; s < S; s++) {
                const f4 d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(w1, bf[s], z, 0, 0, 0);
                const f4 d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(w2, bf[s], z, 0, 0, 0);
                const float dx = d1[0], dy = d1[1], dz = d1[2], qrx = d1[3];
                const float qry = d2[0], qrz = d2[1], m = d2[2], mtr = d2[3];
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float ri = __builtin_amdgcn_rsqf(r2);
                const float ri2 = ri * ri, ri3 = ri * ri2, ri4 = ri2 * ri2, ri5 = ri2 * ri3;
                const float rqr = __builtin_fmaf(qrz, dz, __builtin_fmaf(qry, dy, __builtin_fmaf(qrx, dx, mtr)));
                const float t = rqr * ri4;
                const float meff = __builtin_fmaf(t, 0.5f, m);
                const float meff3 = __builtin_fmaf(t, 2.5f, m) * ri3;
                p[s] = __builtin_fmaf(-meff, ri, p[s]);
                ax[s] = __builtin_fmaf(-meff3, dx, __builtin_fmaf(ri5, qrx, ax[s]));
                ay[s] = __builtin_fmaf(-meff3, dy, __builtin_fmaf(ri5, qry, ay[s]));
                az[s] = __builtin_fmaf(-meff3, dz, __builtin_fmaf(ri5, qrz, az[s]));
            }
        }
#pragma unroll
        for (int s = 0; s < S; s++) { A[s][0] += ax[s]; A[s][1] += ay[s]; A[s][2] += az[s]; A[s][3] += p[s]; }
    }
#pragma unroll
    for (int s = 0; s < S; s++)
        for (int k = 0; k < 4; k++) {
            double v = A[s][k];
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (g == 0) out[(size_t)(base + s * 16 + c16) * 4 + k] = (float)v;
        }
}

static double urand() { return rand() / (RAND_MAX + 1.0); }
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e_), __LINE__); return 1; } } while (0)

int main()
{
    // ---- layout check (exact small integers)
    {
        float A[16][4], B[4][16], a[64], b[64], d[256];
        for (int i = 0; i < 16; i++) for (int k = 0; k < 4; k++) A[i][k] = (float)((i * 7 + k * 3) % 11 - 5);
        for (int k = 0; k < 4; k++) for (int j = 0; j < 16; j++) B[k][j] = (float)((j * 5 + k * 2) % 9 - 4);
        for (int l = 0; l < 64; l++) { a[l] = A[l % 16][l / 16]; b[l] = B[l / 16][l % 16]; }
        float *da, *db, *dd;
        CK(hipMalloc(&da, 256)); CK(hipMalloc(&db, 256)); CK(hipMalloc(&dd, 1024));
        CK(hipMemcpy(da, a, 256, hipMemcpyHostToDevice)); CK(hipMemcpy(db, b, 256, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_layout, dim3(1), dim3(64), 0, 0, da, db, dd);
        CK(hipDeviceSynchronize());
        CK(hipMemcpy(d, dd, 1024, hipMemcpyDeviceToHost));
        int bad = 0;
        for (int l = 0; l < 64; l++) for (int v = 0; v < 4; v++) {
            const int i = 4 * (l / 16) + v, j = l % 16;
            float e = 0; for (int k = 0; k < 4; k++) e += A[i][k] * B[k][j];
            if (e != d[l * 4 + v]) bad++;
        }
        printf("layout check: %d mismatches of 256\n", bad);
        if (bad) return 2;
    }
    const int nj = 4096, nblk = nj / 4, waves = 8192;
    const int total_i = waves * 128;
    std::vector<float4> xi(total_i);
    for (auto &p : xi) p = make_float4((float)(urand() - 0.5) * 0.02f, (float)(urand() - 0.5) * 0.02f, (float)(urand() - 0.5) * 0.02f, 0.f);
    // records: far cells at distance 0.05..0.5, q' traceless and small
    std::vector<double> R(nj * 11);
    std::vector<float> rec(nj * 12), rec16(nj * 16);
    for (int j = 0; j < nj; j++) {
        double *r = &R[j * 11];
        double th = urand() * 6.283, ph = acos(2 * urand() - 1), d = 0.05 + 0.45 * urand();
        r[0] = d * sin(ph) * cos(th); r[1] = d * sin(ph) * sin(th); r[2] = d * cos(ph);
        r[3] = 1e-4 * (0.5 + urand());
        const double s = r[3] * 1e-4;
        r[4] = s * (urand() - 0.5); r[5] = s * (urand() - 0.5); r[6] = -(r[4] + r[5]);
        r[7] = s * (urand() - 0.5); r[8] = s * (urand() - 0.5); r[9] = s * (urand() - 0.5); r[10] = 0.0;
        for (int k = 0; k < 11; k++) rec[j * 12 + k] = (float)r[k];
        rec[j * 12 + 11] = 0.f;
        const float cx = (float)r[0], cy = (float)r[1], cz = (float)r[2];
        const float qxx = (float)r[4], qyy = (float)r[5], qzz = (float)r[6], qxy = (float)r[7], qyz = (float)r[8], qzx = (float)r[9];
        float *o = &rec16[j * 16];
        o[0] = -cx; o[1] = -cy; o[2] = -cz; o[3] = (float)r[3];
        o[4] = qxx; o[5] = qyy; o[6] = qzz; o[7] = qxy; o[8] = qyz; o[9] = qzx; o[10] = 0.f;
        o[11] = (float)-((double)qxx * cx + (double)qxy * cy + (double)qzx * cz);
        o[12] = (float)-((double)qxy * cx + (double)qyy * cy + (double)qyz * cz);
        o[13] = (float)-((double)qzx * cx + (double)qyz * cy + (double)qzz * cz);
        o[14] = 0.f; o[15] = 1.f;
    }
    // weight (q, k) -> record16 field: q = dx dy dz qrx | qry qrz m mtr ; k = x y z 1
    const int F[8][4] = {{15, 14, 14, 0}, {14, 15, 14, 1}, {14, 14, 15, 2}, {4, 7, 9, 11},
                         {7, 5, 8, 12}, {9, 8, 6, 13}, {14, 14, 14, 3}, {14, 14, 14, 10}};
    std::vector<int> fidx(128);
    std::vector<float2> W((size_t)nblk * 64);
    for (int l = 0; l < 64; l++) {
        const int q = l % 4, k = l / 16;
        fidx[l] = F[q][k]; fidx[64 + l] = F[4 + q][k];
    }
    for (int b = 0; b < nblk; b++)
        for (int l = 0; l < 64; l++) {
            const int j = 4 * b + (l % 16) / 4;
            W[b * 64 + l] = make_float2(rec16[j * 16 + fidx[l]], rec16[j * 16 + fidx[64 + l]]);
        }
    float4 *dxi; float *drec, *drec16, *dout; float2 *dW; int *dfi;
    CK(hipMalloc(&dxi, total_i * 16)); CK(hipMalloc(&drec, rec.size() * 4)); CK(hipMalloc(&drec16, rec16.size() * 4));
    CK(hipMalloc(&dW, W.size() * 8)); CK(hipMalloc(&dfi, 512)); CK(hipMalloc(&dout, (size_t)total_i * 16));
    CK(hipMemcpy(dxi, xi.data(), total_i * 16, hipMemcpyHostToDevice));
    CK(hipMemcpy(drec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(drec16, rec16.data(), rec16.size() * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(dW, W.data(), W.size() * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(dfi, fidx.data(), 512, hipMemcpyHostToDevice));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    std::vector<float> ref(total_i * 4), got(total_i * 4);
    struct KV { const char *name; int tile; void (*f)(const float4 *, const float *, int, float *); };
    auto time_it = [&](const char *name, int tile, auto launch) -> int {
        const int nw = total_i / tile;
        float best = 1e30f;
        for (int rep = 0; rep < 4; rep++) {
            CK(hipEventRecord(e0)); launch(nw / 4); CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
            float ms; CK(hipEventElapsedTime(&ms, e0, e1)); best = ms < best ? ms : best;
        }
        CK(hipGetLastError());
        CK(hipMemcpy(got.data(), dout, (size_t)total_i * 16, hipMemcpyDeviceToHost));
        double pairs = (double)total_i * nj;
        double md = 0; for (int i = 0; i < total_i; i++) {
            const double a = sqrt((double)ref[i*4]*ref[i*4] + (double)ref[i*4+1]*ref[i*4+1] + (double)ref[i*4+2]*ref[i*4+2]);
            const double dd = sqrt(pow(got[i*4]-ref[i*4],2) + pow(got[i*4+1]-ref[i*4+1],2) + pow(got[i*4+2]-ref[i*4+2],2));
            md = fmax(md, dd / a);
        }
        printf("%-16s tile %3d  %.3f ms  %.1f Gpairs/s  %.1f TFLOP/s(65/pair)  max acc rel diff vs valu2 %.2e\n", name, tile, best,
               pairs / (best * 1e-3) / 1e9, pairs * 65 / (best * 1e-3) / 1e12, md);
        return 0;
    };
    // reference output: valu IPL2
    hipLaunchKernelGGL(k_valu<2>, dim3(total_i / 128 / 4), dim3(256), 0, 0, dxi, drec, nj, dout);
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(ref.data(), dout, (size_t)total_i * 16, hipMemcpyDeviceToHost));
    if (time_it("valu ipl1", 64, [&](int nb) { hipLaunchKernelGGL(k_valu<1>, dim3(nb), dim3(256), 0, 0, dxi, drec, nj, dout); })) return 1;
    if (time_it("valu ipl2", 128, [&](int nb) { hipLaunchKernelGGL(k_valu<2>, dim3(nb), dim3(256), 0, 0, dxi, drec, nj, dout); })) return 1;
    if (time_it("valu ipl4", 256, [&](int nb) { hipLaunchKernelGGL(k_valu<4>, dim3(nb), dim3(256), 0, 0, dxi, drec, nj, dout); })) return 1;
    if (time_it("mfma S2", 32, [&](int nb) { hipLaunchKernelGGL((k_mfma<2, false>), dim3(nb), dim3(256), 0, 0, dxi, dW, drec16, dfi, nblk, dout); })) return 1;
    if (time_it("mfma S4", 64, [&](int nb) { hipLaunchKernelGGL((k_mfma<4, false>), dim3(nb), dim3(256), 0, 0, dxi, dW, drec16, dfi, nblk, dout); })) return 1;
    if (time_it("mfma S8", 128, [&](int nb) { hipLaunchKernelGGL((k_mfma<8, false>), dim3(nb), dim3(256), 0, 0, dxi, dW, drec16, dfi, nblk, dout); })) return 1;
    if (time_it("mfma S4 gather", 64, [&](int nb) { hipLaunchKernelGGL((k_mfma<4, true>), dim3(nb), dim3(256), 0, 0, dxi, dW, drec16, dfi, nblk, dout); })) return 1;
    if (time_it("mfma S8 gather", 128, [&](int nb) { hipLaunchKernelGGL((k_mfma<8, true>), dim3(nb), dim3(256), 0, 0, dxi, dW, drec16, dfi, nblk, dout); })) return 1;
    printf("done\n");
    return 0;
}
