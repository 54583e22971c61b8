// pth_noalloc.hip -- pth_resident_eval_noalloc: one complete soft-force
// evaluation, the work of pth_resident_create + pth_resident_run +
// pth_resident_results, with every device buffer it needs (the packed EPI /
// EPJ / SPJ records, the plan, the staged lists, the partial sums and the
// outputs) carved out of one file-scope device array, so the call allocates
// no device or pinned memory.  For callers that cannot allocate: tight
// serving loops, harnesses that forbid hipMalloc.
//
// The array `scratch` is a per-device global of the code object: each GPU
// the library runs on holds its own copy, at its own address, from module
// load to the end of the process.  Host code never passes `scratch` itself
// (on the host that name is a stand-in, not a device address): pth_create
// calls noalloc_setup once per device with that device current, which looks
// the address up with hipGetSymbolAddress and keeps it per device; the
// launches below take that address.  Every call on one device shares the
// array and leaves its contents behind; nothing in it has to start at zero
// (every byte a kernel reads was written earlier in the same call).  Calls
// that overlap on one device race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pth_internal.h"

static __device__ __align__(256) unsigned char scratch[PTH_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
void *g_scratch[MAX_DEVICES] = {};   // device address of `scratch` per device
int g_state[MAX_DEVICES] = {};       // 0 not looked up, 1 ready, -1 failed
std::string g_error[MAX_DEVICES];

__global__ __launch_bounds__(256) void k_pack_epi(long long n, const double *__restrict__ pos,
                                                  const double *__restrict__ rs, EpiDev *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    EpiDev d;
    d.x = pos[3 * i];
    d.y = pos[3 * i + 1];
    d.z = pos[3 * i + 2];
    d.rs = (float)rs[i];
    d.pad = 0;
    out[i] = d;
}

// EPJ records, and monopole superparticles (centre and mass, r_search 0) in
// the same record
__global__ __launch_bounds__(256) void k_pack_epj(long long n, const double *__restrict__ pos,
                                                  const double *__restrict__ mass, const double *__restrict__ rs,
                                                  EpjDev *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    EpjDev d;
    d.x = pos[3 * i];
    d.y = pos[3 * i + 1];
    d.z = pos[3 * i + 2];
    d.m = (float)mass[i];
    d.rs = rs ? (float)rs[i] : 0.f;
    out[i] = d;
}

__global__ __launch_bounds__(256) void k_pack_spj(long long n, const double *__restrict__ pos,
                                                  const double *__restrict__ mass, const double *__restrict__ quad,
                                                  SpjDev *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *q = quad + 6 * i;
    SpjDev d;
    d.x = pos[3 * i];
    d.y = pos[3 * i + 1];
    d.z = pos[3 * i + 2];
    d.m = (float)mass[i];
    d.qxx = (float)q[0];
    d.qyy = (float)q[1];
    d.qzz = (float)q[2];
    d.qxy = (float)q[3];
    d.qxz = (float)q[4];
    d.qyz = (float)q[5];
    d.pad[0] = d.pad[1] = d.pad[2] = 0.f;
    out[i] = d;
}

// outputs {acc, pot} + nnb into the caller's layouts: acc [n][3], pot [n],
// nnb [n] (64-bit)
__global__ __launch_bounds__(256) void k_unpack_out(long long n, const double4 *__restrict__ accp,
                                                    const int *__restrict__ nnb, double *__restrict__ acc,
                                                    double *__restrict__ pot, long long *__restrict__ nnb64)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 a = accp[i];
    acc[3 * i] = a.x;
    acc[3 * i + 1] = a.y;
    acc[3 * i + 2] = a.z;
    pot[i] = a.w;
    nnb64[i] = nnb[i];
}

unsigned grid(long long n) { return (unsigned)((n + 255) / 256); }

// Carves consecutive 256-byte aligned ranges out of the scratch array.
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        const size_t o = off;
        off = align256(off + bytes);
        return o;
    }
};

}  // namespace

namespace pth {

void noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_state[device] != 0) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_state[device] = -1;
        g_error[device] = std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorName(e);
        return;
    }
    g_scratch[device] = p;
    g_state[device] = 1;
}

}  // namespace pth

extern "C" int pth_resident_eval_noalloc(pth_ctx *c, int n_walk, const int *n_epi, const int *n_epj,
                                         const int *n_spj, const double *epi_pos, const double *epi_rs,
                                         const int *id_epj, const int *id_spj, int n_epj_tot,
                                         const double *epj_pos, const double *epj_mass, const double *epj_rs,
                                         int n_spj_tot, const double *spj_pos, const double *spj_mass,
                                         const double *spj_quad, double *acc, double *pot, long long *nnb)
{
    static const char *who = "pth_resident_eval_noalloc";
    if (!c) return PTH_ERR_ARG;
    const std::string W(who);
    if (c->par.changeover) return c->fail(PTH_ERR_STATE, W + ": linear-cutoff mode only (the context is in changeover mode)");
    if (c->par.exec_mode != PTH_EXEC_FUSED) return c->fail(PTH_ERR_STATE, W + ": fused form only");
    if (n_walk < 0 || n_epj_tot < 0 || n_spj_tot < 0) return c->fail(PTH_ERR_ARG, W + ": negative count");
    if (n_walk > 0 && (!n_epi || !n_epj || !n_spj)) return c->fail(PTH_ERR_ARG, W + ": null walk sizes");
    const bool mono = c->par.spj_monopole != 0;
    long long ni = 0, n_ide = 0, n_ids = 0;
    for (int w = 0; w < n_walk; w++) {
        if (n_epi[w] < 0 || n_epj[w] < 0 || n_spj[w] < 0)
            return c->fail(PTH_ERR_ARG, W + ": negative size in walk " + std::to_string(w));
        ni += n_epi[w];
        n_ide += n_epj[w];
        n_ids += n_spj[w];
    }
    if ((ni > 0 && (!epi_pos || !epi_rs)) || (n_ide > 0 && !id_epj) || (n_ids > 0 && !id_spj) ||
        (n_epj_tot > 0 && (!epj_pos || !epj_mass || !epj_rs)) ||
        (n_spj_tot > 0 && (!spj_pos || !spj_mass || (!spj_quad && !mono))))
        return c->fail(PTH_ERR_ARG, W + ": null input array");
    for (long long k = 0; k < n_ide; k++)
        if (id_epj[k] < 0 || id_epj[k] >= n_epj_tot) return c->fail(PTH_ERR_ARG, W + ": EPJ index out of range");
    for (long long k = 0; k < n_ids; k++)
        if (id_spj[k] < 0 || id_spj[k] >= n_spj_tot) return c->fail(PTH_ERR_ARG, W + ": SPJ index out of range");

    const int dev = c->device;
    if (dev < 0 || dev >= MAX_DEVICES) return c->fail(PTH_ERR_STATE, W + ": device index beyond the scratch table");
    void *base_p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_state[dev] != 1)
            return c->fail(PTH_ERR_HIP, W + ": " + (g_state[dev] < 0 ? g_error[dev] : std::string("scratch not set up")));
        base_p = g_scratch[dev];
    }
    char *base = static_cast<char *>(base_p);

    // the plan: one segment on the context's stream (a walk's items and
    // results do not depend on the segmentation)
    Plan &P = c->na_plan;
    if (!make_plan(n_walk, n_epi, n_epj, n_spj, PTH_EXEC_FUSED, c->par.chunk_j, c->par.max_ipl, P, 1, mono))
        return c->fail(PTH_ERR_CAPACITY, W + ": sizes exceed the int32 index range");
    Arena A{};
    A.layout(P, n_ide, n_ids, false);
    const bool df = df_enabled(c->par) != 0;
    const long long nw = (long long)P.walks.size();
    const long long ni1 = std::max<long long>(ni, 1), nj1 = std::max(n_epj_tot, 1), ns1 = std::max(n_spj_tot, 1);
    // inputs and outputs for the whole call
    Carve cv;
    const size_t o_arena = cv.take(A.total);
    const size_t o_epj = cv.take(sizeof(EpjDev) * nj1);
    const size_t o_spj = cv.take((mono ? sizeof(EpjDev) : sizeof(SpjDev)) * ns1);
    const size_t o_tab = cv.take(mono ? 0 : sizeof(SpjTab) * ns1);
    const size_t o_accp = cv.take(sizeof(double4) * ni1);
    const size_t o_nnb = cv.take(sizeof(int) * ni1);
    const size_t persist = cv.off;
    // then, in turn over one region: the caller's FP64 arrays while they are
    // packed, the work buffers of the evaluation, the unpacked outputs
    Carve r1{persist};
    const size_t o_ipos = r1.take(24 * (size_t)ni1), o_irs = r1.take(8 * (size_t)ni1);
    const size_t o_jpos = r1.take(24 * (size_t)nj1), o_jm = r1.take(8 * (size_t)nj1), o_jrs = r1.take(8 * (size_t)nj1);
    const size_t o_spos = r1.take(24 * (size_t)ns1), o_sm = r1.take(8 * (size_t)ns1);
    const size_t o_sq = r1.take(mono ? 0 : 48 * (size_t)ns1);
    Carve r2{persist};
    const long long slots = std::max<long long>(P.n_e_slots, 8), recs = std::max<long long>(P.n_s_rec, 1);
    const long long parts = std::max(P.n_part, 1);
    const size_t o_centre = r2.take(sizeof(double4) * std::max<long long>(nw, 1));
    const size_t o_nearb = r2.take(sizeof(int) * std::max<long long>(nw, 1));
    const size_t o_si = r2.take(sizeof(float4) * ni1);
    const size_t o_se = r2.take(sizeof(float) * 5 * slots);
    const size_t o_ss = r2.take(sizeof(float) * PTH_SPJ_STAGED * recs);
    const size_t o_pa = r2.take(sizeof(double4) * 64 * parts);
    const size_t o_pn = r2.take(sizeof(int) * 64 * parts);
    const size_t o_il = r2.take(df ? sizeof(float4) * ni1 : 0);
    const size_t o_el = r2.take(df ? sizeof(float) * (PTH_DF_EXTRA / 8) * slots : 0);
    Carve r3{persist};
    const size_t o_acc = r3.take(24 * (size_t)ni1), o_pot = r3.take(8 * (size_t)ni1), o_nnb64 = r3.take(8 * (size_t)ni1);
    const size_t need = std::max(r1.off, std::max(r2.off, r3.off));
    if (need > (size_t)PTH_NOALLOC_SCRATCH_BYTES)
        return c->fail(PTH_ERR_CAPACITY, W + ": the evaluation needs " + std::to_string(need) +
                                             " bytes of scratch, more than its " +
                                             std::to_string((long long)PTH_NOALLOC_SCRATCH_BYTES) + "-byte maximum");

    HIPCHK(c, hipSetDevice(dev), "hipSetDevice");
    hipStream_t st = c->stream;
    // the batch hooks may still be using the stream: wait for them first
    HIPCHK(c, hipStreamSynchronize(st), "hipStreamSynchronize");
    auto up = [&](size_t off, const void *src, size_t bytes) {
        return bytes ? hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
    };
    hipError_t e;
    if ((e = up(o_arena + A.walks, P.walks.data(), sizeof(WalkDev) * P.walks.size())) ||
        (e = up(o_arena + A.es, P.es_off.data(), sizeof(long long) * P.es_off.size())) ||
        (e = up(o_arena + A.ss, P.ss_off.data(), sizeof(long long) * P.ss_off.size())) ||
        (e = up(o_arena + A.items, P.items.data(), sizeof(Item) * P.items.size())) ||
        (e = up(o_arena + A.ritems, P.ritems.data(), sizeof(RItem) * P.ritems.size())) ||
        (e = up(o_arena + A.ide, id_epj, sizeof(int) * (size_t)n_ide)) ||
        (e = up(o_arena + A.ids, id_spj, sizeof(int) * (size_t)n_ids)) ||
        (e = up(o_ipos, epi_pos, 24 * (size_t)ni)) || (e = up(o_irs, epi_rs, 8 * (size_t)ni)) ||
        (e = up(o_jpos, epj_pos, 24 * (size_t)n_epj_tot)) || (e = up(o_jm, epj_mass, 8 * (size_t)n_epj_tot)) ||
        (e = up(o_jrs, epj_rs, 8 * (size_t)n_epj_tot)) || (e = up(o_spos, spj_pos, 24 * (size_t)n_spj_tot)) ||
        (e = up(o_sm, spj_mass, 8 * (size_t)n_spj_tot)) || (e = up(o_sq, mono ? nullptr : spj_quad, mono ? 0 : 48 * (size_t)n_spj_tot)))
        return c->hip_fail(e, "scratch upload");
    if (ni > 0)
        hipLaunchKernelGGL(k_pack_epi, dim3(grid(ni)), dim3(256), 0, st, ni, (const double *)(base + o_ipos),
                           (const double *)(base + o_irs), (EpiDev *)(base + o_arena + A.epi));
    if (n_epj_tot > 0)
        hipLaunchKernelGGL(k_pack_epj, dim3(grid(n_epj_tot)), dim3(256), 0, st, (long long)n_epj_tot,
                           (const double *)(base + o_jpos), (const double *)(base + o_jm),
                           (const double *)(base + o_jrs), (EpjDev *)(base + o_epj));
    if (n_spj_tot > 0) {
        if (mono)
            hipLaunchKernelGGL(k_pack_epj, dim3(grid(n_spj_tot)), dim3(256), 0, st, (long long)n_spj_tot,
                               (const double *)(base + o_spos), (const double *)(base + o_sm), (const double *)nullptr,
                               (EpjDev *)(base + o_spj));
        else
            hipLaunchKernelGGL(k_pack_spj, dim3(grid(n_spj_tot)), dim3(256), 0, st, (long long)n_spj_tot,
                               (const double *)(base + o_spos), (const double *)(base + o_sm),
                               (const double *)(base + o_sq), (SpjDev *)(base + o_spj));
    }
    HIPCHK(c, hipGetLastError(), "packing kernels");
    if (!mono) HIPCHK(c, launch_spj_table((const SpjDev *)(base + o_spj), n_spj_tot, (float)c->par.eps2,
                                          (SpjTab *)(base + o_tab), st), "SPJ table kernel");

    // the launch descriptor over scratch ranges (DevBuf views: never grown
    // or released)
    DevBuf arena;
    arena.p = base + o_arena;
    arena.cap = A.total;
    Work Wk;
    Wk.centre.p = base + o_centre;
    Wk.near_blk.p = base + o_nearb;
    Wk.staged_i.p = base + o_si;
    Wk.staged_e.p = base + o_se;
    Wk.staged_s.p = base + o_ss;
    Wk.part_accp.p = base + o_pa;
    Wk.part_nnb.p = base + o_pn;
    if (df) {
        Wk.staged_il.p = base + o_il;
        Wk.staged_el.p = base + o_el;
    }
    BatchDev &b = c->na_batch;
    fill_batch(b, P, c->par, arena, A, Wk, (const EpjDev *)(base + o_epj), (const SpjDev *)(base + o_spj),
               (double4 *)(base + o_accp), (int *)(base + o_nnb), nullptr, nullptr,
               mono ? nullptr : (const SpjTab *)(base + o_tab));
    HIPCHK(c, launch_all(b, st), "evaluation kernels");
    if (ni > 0) {
        hipLaunchKernelGGL(k_unpack_out, dim3(grid(ni)), dim3(256), 0, st, ni, (const double4 *)(base + o_accp),
                           (const int *)(base + o_nnb), (double *)(base + o_acc), (double *)(base + o_pot),
                           (long long *)(base + o_nnb64));
        HIPCHK(c, hipGetLastError(), "output kernel");
        if (acc) HIPCHK(c, hipMemcpyAsync(acc, base + o_acc, 24 * (size_t)ni, hipMemcpyDeviceToHost, st), "result download");
        if (pot) HIPCHK(c, hipMemcpyAsync(pot, base + o_pot, 8 * (size_t)ni, hipMemcpyDeviceToHost, st), "result download");
        if (nnb) HIPCHK(c, hipMemcpyAsync(nnb, base + o_nnb64, 8 * (size_t)ni, hipMemcpyDeviceToHost, st), "result download");
    }
    HIPCHK(c, hipStreamSynchronize(st), "hipStreamSynchronize");
    HIPCHK(c, hipGetLastError(), "kernel execution");
    return PTH_OK;
}
