// pth_context.cpp -- host side of libpetar_hip.so: contexts, the FDPS
// multi-walk hooks (send / dispatch / retrieve) and device-resident
// evaluations.  See include/petar_hip.h for the interface and the reference
// functions each entry point replaces, pth_internal.h for the shared state.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <immintrin.h>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pth_internal.h"

extern "C" {

void pth_params_default(pth_params *p)
{
    if (!p) return;
    p->eps2 = 0.0;
    p->rcut2 = 0.0;
    p->G = 1.0;
    p->nnb_mode = 0;
    p->exec_mode = PTH_EXEC_FUSED;
    p->chunk_j = 0;
    p->max_walk = 1000;
    p->max_ipl = 2;  // measured best on the precision workload (steps of 2, 3, 4: 11.05, 11.58, 12.12 ms)
    p->j_vgpr = 0;   // SGPR operands beat VGPR copies at every max_ipl
    p->changeover = 0;
    p->ngb_factor = 0.99;
    p->segments = 8;  // 1 staging + 2 force streams; 16 on 4 streams is faster with GPU_MAX_HW_QUEUES=8, but
    p->streams = 3;   // with HIP's default 4 queues a process's first such pipeline ran 4% slower
    p->df_near = 1;
    p->spj_monopole = 0;
    p->graph = 0;
    p->min_seg_pairs = 0.0;
    p->persistent = 0;
}

const char *pth_last_error(const pth_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pth_device(const pth_ctx *ctx) { return ctx ? ctx->device : -1; }

static double wall_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static thread_local std::string g_create_err;
const char *pth_create_error(void) { return g_create_err.c_str(); }

int pth_create(pth_ctx **out, int rank, const pth_params *params)
{
    if (!out) {
        g_create_err = "pth_create: null output pointer";
        return PTH_ERR_ARG;
    }
    *out = nullptr;
    pth_params par;
    if (params) par = *params;
    else pth_params_default(&par);
    if (par.nnb_mode != 0 && par.nnb_mode != 1) {
        g_create_err = "pth_create: nnb_mode must be 0 or 1";
        return PTH_ERR_ARG;
    }
    if (par.exec_mode != PTH_EXEC_FUSED && par.exec_mode != PTH_EXEC_PLAIN) {
        g_create_err = "pth_create: unknown exec_mode";
        return PTH_ERR_ARG;
    }
    if (!(par.rcut2 >= 0.0) || !(par.eps2 >= 0.0)) {
        g_create_err = "pth_create: eps2 and rcut2 must be >= 0";
        return PTH_ERR_ARG;
    }
    if (par.changeover != 0 && par.changeover != 1) {
        g_create_err = "pth_create: changeover must be 0 or 1";
        return PTH_ERR_ARG;
    }
    if (par.changeover && par.exec_mode != PTH_EXEC_FUSED) {
        g_create_err = "pth_create: changeover mode needs the fused form";
        return PTH_ERR_ARG;
    }
    if (!(par.ngb_factor >= 0.0)) {
        g_create_err = "pth_create: ngb_factor must be >= 0";
        return PTH_ERR_ARG;
    }
    if (par.max_walk <= 0) par.max_walk = 1000;
    par.df_near = par.df_near ? 1 : 0;
    par.spj_monopole = par.spj_monopole ? 1 : 0;
    par.graph = par.graph ? 1 : 0;
    par.min_seg_pairs = par.min_seg_pairs > 0 ? par.min_seg_pairs : 0.0;
    par.segments = std::min(std::max(par.segments, 1), PTH_MAX_SEGMENTS);
    par.streams = std::min(std::max(par.streams, 1), 8);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        g_create_err = std::string("pth_create: no HIP device (") +
                       (e != hipSuccess ? hipGetErrorName(e) : "count 0") + ")";
        return PTH_ERR_HIP;
    }
    pth_ctx *c = new pth_ctx;
    c->par = par;
    {
        // how pth_dispatch sends index lists (a diagnostic switch):
        // PTH_LIST_RUNS=0 as they are, =1 always as runs, unset: as runs
        // when that at least halves their bytes
        const char *lr = std::getenv("PTH_LIST_RUNS");
        c->list_runs = (lr && lr[0] == '0') ? 0 : (lr && lr[0] == '1') ? 1 : -1;
        const char *cs = std::getenv("PTH_CLASS_STREAMS");
        c->class_streams = !(cs && cs[0] == '0');
        const char *zc = std::getenv("PTH_ZERO_COPY");
        c->zero_copy = !(zc && zc[0] == '0');
    }
    c->device = (rank >= 0 ? rank : -rank) % ndev;  // src/force_gpu_cuda.cu:550-553
    if ((e = hipSetDevice(c->device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_send, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_upload, hipEventDisableTiming)) != hipSuccess) {
        g_create_err = std::string("pth_create: ") + hipGetErrorName(e);
        pth_destroy(c);
        return PTH_ERR_HIP;
    }
    for (int t = 0; t < MAX_TAG; t++) {
        if ((e = hipEventCreateWithFlags(&c->tag[t].done, hipEventDisableTiming)) != hipSuccess) {
            g_create_err = std::string("pth_create: ") + hipGetErrorName(e);
            pth_destroy(c);
            return PTH_ERR_HIP;
        }
    }
    // the scratch array of pth_resident_eval_noalloc on this device (a
    // failure is reported by that entry point, not here)
    noalloc_setup(c->device);
    *out = c;
    return PTH_OK;
}

int pth_set_params(pth_ctx *c, const pth_params *p)
{
    if (!c) return PTH_ERR_ARG;
    if (!p) return c->fail(PTH_ERR_ARG, "pth_set_params: null params");
    if (p->nnb_mode != 0 && p->nnb_mode != 1) return c->fail(PTH_ERR_ARG, "pth_set_params: nnb_mode must be 0 or 1");
    if (p->exec_mode != PTH_EXEC_FUSED && p->exec_mode != PTH_EXEC_PLAIN)
        return c->fail(PTH_ERR_ARG, "pth_set_params: unknown exec_mode");
    if (!(p->rcut2 >= 0.0) || !(p->eps2 >= 0.0))
        return c->fail(PTH_ERR_ARG, "pth_set_params: eps2 and rcut2 must be >= 0");
    if (p->changeover != 0 && p->changeover != 1)
        return c->fail(PTH_ERR_ARG, "pth_set_params: changeover must be 0 or 1");
    if (p->changeover && p->exec_mode != PTH_EXEC_FUSED)
        return c->fail(PTH_ERR_ARG, "pth_set_params: changeover mode needs the fused form");
    if (!(p->ngb_factor >= 0.0)) return c->fail(PTH_ERR_ARG, "pth_set_params: ngb_factor must be >= 0");
    // every value is read when a batch is dispatched (its plan and launch
    // arguments are made then), so batches already in flight keep theirs
    c->par.eps2 = p->eps2;
    c->par.rcut2 = p->rcut2;
    c->par.G = p->G;
    c->par.nnb_mode = p->nnb_mode;
    c->par.exec_mode = p->exec_mode;
    c->par.chunk_j = p->chunk_j;
    c->par.max_walk = p->max_walk > 0 ? p->max_walk : 1000;
    c->par.max_ipl = p->max_ipl;
    c->par.j_vgpr = p->j_vgpr;
    c->par.changeover = p->changeover;
    c->par.ngb_factor = p->ngb_factor;
    c->par.segments = std::min(std::max(p->segments, 1), PTH_MAX_SEGMENTS);
    c->par.streams = std::min(std::max(p->streams, 1), 8);
    c->par.df_near = p->df_near ? 1 : 0;
    c->par.spj_monopole = p->spj_monopole ? 1 : 0;
    c->par.graph = p->graph ? 1 : 0;
    c->par.min_seg_pairs = p->min_seg_pairs > 0 ? p->min_seg_pairs : 0.0;
    c->par.persistent = p->persistent ? 1 : 0;
    return PTH_OK;
}

int pth_get_params(const pth_ctx *c, pth_params *p)
{
    if (!c || !p) return PTH_ERR_ARG;
    *p = c->par;
    return PTH_OK;
}

void pth_destroy(pth_ctx *c)
{
    if (!c) return;
    if (c->device >= 0) (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->d_epj.release();
    c->d_spj.release();
    c->d_epj_co.release();
    c->h_epj.release();
    c->h_spj.release();
    c->h_epj_co.release();
    c->h_arena.release();
    c->d_arena.release();
    c->work.release();
    for (int t = 0; t < MAX_TAG; t++) {
        c->tag[t].out_accp.release();
        c->tag[t].out_nnb.release();
        c->tag[t].out_pott.release();
        c->tag[t].h_accp.release();
        c->tag[t].h_nnb.release();
        c->tag[t].h_pott.release();
        c->tag[t].h_err.release();
        if (c->tag[t].done) (void)hipEventDestroy(c->tag[t].done);
        for (hipEvent_t &x : c->tag[t].prof)
            if (x) (void)hipEventDestroy(x);
    }
    for (hipEvent_t &x : c->prof_send)
        if (x) (void)hipEventDestroy(x);
    c->d_ids.release();
    c->d_spj_tab.release();
    if (c->ev_send) (void)hipEventDestroy(c->ev_send);
    if (c->ev_upload) (void)hipEventDestroy(c->ev_upload);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"

namespace {

// Runs of consecutive indices of one list p[0 .. n): run k starts at
// position rpos[k] with index rid[k]; returns the number of runs and the
// smallest and largest index (for the range check).  rid / rpos have room
// for n entries: every entry writes the slot after the last run (at most one
// run per entry, so the writes stay inside), and only a run start advances
// the count, so no branch depends on the data.
int encode_runs_scalar(const int *p, int n, int *rid, int *rpos, int &mn, int &mx)
{
    int nr = 0;
    mn = 0;
    mx = -1;
    long long next = -1;
    for (int j = 0; j < n; j++) {
        const int id = p[j];
        mn = std::min(mn, id);
        mx = std::max(mx, id);
        rid[nr] = id;
        rpos[nr] = j;
        nr += (id != next);
        next = (long long)id + 1;
    }
    return nr;
}

// The same with AVX-512: 16 entries at a time, run starts as a lane mask,
// the starts compressed in registers and stored 16 lanes wide (nr <= j, so
// the 16 lanes at nr stay inside the n entries while j + 16 <= n).
__attribute__((target("avx512f"))) int encode_runs_avx512(const int *p, int n, int *rid, int *rpos, int &mn,
                                                              int &mx)
{
    if (n < 16) return encode_runs_scalar(p, n, rid, rpos, mn, mx);
    const __m512i one = _mm512_set1_epi32(1), step = _mm512_set1_epi32(16);
    __m512i pos = _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
    __m512i vmn = _mm512_set1_epi32(INT_MAX), vmx = _mm512_set1_epi32(INT_MIN);
    int nr = 0;
    // entry j - 1 of the first vector's lane 0: p[0] - 2, so lane 0 starts a run
    __m512i prev = _mm512_alignr_epi32(_mm512_loadu_si512(p), _mm512_set1_epi32(p[0] - 2), 15);
    int j = 0;
    for (; j + 16 <= n; j += 16) {
        const __m512i v = _mm512_loadu_si512(p + j);
        if (j > 0) prev = _mm512_loadu_si512(p + j - 1);
        const __mmask16 m = _mm512_cmpneq_epi32_mask(v, _mm512_add_epi32(prev, one));
        _mm512_storeu_si512(rid + nr, _mm512_maskz_compress_epi32(m, v));
        _mm512_storeu_si512(rpos + nr, _mm512_maskz_compress_epi32(m, pos));
        nr += __builtin_popcount((unsigned)m);
        vmn = _mm512_min_epi32(vmn, v);
        vmx = _mm512_max_epi32(vmx, v);
        pos = _mm512_add_epi32(pos, step);
    }
    mn = std::min(0, _mm512_reduce_min_epi32(vmn));
    mx = _mm512_reduce_max_epi32(vmx);
    long long next = (long long)p[j - 1] + 1;
    for (; j < n; j++) {
        const int id = p[j];
        mn = std::min(mn, id);
        mx = std::max(mx, id);
        rid[nr] = id;
        rpos[nr] = j;
        nr += (id != next);
        next = (long long)id + 1;
    }
    return nr;
}

int encode_runs(const int *p, int n, int *rid, int *rpos, int &mn, int &mx)
{
    static const bool avx512 = __builtin_cpu_supports("avx512f");
    return avx512 ? encode_runs_avx512(p, n, rid, rpos, mn, mx) : encode_runs_scalar(p, n, rid, rpos, mn, mx);
}

// One EPJ, monopole-superparticle or quadrupole-superparticle record from
// the caller's struct through its view.
inline EpjDev epj_record(const char *r, const pth_epj_view *ev)
{
    const double *p = reinterpret_cast<const double *>(r + ev->off_pos);
    EpjDev d;
    d.x = p[0];
    d.y = p[1];
    d.z = p[2];
    d.m = (float)*reinterpret_cast<const double *>(r + ev->off_mass);
    d.rs = (float)*reinterpret_cast<const double *>(r + ev->off_rsearch);
    return d;
}
inline EpjDev spj_mono_record(const char *r, const pth_spj_view *sv)
{
    const double *p = reinterpret_cast<const double *>(r + sv->off_pos);
    EpjDev d;
    d.x = p[0];
    d.y = p[1];
    d.z = p[2];
    d.m = (float)*reinterpret_cast<const double *>(r + sv->off_mass);
    d.rs = 0.f;
    return d;
}
inline SpjDev spj_record(const char *r, const pth_spj_view *sv)
{
    const double *p = reinterpret_cast<const double *>(r + sv->off_pos);
    SpjDev d;
    d.x = p[0];
    d.y = p[1];
    d.z = p[2];
    d.m = (float)*reinterpret_cast<const double *>(r + sv->off_mass);
    d.qxx = (float)*reinterpret_cast<const double *>(r + sv->off_qxx);
    d.qyy = (float)*reinterpret_cast<const double *>(r + sv->off_qyy);
    d.qzz = (float)*reinterpret_cast<const double *>(r + sv->off_qzz);
    d.qxy = (float)*reinterpret_cast<const double *>(r + sv->off_qxy);
    d.qxz = (float)*reinterpret_cast<const double *>(r + sv->off_qxz);
    d.qyz = (float)*reinterpret_cast<const double *>(r + sv->off_qyz);
    d.pad[0] = d.pad[1] = d.pad[2] = 0.f;
    return d;
}
inline EpjCoDev epj_co_record(const char *r, const pth_epj_view *ev)
{
    return co_record(*reinterpret_cast<const double *>(r + ev->off_rin),
                     *reinterpret_cast<const double *>(r + ev->off_rout),
                     *reinterpret_cast<const long long *>(r + ev->off_id),
                     ev->off_group >= 0 ? reinterpret_cast<const double *>(r + ev->off_group) : nullptr);
}

// The j arrays of a force call: waits for the previous call's use of the
// buffers, sizes them, packs record k (k < n_epj_tot, s < n_spj_tot) with
// rec_e(k) / rec_s(s) giving the caller's record, uploads them and marks
// them as sent.
template <class RE, class RS>
int send_records(pth_ctx *c, int n_epj_tot, const pth_epj_view *ev, int n_spj_tot, const pth_spj_view *sv,
                 int n_walk, RE rec_e, RS rec_s)
{
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    // the pinned staging and the resident arrays may still be in use by the
    // previous force call's copies and kernels
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    const bool mono = c->par.spj_monopole != 0;
    const bool co = c->par.changeover != 0;
    const size_t be = sizeof(EpjDev) * (size_t)std::max(n_epj_tot, 1);
    // monopole superparticles are kept as EPJ-like records (centre, mass)
    const size_t bs = (mono ? sizeof(EpjDev) : sizeof(SpjDev)) * (size_t)std::max(n_spj_tot, 1);
    HIPCHK(c, c->h_epj.ensure(be), "pinned EPJ allocation");
    HIPCHK(c, c->h_spj.ensure(bs), "pinned SPJ allocation");
    HIPCHK(c, c->d_epj.ensure(be), "device EPJ allocation");
    HIPCHK(c, c->d_spj.ensure(bs), "device SPJ allocation");
    if (co) {
        const size_t bc = sizeof(EpjCoDev) * (size_t)std::max(n_epj_tot, 1);
        HIPCHK(c, c->h_epj_co.ensure(bc), "pinned EPJ allocation");
        HIPCHK(c, c->d_epj_co.ensure(bc), "device EPJ allocation");
    }
    const double t_pack = wall_s();
    EpjDev *he = c->h_epj.as<EpjDev>();
    EpjCoDev *hc = c->h_epj_co.as<EpjCoDev>();
    rec_e(he, co ? hc : nullptr, ev);
    if (mono) rec_s(c->h_spj.as<EpjDev>(), (SpjDev *)nullptr, sv);
    else rec_s((EpjDev *)nullptr, c->h_spj.as<SpjDev>(), sv);
    (void)n_walk;
    if (co && n_epj_tot > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_epj_co.p, hc, sizeof(EpjCoDev) * n_epj_tot, hipMemcpyHostToDevice, c->stream),
               "EPJ upload");
    c->co_sent = co;
    c->mono_sent = mono;
    c->prof.copy_s += wall_s() - t_pack;
    if (c->prof_on) HIPCHK(c, hipEventRecord(c->prof_send[0], c->stream), "hipEventRecord");
    if (n_epj_tot > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_epj.p, he, sizeof(EpjDev) * n_epj_tot, hipMemcpyHostToDevice, c->stream),
               "EPJ upload");
    if (n_spj_tot > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_spj.p, c->h_spj.p, (mono ? sizeof(EpjDev) : sizeof(SpjDev)) * n_spj_tot,
                                 hipMemcpyHostToDevice, c->stream), "SPJ upload");
    if (c->prof_on) {
        HIPCHK(c, hipEventRecord(c->prof_send[1], c->stream), "hipEventRecord");
        c->prof_send_pending = true;
    }
    HIPCHK(c, hipEventRecord(c->ev_send, c->stream), "hipEventRecord");
    if (c->par.exec_mode == PTH_EXEC_PLAIN) HIPCHK(c, hipStreamSynchronize(c->stream), "send sync");
    c->n_epj = n_epj_tot;
    c->n_spj = n_spj_tot;
    c->j_sent = true;
    c->tab_valid = false;
    return PTH_OK;
}

}  // namespace

extern "C" {

// ---- send_flag == true: src/force_gpu_cuda.cu:577-620 -----------------------
int pth_send_j(pth_ctx *c, const void *epj, int n_epj_tot, const pth_epj_view *ev, const void *spj,
               int n_spj_tot, const pth_spj_view *sv)
{
    if (!c) return PTH_ERR_ARG;
    if (n_epj_tot < 0 || n_spj_tot < 0) return c->fail(PTH_ERR_ARG, "pth_send_j: negative count");
    if ((n_epj_tot > 0 && (!epj || !ev)) || (n_spj_tot > 0 && (!spj || !sv)))
        return c->fail(PTH_ERR_ARG, "pth_send_j: null records or view");
    const char *eb = static_cast<const char *>(epj);
    const char *sb = static_cast<const char *>(spj);
    return send_records(
        c, n_epj_tot, ev, n_spj_tot, sv, 0,
        [&](EpjDev *he, EpjCoDev *hc, const pth_epj_view *v) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n_epj_tot; i++) {
                const char *r = eb + (size_t)i * v->stride;
                he[i] = epj_record(r, v);
                if (hc) hc[i] = epj_co_record(r, v);
            }
        },
        [&](EpjDev *hm, SpjDev *hs, const pth_spj_view *v) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n_spj_tot; i++) {
                const char *r = sb + (size_t)i * v->stride;
                if (hm) hm[i] = spj_mono_record(r, v);
                else hs[i] = spj_record(r, v);
            }
        });
}

// ---- send_flag == false: src/force_gpu_cuda.cu:621-699 -----------------------
int pth_dispatch(pth_ctx *c, int tag, int n_walk, const void *const *epi, const int *n_epi,
                 const pth_epi_view *iv, const int *const *id_epj, const int *n_epj,
                 const int *const *id_spj, const int *n_spj)
{
    if (!c) return PTH_ERR_ARG;
    if (tag < 0 || tag >= MAX_TAG) return c->fail(PTH_ERR_ARG, "pth_dispatch: tag out of range [0,8)");
    if (n_walk < 0 || n_walk > c->par.max_walk)
        return c->fail(PTH_ERR_CAPACITY, "pth_dispatch: n_walk " + std::to_string(n_walk) +
                                             " outside [0, " + std::to_string(c->par.max_walk) + "]");
    if (!c->j_sent) return c->fail(PTH_ERR_STATE, "pth_dispatch: no j arrays sent (send_flag=true call missing)");
    if (n_walk > 0 && (!epi || !n_epi || !iv || !id_epj || !n_epj || !id_spj || !n_spj))
        return c->fail(PTH_ERR_ARG, "pth_dispatch: null argument");
    TagState &T = c->tag[tag];
    if (T.pending) return c->fail(PTH_ERR_STATE, "pth_dispatch: previous batch of this tag not retrieved");
    // where each walk's i records, EPJ and SPJ list entries start
    c->w_off.resize(3 * (size_t)(n_walk + 1));
    long long *oi = c->w_off.data(), *oe = oi + (n_walk + 1), *os = oe + (n_walk + 1);
    oi[0] = oe[0] = os[0] = 0;
    for (int w = 0; w < n_walk; w++) {
        if (n_epi[w] < 0 || n_epj[w] < 0 || n_spj[w] < 0)
            return c->fail(PTH_ERR_ARG, "pth_dispatch: negative size in walk " + std::to_string(w));
        oi[w + 1] = oi[w] + n_epi[w];
        oe[w + 1] = oe[w] + n_epj[w];
        os[w + 1] = os[w] + n_spj[w];
    }
    const long long ni_tot = oi[n_walk], n_ide = oe[n_walk], n_ids = os[n_walk];
    if (ni_tot > INT_MAX || n_ide > INT_MAX || n_ids > INT_MAX)
        return c->fail(PTH_ERR_CAPACITY, "pth_dispatch: batch exceeds the int32 index range");
    const bool co = c->par.changeover != 0;
    if (co && !c->co_sent)
        return c->fail(PTH_ERR_STATE, "pth_dispatch: changeover mode, but the j arrays were sent without it");
    if ((c->par.spj_monopole != 0) != c->mono_sent)
        return c->fail(PTH_ERR_STATE, "pth_dispatch: spj_monopole differs from the last pth_send_j");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    // the pinned arena may still feed the previous batch's copy
    HIPCHK(c, hipEventSynchronize(c->ev_upload), "hipEventSynchronize");
    // the i records lead the arena: packed before its full size is known
    HIPCHK(c, c->h_arena.ensure(Arena::i_part(ni_tot, co)), "pinned arena allocation");
    double t_pack = wall_s();
    const bool encode = c->list_runs != 0;
    const int nj_lim = c->n_epj, ns_lim = c->n_spj;
    if (encode && (long long)c->h_rid.size() < n_ide + n_ids) {
        c->h_rid.resize((size_t)(n_ide + n_ids));
        c->h_rpos.resize((size_t)(n_ide + n_ids));
    }
    c->run_ne.assign(n_walk + 1, 0);
    c->run_ns.assign(n_walk + 1, 0);
    int *hrid = c->h_rid.data(), *hrpos = c->h_rpos.data();
    EpiDev *he = c->h_arena.as<EpiDev>();
    long long *hid = reinterpret_cast<long long *>(c->h_arena.as<char>() + Arena::i_id_offset(ni_tot));
    Plan &P = c->d_plan;
    bool plan_ok = true;
    int bad = -1;
    // one thread makes the plan while the others validate the index lists,
    // encode them as runs of consecutive indices (FDPS appends whole leaf
    // cells, so EPJ lists are mostly runs; each walk's runs go where its
    // list would start) and pack the i records
#pragma omp parallel
    {
#pragma omp single nowait
        plan_ok = make_plan(n_walk, n_epi, n_epj, n_spj, c->par.exec_mode, c->par.chunk_j, c->par.max_ipl, P, 1,
                            c->mono_sent);
#pragma omp for schedule(dynamic) nowait
        for (int w = 0; w < n_walk; w++) {
            const int ne = n_epj[w], ns = n_spj[w];
            bool ok = true;
            if (encode) {
                int mn, mx;
                c->run_ne[w] = encode_runs(id_epj[w], ne, hrid + oe[w], hrpos + oe[w], mn, mx);
                ok = mn >= 0 && mx < nj_lim;
                c->run_ns[w] = encode_runs(id_spj[w], ns, hrid + n_ide + os[w], hrpos + n_ide + os[w], mn, mx);
                ok = ok && mn >= 0 && mx < ns_lim;
            } else {
                for (int k = 0; k < ne; k++) ok = ok && id_epj[w][k] >= 0 && id_epj[w][k] < nj_lim;
                for (int k = 0; k < ns; k++) ok = ok && id_spj[w][k] >= 0 && id_spj[w][k] < ns_lim;
            }
            if (!ok) {
#pragma omp atomic write
                bad = w;
            }
            const char *base = static_cast<const char *>(epi[w]);
            for (int i = 0; i < n_epi[w]; i++) {
                const char *r = base + (size_t)i * iv->stride;
                const double *q = reinterpret_cast<const double *>(r + iv->off_pos);
                EpiDev d;
                d.x = q[0];
                d.y = q[1];
                d.z = q[2];
                d.rs = (float)*reinterpret_cast<const double *>(r + iv->off_rsearch);
                d.pad = 0;
                he[oi[w] + i] = d;
                if (co) hid[oi[w] + i] = *reinterpret_cast<const long long *>(r + iv->off_id);
            }
        }
    }
    if (!plan_ok) return c->fail(PTH_ERR_CAPACITY, "pth_dispatch: batch exceeds the int32 index range");
    if (co) {
        const char *m = co_plan_error(P);
        if (m) return c->fail(PTH_ERR_CAPACITY, std::string("pth_dispatch: ") + m);
    }
    if (bad >= 0)
        return c->fail(PTH_ERR_ARG, "pth_dispatch: list index out of range in walk " + std::to_string(bad));
    long long n_re = 0, n_rs = 0;
    for (int w = 0; w < n_walk; w++) {
        n_re += c->run_ne[w];
        n_rs += c->run_ns[w];
    }
    // runs travel when they take at most half the bytes of the lists
    const bool runs = c->list_runs == 1 ||
                      (c->list_runs < 0 && 8 * (n_re + n_rs + n_walk + 1) <= 2 * (n_ide + n_ids));
    c->prof.copy_s += wall_s() - t_pack;
    Arena A;
    if (runs) A.layout(P, n_ide, n_ids, co, n_re, n_rs);
    else A.layout(P, n_ide, n_ids, co);
    HIPCHK(c, c->h_arena.ensure_keep(A.total, A.walks), "pinned arena allocation");
    const bool grow = A.total > c->d_arena.cap;
    if (grow) HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    HIPCHK(c, c->d_arena.ensure(A.total), "device arena allocation");
    {
        // the work buffers are shared by all tags: reallocation waits for the stream
        bool need = false;
        need |= sizeof(double4) * P.walks.size() > c->work.centre.cap;
        need |= sizeof(float4) * (size_t)P.ni_tot > c->work.staged_i.cap;
        need |= sizeof(float) * 5 * (size_t)P.n_e_slots > c->work.staged_e.cap;
        need |= sizeof(float) * PTH_SPJ_STAGED * (size_t)P.n_s_rec > c->work.staged_s.cap;
        need |= sizeof(double4) * 64 * (size_t)P.n_part > c->work.part_accp.cap;
        if (co) {
            need |= sizeof(float4) * (size_t)P.ni_tot > c->work.staged_ic.cap;
            need |= sizeof(float4) * (size_t)P.ni_tot > c->work.staged_il.cap;
            need |= sizeof(float) * (PTH_CO_EXTRA / 8) * (size_t)P.n_e_slots > c->work.staged_ec.cap;
            need |= sizeof(double) * 64 * (size_t)P.n_part > c->work.part_pott.cap;
            need |= sizeof(int) * P.walks.size() > c->work.walk_err.cap;
        }
        const bool df = df_enabled(c->par) != 0;
        if (df) {
            need |= sizeof(float4) * (size_t)P.ni_tot > c->work.staged_il.cap;
            need |= sizeof(float) * (PTH_DF_EXTRA / 8) * (size_t)P.n_e_slots > c->work.staged_el.cap;
        }
        if (runs) need |= sizeof(int) * (size_t)(n_ide + n_ids) > c->d_ids.cap;
        if (need) HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        HIPCHK(c, c->work.ensure(P, co, df), "work buffer allocation");
        if (runs) HIPCHK(c, c->d_ids.ensure(sizeof(int) * (size_t)std::max<long long>(n_ide + n_ids, 1)),
                         "device list allocation");
    }
    // the tag's pinned result buffers (nothing of this tag is in flight);
    // with zero_copy the kernels write the results straight into them (host
    // memory the device addresses; the writes are visible to the host once
    // T.done, recorded after the kernels without hipEventDisableSystemFence,
    // has completed), else into device buffers then copied
    HIPCHK(c, T.h_accp.ensure(sizeof(double4) * std::max<long long>(P.ni_tot, 1)), "pinned output");
    HIPCHK(c, T.h_nnb.ensure(sizeof(int) * std::max<long long>(P.ni_tot, 1)), "pinned output");
    if (co) {
        HIPCHK(c, T.h_pott.ensure(sizeof(double) * std::max<long long>(P.ni_tot, 1)), "pinned output");
        HIPCHK(c, T.h_err.ensure(sizeof(int) * std::max<size_t>(P.walks.size(), 1)), "pinned output");
    }
    const bool zc = c->zero_copy;
    if (!zc) {
        HIPCHK(c, T.out_accp.ensure(sizeof(double4) * std::max<long long>(P.ni_tot, 1)), "output allocation");
        HIPCHK(c, T.out_nnb.ensure(sizeof(int) * std::max<long long>(P.ni_tot, 1)), "output allocation");
        if (co) HIPCHK(c, T.out_pott.ensure(sizeof(double) * std::max<long long>(P.ni_tot, 1)), "output allocation");
    }

    t_pack = wall_s();
    char *h = c->h_arena.as<char>();
    int *hide = reinterpret_cast<int *>(h + A.ide);
    int *hids = reinterpret_cast<int *>(h + A.ids);
    // runs: first indices of all of the batch's runs, then their positions
    int *hre = reinterpret_cast<int *>(h + A.run_e);
    int *hrs = reinterpret_cast<int *>(h + A.run_s);
    int *hreo = reinterpret_cast<int *>(h + A.run_eo);
    int *hrso = reinterpret_cast<int *>(h + A.run_so);
    if (runs) {
        // run offsets: walk w's runs start at the count of the walks before it
        int oe = 0, os = 0;
        for (int w = 0; w < n_walk; w++) {
            const int ne = c->run_ne[w], ns = c->run_ns[w];
            c->run_ne[w] = oe;
            c->run_ns[w] = os;
            oe += ne;
            os += ns;
        }
        c->run_ne[n_walk] = oe;
        c->run_ns[n_walk] = os;
        std::memcpy(hreo, c->run_ne.data(), sizeof(int) * (n_walk + 1));
        std::memcpy(hrso, c->run_ns.data(), sizeof(int) * (n_walk + 1));
    }
#pragma omp parallel for schedule(dynamic)
    for (int w = 0; w < n_walk; w++) {
        const WalkDev &W = P.walks[w];
        if (runs) {
            const int ne = c->run_ne[w + 1] - c->run_ne[w], ns = c->run_ns[w + 1] - c->run_ns[w];
            if (ne) {
                std::memcpy(hre + c->run_ne[w], hrid + W.e_off, sizeof(int) * ne);
                std::memcpy(hre + n_re + c->run_ne[w], hrpos + W.e_off, sizeof(int) * ne);
            }
            if (ns) {
                std::memcpy(hrs + c->run_ns[w], hrid + n_ide + W.s_off, sizeof(int) * ns);
                std::memcpy(hrs + n_rs + c->run_ns[w], hrpos + n_ide + W.s_off, sizeof(int) * ns);
            }
        } else {
            if (W.n_e) std::memcpy(hide + W.e_off, id_epj[w], sizeof(int) * W.n_e);
            if (W.n_s) std::memcpy(hids + W.s_off, id_spj[w], sizeof(int) * W.n_s);
        }
    }
    copy_plan_to_arena(h, P, A);
    c->prof.copy_s += wall_s() - t_pack;
    c->prof.n_walk += n_walk;
    c->prof.n_epi += P.ni_tot;
    c->prof.n_epj += n_ide;
    c->prof.n_spj += n_ids;
    c->prof.n_call += 1;
    T.timed = c->prof_on;
    if (T.timed) HIPCHK(c, hipEventRecord(T.prof[0], c->stream), "hipEventRecord");
    HIPCHK(c, hipMemcpyAsync(c->d_arena.p, h, A.total, hipMemcpyHostToDevice, c->stream), "batch upload");
    HIPCHK(c, hipEventRecord(c->ev_upload, c->stream), "hipEventRecord");
    if (T.timed) HIPCHK(c, hipEventRecord(T.prof[1], c->stream), "hipEventRecord");
    if (c->par.exec_mode == PTH_EXEC_PLAIN) HIPCHK(c, hipStreamSynchronize(c->stream), "upload sync");
    BatchDev b;
    // the SPJ table of the sent superparticles (quadrupole form), for this
    // batch's eps2
    if (!c->mono_sent && c->n_spj > 0 && (!c->tab_valid || c->tab_eps2 != (float)c->par.eps2)) {
        if (sizeof(SpjTab) * (size_t)c->n_spj > c->d_spj_tab.cap)
            HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        HIPCHK(c, c->d_spj_tab.ensure(sizeof(SpjTab) * (size_t)c->n_spj), "SPJ table allocation");
        HIPCHK(c, launch_spj_table(c->d_spj.as<SpjDev>(), c->n_spj, (float)c->par.eps2, c->d_spj_tab.as<SpjTab>(),
                                   c->stream), "SPJ table kernel");
        c->tab_valid = true;
        c->tab_eps2 = (float)c->par.eps2;
    }
    fill_batch(b, P, c->par, c->d_arena, A, c->work, c->d_epj.as<EpjDev>(), c->d_spj.as<SpjDev>(),
               zc ? T.h_accp.as<double4>() : T.out_accp.as<double4>(), zc ? T.h_nnb.as<int>() : T.out_nnb.as<int>(),
               c->d_epj_co.as<EpjCoDev>(), zc ? T.h_pott.as<double>() : T.out_pott.as<double>(),
               c->d_spj_tab.as<SpjTab>());
    if (runs) {
        char *d = c->d_arena.as<char>();
        int *ide_dev = c->d_ids.as<int>();
        const int *re = reinterpret_cast<const int *>(d + A.run_e);
        const int *rs = reinterpret_cast<const int *>(d + A.run_s);
        HIPCHK(c, launch_expand_runs(b.walks, 0, n_walk, re, re + n_re, reinterpret_cast<const int *>(d + A.run_eo),
                                     rs, rs + n_rs, reinterpret_cast<const int *>(d + A.run_so), ide_dev,
                                     ide_dev + n_ide, c->stream),
               "list expansion kernel");
        b.id_epj = ide_dev;
        b.id_spj = ide_dev + n_ide;
    }
    // a batch of up to a few hundred walks holds fewer waves of each
    // i-per-lane class than the chip has slots: each class launch is one
    // latency-bound round, so the 1-i items run on a second stream beside
    // the longer classes instead of after them
    const bool split = c->class_streams && !b.plain && P.seg_count[0] > 0 &&
                       P.seg_count[1] + P.seg_count[2] + P.seg_count[3] > 0;
    if (split) {
        if (!c->side) HIPCHK(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking), "hipStreamCreate");
        if (!c->ev_fork) HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "hipEventCreate");
        if (!c->ev_join) HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "hipEventCreate");
        HIPCHK(c, launch_stage(b, 0, c->stream), "staging kernel");
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream), "hipEventRecord");
        HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0), "hipStreamWaitEvent");
        HIPCHK(c, launch_force(b, 0, c->stream, 0xEu), "soft-force kernels");
        HIPCHK(c, launch_force(b, 0, c->side, 0x1u), "soft-force kernels");
        HIPCHK(c, hipEventRecord(c->ev_join, c->side), "hipEventRecord");
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0), "hipStreamWaitEvent");
        HIPCHK(c, launch_reduce(b, 0, c->stream), "reduction kernel");
    } else {
        HIPCHK(c, launch_all(b, c->stream), "soft-force kernels");
    }
    if (T.timed) HIPCHK(c, hipEventRecord(T.prof[2], c->stream), "hipEventRecord");
    if (!zc && P.ni_tot > 0) {
        HIPCHK(c, hipMemcpyAsync(T.h_accp.p, T.out_accp.p, sizeof(double4) * P.ni_tot, hipMemcpyDeviceToHost,
                                 c->stream), "result download");
        HIPCHK(c, hipMemcpyAsync(T.h_nnb.p, T.out_nnb.p, sizeof(int) * P.ni_tot, hipMemcpyDeviceToHost,
                                 c->stream), "result download");
        if (co)
            HIPCHK(c, hipMemcpyAsync(T.h_pott.p, T.out_pott.p, sizeof(double) * P.ni_tot, hipMemcpyDeviceToHost,
                                     c->stream), "result download");
    }
    if (co && n_walk > 0)
        HIPCHK(c, hipMemcpyAsync(T.h_err.p, c->work.walk_err.p, sizeof(int) * n_walk, hipMemcpyDeviceToHost,
                                 c->stream), "result download");
    T.co = co;
    if (T.timed) HIPCHK(c, hipEventRecord(T.prof[3], c->stream), "hipEventRecord");
    HIPCHK(c, hipEventRecord(T.done, c->stream), "hipEventRecord");
    T.n_i.assign(n_epi, n_epi + n_walk);
    T.ni_tot = P.ni_tot;
    T.pending = true;
    return PTH_OK;
}

// ---- the non-index multi-walk functor: src/force_gpu_cuda.cu:704-816 ----------
// The batch's own EPJ / SPJ records per walk become the j arrays of this
// batch (packed and sent as pth_send_j sends a force call's), and the batch
// is dispatched with each walk's list naming its own records in order (one
// run of consecutive indices per walk).
int pth_dispatch_copies(pth_ctx *c, int tag, int n_walk, const void *const *epi, const int *n_epi,
                        const pth_epi_view *iv, const void *const *epj, const int *n_epj, const pth_epj_view *ev,
                        const void *const *spj, const int *n_spj, const pth_spj_view *sv)
{
    if (!c) return PTH_ERR_ARG;
    if (tag < 0 || tag >= MAX_TAG) return c->fail(PTH_ERR_ARG, "pth_dispatch_copies: tag out of range [0,8)");
    if (n_walk < 0 || n_walk > c->par.max_walk)
        return c->fail(PTH_ERR_CAPACITY, "pth_dispatch_copies: n_walk " + std::to_string(n_walk) +
                                             " outside [0, " + std::to_string(c->par.max_walk) + "]");
    if (n_walk > 0 && (!n_epi || !n_epj || !n_spj || !epj || !spj || !ev || !sv))
        return c->fail(PTH_ERR_ARG, "pth_dispatch_copies: null argument");
    if (c->tag[tag].pending)
        return c->fail(PTH_ERR_STATE, "pth_dispatch_copies: previous batch of this tag not retrieved");
    std::vector<long long> oe(n_walk + 1, 0), os(n_walk + 1, 0);
    for (int w = 0; w < n_walk; w++) {
        if (n_epj[w] < 0 || n_spj[w] < 0)
            return c->fail(PTH_ERR_ARG, "pth_dispatch_copies: negative size in walk " + std::to_string(w));
        if ((n_epj[w] > 0 && !epj[w]) || (n_spj[w] > 0 && !spj[w]))
            return c->fail(PTH_ERR_ARG, "pth_dispatch_copies: null records in walk " + std::to_string(w));
        oe[w + 1] = oe[w] + n_epj[w];
        os[w + 1] = os[w] + n_spj[w];
    }
    if (oe[n_walk] > INT_MAX || os[n_walk] > INT_MAX)
        return c->fail(PTH_ERR_CAPACITY, "pth_dispatch_copies: batch exceeds the int32 index range");
    const int ne = (int)oe[n_walk], ns = (int)os[n_walk];
    const int rc = send_records(
        c, ne, ev, ns, sv, n_walk,
        [&](EpjDev *he, EpjCoDev *hc, const pth_epj_view *v) {
#pragma omp parallel for schedule(dynamic)
            for (int w = 0; w < n_walk; w++) {
                const char *b = static_cast<const char *>(epj[w]);
                for (int j = 0; j < n_epj[w]; j++) {
                    const char *r = b + (size_t)j * v->stride;
                    he[oe[w] + j] = epj_record(r, v);
                    if (hc) hc[oe[w] + j] = epj_co_record(r, v);
                }
            }
        },
        [&](EpjDev *hm, SpjDev *hs, const pth_spj_view *v) {
#pragma omp parallel for schedule(dynamic)
            for (int w = 0; w < n_walk; w++) {
                const char *b = static_cast<const char *>(spj[w]);
                for (int j = 0; j < n_spj[w]; j++) {
                    const char *r = b + (size_t)j * v->stride;
                    if (hm) hm[os[w] + j] = spj_mono_record(r, v);
                    else hs[os[w] + j] = spj_record(r, v);
                }
            }
        });
    if (rc != PTH_OK) return rc;
    const size_t n_id = (size_t)std::max(ne, ns);
    for (size_t k = c->ident.size(); k < n_id; k++) c->ident.push_back((int)k);
    std::vector<const int *> le(n_walk), ls(n_walk);
    for (int w = 0; w < n_walk; w++) {
        le[w] = c->ident.data() + oe[w];
        ls[w] = c->ident.data() + os[w];
    }
    return pth_dispatch(c, tag, n_walk, epi, n_epi, iv, le.data(), n_epj, ls.data(), n_spj);
}

// ---- RetrieveForceCUDA: src/force_gpu_cuda.cu:831-880 -------------------------
int pth_retrieve(pth_ctx *c, int tag, int n_walk, const int *ni, void *const *force,
                 const pth_force_view *fv)
{
    if (!c) return PTH_ERR_ARG;
    if (tag < 0 || tag >= MAX_TAG) return c->fail(PTH_ERR_ARG, "pth_retrieve: tag out of range [0,8)");
    TagState &T = c->tag[tag];
    if (!T.pending) return c->fail(PTH_ERR_STATE, "pth_retrieve: nothing dispatched for this tag");
    if (n_walk != (int)T.n_i.size())
        return c->fail(PTH_ERR_ARG, "pth_retrieve: n_walk differs from the dispatched batch");
    for (int w = 0; w < n_walk; w++)
        if (!ni || ni[w] != T.n_i[w])
            return c->fail(PTH_ERR_ARG, "pth_retrieve: ni differs from the dispatched batch in walk " +
                                            std::to_string(w));
    if (n_walk > 0 && (!force || !fv)) return c->fail(PTH_ERR_ARG, "pth_retrieve: null force records");
    if (fv && fv->nngb_bytes != 4 && fv->nngb_bytes != 8)
        return c->fail(PTH_ERR_ARG, "pth_retrieve: nngb_bytes must be 4 or 8");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    T.pending = false;
    HIPCHK(c, hipEventSynchronize(T.done), "hipEventSynchronize");
    HIPCHK(c, hipGetLastError(), "kernel execution");
    if (T.timed) {
        float ms[3] = {0.f, 0.f, 0.f};
        for (int k = 0; k < 3; k++)
            HIPCHK(c, hipEventElapsedTime(&ms[k], T.prof[k], T.prof[k + 1]), "hipEventElapsedTime");
        c->prof.send_s += 1e-3 * ms[0];
        c->prof.calc_s += 1e-3 * ms[1];
        c->prof.recv_s += 1e-3 * ms[2];
        T.timed = false;
    }
    if (c->prof_send_pending) {
        // the send's copies ran on this stream before the batch: complete
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->prof_send[0], c->prof_send[1]), "hipEventElapsedTime");
        c->prof.send_s += 1e-3 * ms;
        c->prof_send_pending = false;
    }
    const double t_scatter = wall_s();
    const double4 *a = T.h_accp.as<double4>();
    const int *nb = T.h_nnb.as<int>();
    const double *pt = T.h_pott.as<double>();
    if (T.co)
        for (int w = 0; w < n_walk; w++)
            if (T.h_err.as<int>()[w] != 0)
                return c->fail(PTH_ERR_ARG, "pth_retrieve: changeover mode: an i-particle of walk " +
                                                std::to_string(w) + " has no EPJ with its id in the walk's list");
    const bool write_pt = T.co && fv && fv->off_pot_tot >= 0;
    std::vector<long long> off(n_walk + 1, 0);
    for (int w = 0; w < n_walk; w++) off[w + 1] = off[w] + ni[w];
#pragma omp parallel for schedule(dynamic)
    for (int w = 0; w < n_walk; w++) {
        char *base = static_cast<char *>(force[w]);
        for (int i = 0; i < ni[w]; i++) {
            char *r = base + (size_t)i * fv->stride;
            const double4 v = a[off[w] + i];
            double *acc = reinterpret_cast<double *>(r + fv->off_acc);
            acc[0] = v.x;
            acc[1] = v.y;
            acc[2] = v.z;
            *reinterpret_cast<double *>(r + fv->off_pot) = v.w;
            if (write_pt) *reinterpret_cast<double *>(r + fv->off_pot_tot) = pt[off[w] + i];
            if (fv->nngb_bytes == 8) *reinterpret_cast<long long *>(r + fv->off_nngb) = nb[off[w] + i];
            else *reinterpret_cast<int *>(r + fv->off_nngb) = nb[off[w] + i];
        }
    }
    c->prof.copy_s += wall_s() - t_scatter;
    return PTH_OK;
}

int pth_profile_enable(pth_ctx *c, int on)
{
    if (!c) return PTH_ERR_ARG;
    if (on) {
        HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
        for (int t = 0; t < MAX_TAG; t++)
            for (hipEvent_t &x : c->tag[t].prof)
                if (!x) HIPCHK(c, hipEventCreate(&x), "hipEventCreate");
        for (hipEvent_t &x : c->prof_send)
            if (!x) HIPCHK(c, hipEventCreate(&x), "hipEventCreate");
    }
    c->prof_on = on != 0;
    return PTH_OK;
}

int pth_profile_take(pth_ctx *c, pth_profile *out)
{
    if (!c || !out) return PTH_ERR_ARG;
    *out = c->prof;
    c->prof = pth_profile{};
    return PTH_OK;
}

// ---- device-resident evaluation ------------------------------------------------
int pth_resident_create(pth_ctx *c, pth_resident **out, int n_walk, const int *n_epi,
                        const int *n_epj, const int *n_spj, const double *epi_pos,
                        const double *epi_rs, const int *id_epj, const int *id_spj,
                        int n_epj_tot, const double *epj_pos, const double *epj_mass,
                        const double *epj_rs, int n_spj_tot, const double *spj_pos,
                        const double *spj_mass, const double *spj_quad)
{
    if (!c) return PTH_ERR_ARG;
    if (!out) return c->fail(PTH_ERR_ARG, "pth_resident_create: null output");
    *out = nullptr;
    if (n_walk < 0 || n_epj_tot < 0 || n_spj_tot < 0)
        return c->fail(PTH_ERR_ARG, "pth_resident_create: negative count");
    if (n_spj_tot > 0 && (!spj_pos || !spj_mass || (!spj_quad && !c->par.spj_monopole)))
        return c->fail(PTH_ERR_ARG, "pth_resident_create: null SPJ array");
    long long n_ide = 0, n_ids = 0;
    for (int w = 0; w < n_walk; w++) {
        if (n_epi[w] < 0 || n_epj[w] < 0 || n_spj[w] < 0)
            return c->fail(PTH_ERR_ARG, "pth_resident_create: negative size in walk " + std::to_string(w));
        n_ide += n_epj[w];
        n_ids += n_spj[w];
    }
    for (long long k = 0; k < n_ide; k++)
        if (id_epj[k] < 0 || id_epj[k] >= n_epj_tot)
            return c->fail(PTH_ERR_ARG, "pth_resident_create: EPJ index out of range");
    for (long long k = 0; k < n_ids; k++)
        if (id_spj[k] < 0 || id_spj[k] >= n_spj_tot)
            return c->fail(PTH_ERR_ARG, "pth_resident_create: SPJ index out of range");
    pth_resident *r = nullptr;
    int rc = resident_prepare(c, r, n_walk, n_epi, n_epj, n_spj, n_ide, n_ids, n_epj_tot, n_spj_tot,
                              "pth_resident_create");
    if (rc != PTH_OK) return rc;
    const Plan &P = r->plan;
    auto bail = [&](hipError_t e, const char *what) {
        int rc2 = c->hip_fail(e, what);
        pth_resident_free(r);
        return rc2;
    };
    hipError_t e;
    // host mirrors (pageable is fine: uploaded once)
    std::vector<char> h(r->A.total);
    EpiDev *he = reinterpret_cast<EpiDev *>(h.data() + r->A.epi);
    for (long long i = 0; i < P.ni_tot; i++) {
        he[i].x = epi_pos[3 * i];
        he[i].y = epi_pos[3 * i + 1];
        he[i].z = epi_pos[3 * i + 2];
        he[i].rs = (float)epi_rs[i];
        he[i].pad = 0;
    }
    if (n_ide) std::memcpy(h.data() + r->A.ide, id_epj, sizeof(int) * n_ide);
    if (n_ids) std::memcpy(h.data() + r->A.ids, id_spj, sizeof(int) * n_ids);
    copy_plan_to_arena(h.data(), P, r->A);
    std::vector<EpjDev> ej(std::max(n_epj_tot, 0));
    for (int i = 0; i < n_epj_tot; i++) {
        ej[i].x = epj_pos[3 * (size_t)i];
        ej[i].y = epj_pos[3 * (size_t)i + 1];
        ej[i].z = epj_pos[3 * (size_t)i + 2];
        ej[i].m = (float)epj_mass[i];
        ej[i].rs = (float)epj_rs[i];
    }
    std::vector<SpjDev> sj(P.mono ? 0 : std::max(n_spj_tot, 0));
    std::vector<EpjDev> sm(P.mono ? std::max(n_spj_tot, 0) : 0);
    for (int i = 0; i < (P.mono ? n_spj_tot : 0); i++) {
        sm[i].x = spj_pos[3 * (size_t)i];
        sm[i].y = spj_pos[3 * (size_t)i + 1];
        sm[i].z = spj_pos[3 * (size_t)i + 2];
        sm[i].m = (float)spj_mass[i];
        sm[i].rs = 0.f;
    }
    for (int i = 0; i < (P.mono ? 0 : n_spj_tot); i++) {
        const double *q = spj_quad + 6 * (size_t)i;
        sj[i].x = spj_pos[3 * (size_t)i];
        sj[i].y = spj_pos[3 * (size_t)i + 1];
        sj[i].z = spj_pos[3 * (size_t)i + 2];
        sj[i].m = (float)spj_mass[i];
        sj[i].qxx = (float)q[0];
        sj[i].qyy = (float)q[1];
        sj[i].qzz = (float)q[2];
        sj[i].qxy = (float)q[3];
        sj[i].qxz = (float)q[4];
        sj[i].qyz = (float)q[5];
        sj[i].pad[0] = sj[i].pad[1] = sj[i].pad[2] = 0.f;
    }
    if ((e = hipMemcpy(r->arena.p, h.data(), r->A.total, hipMemcpyHostToDevice))) return bail(e, "resident upload");
    if (n_epj_tot && (e = hipMemcpy(r->epj.p, ej.data(), sizeof(EpjDev) * n_epj_tot, hipMemcpyHostToDevice)))
        return bail(e, "resident upload");
    if (n_spj_tot && (e = P.mono ? hipMemcpy(r->spj.p, sm.data(), sizeof(EpjDev) * n_spj_tot, hipMemcpyHostToDevice)
                                 : hipMemcpy(r->spj.p, sj.data(), sizeof(SpjDev) * n_spj_tot, hipMemcpyHostToDevice)))
        return bail(e, "resident upload");
    if ((e = resident_finish(r))) return bail(e, "resident SPJ table");
    *out = r;
    return PTH_OK;
}

int pth_resident_changeover(pth_resident *r, const long long *epi_id, const long long *epj_id,
                            const double *epj_rin, const double *epj_rout, const double *epj_group)
{
    if (!r) return PTH_ERR_ARG;
    pth_ctx *c = r->ctx;
    if (!r->co) return c->fail(PTH_ERR_STATE, "pth_resident_changeover: the context is not in changeover mode");
    const long long ni = r->plan.ni_tot;
    if ((ni > 0 && !epi_id) || (r->n_epj_tot > 0 && (!epj_id || !epj_rin || !epj_rout)))
        return c->fail(PTH_ERR_ARG, "pth_resident_changeover: null array");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    std::vector<EpjCoDev> co(std::max(r->n_epj_tot, 1));
    for (int i = 0; i < r->n_epj_tot; i++)
        co[i] = co_record(epj_rin[i], epj_rout[i], epj_id[i], epj_group ? epj_group + 2 * (size_t)i : nullptr);
    if (r->n_epj_tot > 0)
        HIPCHK(c, hipMemcpy(r->epj_co.p, co.data(), sizeof(EpjCoDev) * r->n_epj_tot, hipMemcpyHostToDevice),
               "resident upload");
    if (ni > 0)
        HIPCHK(c, hipMemcpy(r->arena.as<char>() + r->A.epi_id, epi_id, sizeof(long long) * ni,
                            hipMemcpyHostToDevice), "resident upload");
    r->co_ready = true;
    return PTH_OK;
}

// The launches of one evaluation on stream st: on st alone, or pipelined
// over the side streams (forked from st by fork_ev, joined back into st).
// With timed (st the context's stream), events bracket the stage, force and
// reduce phases of the one-stream form.
static int resident_enqueue(pth_resident *r, hipStream_t st, hipEvent_t fork_ev, bool timed)
{
    pth_ctx *c = r->ctx;
    if (r->side.empty() || r->b.n_seg < 2 || (int)r->staged.size() < r->b.n_seg) {
        // one segment: its 1-i items run on a side stream beside the longer
        // classes (the 1-i launch is about one round of waves, whose run-down
        // tail the longer launch then fills)
        const SegDev &g0 = r->b.seg[0];
        const bool split = r->b.n_seg == 1 && c->class_streams && !r->b.plain && !r->side.empty() &&
                           !r->staged.empty() && g0.seg_count[0] > 0 &&
                           g0.seg_count[1] + g0.seg_count[2] + g0.seg_count[3] > 0;
        for (int k = 0; k < r->b.n_seg; k++) HIPCHK(c, launch_stage(r->b, k, st), "stage kernels");
        if (timed) HIPCHK(c, hipEventRecord(r->ev[1], st), "hipEventRecord");
        if (split) {
            HIPCHK(c, hipEventRecord(r->staged[0], st), "hipEventRecord");
            HIPCHK(c, hipStreamWaitEvent(r->side[0], r->staged[0], 0), "hipStreamWaitEvent");
            HIPCHK(c, launch_force(r->b, 0, st, 0xEu), "force kernel");
            HIPCHK(c, launch_force(r->b, 0, r->side[0], 0x1u), "force kernel");
            HIPCHK(c, hipEventRecord(r->side_done[0], r->side[0]), "hipEventRecord");
            HIPCHK(c, hipStreamWaitEvent(st, r->side_done[0], 0), "hipStreamWaitEvent");
        } else {
            for (int k = 0; k < r->b.n_seg; k++) HIPCHK(c, launch_force(r->b, k, st), "force kernel");
        }
        if (timed) HIPCHK(c, hipEventRecord(r->ev[2], st), "hipEventRecord");
        for (int k = 0; k < r->b.n_seg; k++) HIPCHK(c, launch_reduce(r->b, k, st), "reduce kernel");
        r->pipelined = false;
        return PTH_OK;
    }
    // a pipeline: the first side stream stages the segments one after
    // another; segment k's force and reduce run on side stream
    // 1 + k mod (n - 1) once its staging is done, so the staging of the
    // segments after k overlaps k's force kernels
    const int ns = (int)r->side.size();
    if (!timed) HIPCHK(c, hipEventRecord(fork_ev, st), "hipEventRecord");
    for (int s = 0; s < ns; s++) HIPCHK(c, hipStreamWaitEvent(r->side[s], fork_ev, 0), "hipStreamWaitEvent");
    for (int k = 0; k < r->b.n_seg; k++) {
        HIPCHK(c, launch_stage(r->b, k, r->side[0]), "stage kernels");
        HIPCHK(c, hipEventRecord(r->staged[k], r->side[0]), "hipEventRecord");
    }
    for (int k = 0; k < r->b.n_seg; k++) {
        hipStream_t fs = r->side[ns > 1 ? 1 + k % (ns - 1) : 0];
        HIPCHK(c, hipStreamWaitEvent(fs, r->staged[k], 0), "hipStreamWaitEvent");
        // the last segment's 1-i launch on the staging stream (idle by then)
        // beside its longer classes: nothing else is left to fill its tail
        const SegDev &g = r->b.seg[k];
        const bool split = k == r->b.n_seg - 1 && ns > 1 && c->class_streams && !r->b.plain &&
                           g.seg_count[0] > 0 && g.seg_count[1] + g.seg_count[2] + g.seg_count[3] > 0;
        if (split) {
            HIPCHK(c, launch_force(r->b, k, fs, 0xEu), "force kernel");
            HIPCHK(c, launch_force(r->b, k, r->side[0], 0x1u), "force kernel");
            HIPCHK(c, hipEventRecord(r->staged[k], r->side[0]), "hipEventRecord");
            HIPCHK(c, hipStreamWaitEvent(fs, r->staged[k], 0), "hipStreamWaitEvent");
        } else {
            HIPCHK(c, launch_force(r->b, k, fs), "force kernel");
        }
        HIPCHK(c, launch_reduce(r->b, k, fs), "reduce kernel");
    }
    for (int s = 0; s < ns; s++) {
        HIPCHK(c, hipEventRecord(r->side_done[s], r->side[s]), "hipEventRecord");
        HIPCHK(c, hipStreamWaitEvent(st, r->side_done[s], 0), "hipStreamWaitEvent");
    }
    r->pipelined = true;
    return PTH_OK;
}

// Graph mode: records the evaluation's launches on the capture stream and
// instantiates them (again after any rebuild of the launch descriptor).
static int resident_capture(pth_resident *r)
{
    pth_ctx *c = r->ctx;
    if (r->gexec && r->g_gen == r->b_gen) return PTH_OK;
    if (r->gexec) {
        HIPCHK(c, hipGraphExecDestroy(r->gexec), "hipGraphExecDestroy");
        r->gexec = nullptr;
    }
    if (!r->cap) HIPCHK(c, hipStreamCreateWithFlags(&r->cap, hipStreamNonBlocking), "hipStreamCreate");
    if (!r->fork) HIPCHK(c, hipEventCreateWithFlags(&r->fork, hipEventDisableTiming), "hipEventCreate");
    // the capture stream starts after the context's stream's work
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    HIPCHK(c, hipStreamBeginCapture(r->cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const int rc = resident_enqueue(r, r->cap, r->fork, false);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(r->cap, &graph);
    if (rc != PTH_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return c->hip_fail(e, "hipStreamEndCapture");
    }
    const hipError_t ei = hipGraphInstantiate(&r->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        r->gexec = nullptr;
        return c->hip_fail(ei, "hipGraphInstantiate");
    }
    r->g_gen = r->b_gen;
    return PTH_OK;
}

int pth_resident_run(pth_resident *r)
{
    if (!r) return PTH_ERR_ARG;
    pth_ctx *c = r->ctx;
    if (r->co && !r->co_ready)
        return c->fail(PTH_ERR_STATE, "pth_resident_run: changeover mode needs pth_resident_changeover first");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    if (c->par.graph) {
        const int rc = resident_capture(r);
        if (rc != PTH_OK) return rc;
        HIPCHK(c, hipEventRecord(r->ev[0], c->stream), "hipEventRecord");
        HIPCHK(c, hipGraphLaunch(r->gexec, c->stream), "hipGraphLaunch");
        HIPCHK(c, hipEventRecord(r->ev[3], c->stream), "hipEventRecord");
        r->graph_run = true;
        r->ran = true;
        return PTH_OK;
    }
    HIPCHK(c, hipEventRecord(r->ev[0], c->stream), "hipEventRecord");
    const int rc = resident_enqueue(r, c->stream, r->ev[0], true);
    if (rc != PTH_OK) return rc;
    HIPCHK(c, hipEventRecord(r->ev[3], c->stream), "hipEventRecord");
    r->graph_run = false;
    r->ran = true;
    return PTH_OK;
}

int pth_resident_sync(pth_resident *r)
{
    if (!r) return PTH_ERR_ARG;
    pth_ctx *c = r->ctx;
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    HIPCHK(c, hipGetLastError(), "kernel execution");
    return PTH_OK;
}

int pth_resident_results(pth_resident *r, double *acc, double *pot, long long *nnb, double *pot_tot)
{
    if (!r) return PTH_ERR_ARG;
    pth_ctx *c = r->ctx;
    if (!r->ran) return c->fail(PTH_ERR_STATE, "pth_resident_results: no evaluation run");
    int rc = pth_resident_sync(r);
    if (rc) return rc;
    const long long n = r->plan.ni_tot;
    std::vector<double4> a(std::max<long long>(n, 1));
    std::vector<int> nb(std::max<long long>(n, 1));
    if (n > 0) {
        HIPCHK(c, hipMemcpy(a.data(), r->out_accp.p, sizeof(double4) * n, hipMemcpyDeviceToHost), "result download");
        HIPCHK(c, hipMemcpy(nb.data(), r->out_nnb.p, sizeof(int) * n, hipMemcpyDeviceToHost), "result download");
    }
    if (r->co) {
        const int nw = (int)r->plan.walks.size();
        std::vector<int> err(std::max(nw, 1));
        if (nw > 0)
            HIPCHK(c, hipMemcpy(err.data(), r->work.walk_err.p, sizeof(int) * nw, hipMemcpyDeviceToHost),
                   "result download");
        for (int w = 0; w < nw; w++)
            if (err[w])
                return c->fail(PTH_ERR_ARG, "pth_resident_results: changeover mode: an i-particle of walk " +
                                                std::to_string(w) + " has no EPJ with its id in the walk's list");
        if (pot_tot && n > 0)
            HIPCHK(c, hipMemcpy(pot_tot, r->out_pott.p, sizeof(double) * n, hipMemcpyDeviceToHost),
                   "result download");
    }
    for (long long i = 0; i < n; i++) {
        if (acc) {
            acc[3 * i] = a[i].x;
            acc[3 * i + 1] = a[i].y;
            acc[3 * i + 2] = a[i].z;
        }
        if (pot) pot[i] = a[i].w;
        if (nnb) nnb[i] = nb[i];
    }
    return PTH_OK;
}

int pth_resident_timing(pth_resident *r, float *ms_stage, float *ms_force, float *ms_reduce)
{
    if (!r) return PTH_ERR_ARG;
    pth_ctx *c = r->ctx;
    if (!r->ran) return c->fail(PTH_ERR_STATE, "pth_resident_timing: no evaluation run");
    HIPCHK(c, hipEventSynchronize(r->ev[3]), "hipEventSynchronize");
    float t = 0;
    if (r->pipelined || r->graph_run) {
        if (ms_stage) *ms_stage = 0.f;
        if (ms_reduce) *ms_reduce = 0.f;
        if (ms_force) { HIPCHK(c, hipEventElapsedTime(&t, r->ev[0], r->ev[3]), "hipEventElapsedTime"); *ms_force = t; }
        return PTH_OK;
    }
    if (ms_stage) { HIPCHK(c, hipEventElapsedTime(&t, r->ev[0], r->ev[1]), "hipEventElapsedTime"); *ms_stage = t; }
    if (ms_force) { HIPCHK(c, hipEventElapsedTime(&t, r->ev[1], r->ev[2]), "hipEventElapsedTime"); *ms_force = t; }
    if (ms_reduce) { HIPCHK(c, hipEventElapsedTime(&t, r->ev[2], r->ev[3]), "hipEventElapsedTime"); *ms_reduce = t; }
    return PTH_OK;
}

int pth_resident_replan(pth_resident *r, int max_ipl, int chunk_j)
{
    if (!r) return PTH_ERR_ARG;
    pth_ctx *c = r->ctx;
    if (c->par.exec_mode != PTH_EXEC_FUSED) return c->fail(PTH_ERR_STATE, "pth_resident_replan: fused form only");
    const Plan &P0 = r->plan;
    const int nw = (int)P0.walks.size();
    std::vector<int> ne(nw), nj(nw), ns(nw);
    for (int w = 0; w < nw; w++) {
        ne[w] = P0.walks[w].n_i;
        nj[w] = P0.walks[w].n_e;
        ns[w] = P0.walks[w].n_s;
    }
    Plan P;
    if (!make_plan(nw, ne.data(), nj.data(), ns.data(), PTH_EXEC_FUSED, chunk_j, max_ipl, P, r->n_seg, P0.mono,
                   seg_pairs(c->par)))
        return c->fail(PTH_ERR_CAPACITY, "pth_resident_replan: sizes exceed the int32 index range");
    // the staged layout depends on the walk sizes only
    if (P.es_off != P0.es_off || P.ss_off != P0.ss_off || P.ni_tot != P0.ni_tot)
        return c->fail(PTH_ERR_STATE, "pth_resident_replan: staged layout changed");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    const size_t bi = sizeof(Item) * P.items.size(), br = sizeof(RItem) * P.ritems.size();
    const size_t off_r = align256(bi);
    HIPCHK(c, r->plan_dev.ensure(std::max<size_t>(off_r + br, 256)), "plan allocation");
    HIPCHK(c, r->work.ensure(P, r->co, df_enabled(c->par) != 0), "resident work buffers");
    if (bi) HIPCHK(c, hipMemcpy(r->plan_dev.p, P.items.data(), bi, hipMemcpyHostToDevice), "plan upload");
    if (br) HIPCHK(c, hipMemcpy(r->plan_dev.as<char>() + off_r, P.ritems.data(), br, hipMemcpyHostToDevice),
                   "plan upload");
    fill_batch(r->b, P, c->par, r->arena, r->A, r->work, r->epj.as<EpjDev>(), r->spj.as<SpjDev>(),
               r->out_accp.as<double4>(), r->out_nnb.as<int>(), r->epj_co.as<EpjCoDev>(), r->out_pott.as<double>(),
               r->spj_tab.as<SpjTab>());
    r->b.items = r->plan_dev.as<const Item>();
    r->b.ritems = reinterpret_cast<const RItem *>(r->plan_dev.as<char>() + off_r);
    r->b_gen++;
    r->plan = std::move(P);
    r->max_ipl = max_ipl;
    r->chunk_j = chunk_j;
    return PTH_OK;
}

int pth_resident_autotune(pth_resident *r, const int *ipl, int n_ipl, const int *chunk, int n_chunk, int reps,
                          int *best_ipl, int *best_chunk, float *best_ms)
{
    if (!r) return PTH_ERR_ARG;
    pth_ctx *c = r->ctx;
    if (n_ipl <= 0 || n_chunk <= 0 || !ipl || !chunk || reps <= 0)
        return c->fail(PTH_ERR_ARG, "pth_resident_autotune: empty candidate list");
    if (r->co && !r->co_ready)
        return c->fail(PTH_ERR_STATE, "pth_resident_autotune: changeover mode needs pth_resident_changeover first");
    float bt = 1e30f;
    int bi = ipl[0], bc = chunk[0];
    for (int a = 0; a < n_ipl; a++)
        for (int k = 0; k < n_chunk; k++) {
            int rc = pth_resident_replan(r, ipl[a], chunk[k]);
            if (rc) return rc;
            if ((rc = pth_resident_run(r))) return rc;  // warm-up
            std::vector<float> t(reps);
            for (int q = 0; q < reps; q++) {
                if ((rc = pth_resident_run(r))) return rc;
                float s = 0, f = 0, d = 0;
                if ((rc = pth_resident_timing(r, &s, &f, &d))) return rc;
                t[q] = s + f + d;
            }
            std::nth_element(t.begin(), t.begin() + reps / 2, t.end());
            if (t[reps / 2] < bt) {
                bt = t[reps / 2];
                bi = ipl[a];
                bc = chunk[k];
            }
        }
    const int rc = pth_resident_replan(r, bi, bc);
    if (rc) return rc;
    if (best_ipl) *best_ipl = bi;
    if (best_chunk) *best_chunk = bc;
    if (best_ms) *best_ms = bt;
    return PTH_OK;
}

int pth_resident_stats(const pth_resident *r, long long *n_epep, long long *n_epsp, long long *n_items,
                       long long *staged_bytes)
{
    if (!r) return PTH_ERR_ARG;
    if (n_epep) *n_epep = r->plan.n_epep;
    if (n_epsp) *n_epsp = r->plan.n_epsp;
    if (n_items) *n_items = r->plan.n_real_items;
    if (staged_bytes) *staged_bytes = r->staged_bytes;
    return PTH_OK;
}

void pth_resident_free(pth_resident *r)
{
    if (!r) return;
    if (r->ctx && r->ctx->device >= 0) {
        (void)hipSetDevice(r->ctx->device);
        if (r->ctx->stream) (void)hipStreamSynchronize(r->ctx->stream);
    }
    r->arena.release();
    r->epj.release();
    r->spj.release();
    r->out_accp.release();
    r->out_nnb.release();
    r->epj_co.release();
    r->out_pott.release();
    r->plan_dev.release();
    r->work.release();
    for (int k = 0; k < 4; k++)
        if (r->ev[k]) (void)hipEventDestroy(r->ev[k]);
    for (hipEvent_t x : r->side_done)
        if (x) (void)hipEventDestroy(x);
    for (hipEvent_t x : r->staged)
        if (x) (void)hipEventDestroy(x);
    for (hipStream_t x : r->side)
        if (x) (void)hipStreamDestroy(x);
    if (r->gexec) (void)hipGraphExecDestroy(r->gexec);
    if (r->fork) (void)hipEventDestroy(r->fork);
    if (r->cap) (void)hipStreamDestroy(r->cap);
    delete r;
}

long long pth_plan_debug(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj, int exec_mode,
                         int chunk_j, int max_ipl, int n_seg, int *items, long long max_items, long long *es_off,
                         long long *ss_off, int *ritems, long long max_ritems, long long *n_ritems, int *segs)
{
    if (n_walk < 0 || (n_walk > 0 && (!n_epi || !n_epj || !n_spj))) return PTH_ERR_ARG;
    Plan P;
    if (!make_plan(n_walk, n_epi, n_epj, n_spj, exec_mode, chunk_j, max_ipl, P, n_seg)) return PTH_ERR_CAPACITY;
    if (segs) {
        for (int k = 0; k < PTH_MAX_SEGMENTS; k++) {
            int *o = segs + 12 * k;
            for (int q = 0; q < 12; q++) o[q] = -1;
            if (k >= (int)P.segs.size()) continue;
            const PlanSeg &g = P.segs[k];
            o[0] = g.walk_begin;
            o[1] = g.walk_end;
            for (int c = 0; c < PLAN_CLASSES; c++) {
                o[2 + c] = g.seg_begin[c];
                o[6 + c] = g.seg_count[c];
            }
            o[10] = g.ritem_begin;
            o[11] = g.ritem_count;
        }
    }
    if (es_off) std::copy(P.es_off.begin(), P.es_off.end(), es_off);
    if (ss_off) std::copy(P.ss_off.begin(), P.ss_off.end(), ss_off);
    if (n_ritems) *n_ritems = (long long)P.ritems.size();
    if (ritems) {
        if ((long long)P.ritems.size() > max_ritems) return PTH_ERR_CAPACITY;
        std::memcpy(ritems, P.ritems.data(), sizeof(RItem) * P.ritems.size());
    }
    if (items) {
        if ((long long)P.items.size() > max_items) return PTH_ERR_CAPACITY;
        std::memcpy(items, P.items.data(), sizeof(Item) * P.items.size());
    }
    return (long long)P.items.size();
}

}  // extern "C"

namespace pth {

int resident_setup(pth_resident *r, int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                   long long n_ide, long long n_ids, int n_epj_tot, int n_spj_tot, const char *who, Plan *pre,
                   bool pre_ok)
{
    pth_ctx *c = r->ctx;
    r->ran = false;
    r->co_ready = false;
    r->pipelined = false;
    r->n_seg = (c->par.exec_mode == PTH_EXEC_FUSED) ? c->par.segments : 1;
    if (pre) {
        if (!pre_ok) return c->fail(PTH_ERR_CAPACITY, std::string(who) + ": sizes exceed the int32 index range");
        std::swap(r->plan, *pre);
    } else if (!make_plan(n_walk, n_epi, n_epj, n_spj, c->par.exec_mode, c->par.chunk_j, c->par.max_ipl, r->plan,
                          r->n_seg, c->par.spj_monopole != 0, seg_pairs(c->par)))
        return c->fail(PTH_ERR_CAPACITY, std::string(who) + ": sizes exceed the int32 index range");
    const Plan &P = r->plan;
    r->co = c->par.changeover != 0;
    if (r->co) {
        const char *m = co_plan_error(P);
        if (m) return c->fail(PTH_ERR_CAPACITY, std::string(who) + ": " + m);
    }
    r->A.layout(P, n_ide, n_ids, r->co);
    auto bail = [&](hipError_t e, const char *what) { return c->hip_fail(e, what); };
    hipError_t e;
    if ((e = hipSetDevice(c->device))) return bail(e, "hipSetDevice");
    if ((e = r->arena.ensure(r->A.total))) return bail(e, "resident arena");
    if ((e = r->epj.ensure(sizeof(EpjDev) * std::max(n_epj_tot, 1)))) return bail(e, "resident EPJ");
    if ((e = r->spj.ensure((P.mono ? sizeof(EpjDev) : sizeof(SpjDev)) * std::max(n_spj_tot, 1))))
        return bail(e, "resident SPJ");
    if ((e = r->out_accp.ensure(sizeof(double4) * std::max<long long>(P.ni_tot, 1)))) return bail(e, "resident output");
    if ((e = r->out_nnb.ensure(sizeof(int) * std::max<long long>(P.ni_tot, 1)))) return bail(e, "resident output");
    if ((e = r->work.ensure(P, r->co, df_enabled(c->par) != 0))) return bail(e, "resident work buffers");
    if (r->co) {
        if ((e = r->epj_co.ensure(sizeof(EpjCoDev) * std::max(n_epj_tot, 1)))) return bail(e, "resident EPJ");
        if ((e = r->out_pott.ensure(sizeof(double) * std::max<long long>(P.ni_tot, 1)))) return bail(e, "resident output");
    }
    for (int k = 0; k < 4; k++)
        if (!r->ev[k] && (e = hipEventCreate(&r->ev[k]))) return bail(e, "hipEventCreate");
    {
        // (one side stream also for a single segment: its 1-i class)
        const int ns = std::max(1, std::min(c->par.streams, (int)P.segs.size()));
        // side streams are only ever added (a later plan may use fewer)
        while ((int)r->side.size() < ns) {
            r->side.push_back(nullptr);
            r->side_done.push_back(nullptr);
            if ((e = hipStreamCreateWithFlags(&r->side.back(), hipStreamNonBlocking))) return bail(e, "hipStreamCreate");
            if ((e = hipEventCreateWithFlags(&r->side_done.back(), hipEventDisableTiming)))
                return bail(e, "hipEventCreate");
        }
        while (r->staged.size() < P.segs.size()) {
            r->staged.push_back(nullptr);
            if ((e = hipEventCreateWithFlags(&r->staged.back(), hipEventDisableTiming))) return bail(e, "hipEventCreate");
        }
    }
    r->n_epj_tot = n_epj_tot;
    r->n_spj_tot = n_spj_tot;
    if (!P.mono && (e = r->spj_tab.ensure(sizeof(SpjTab) * (size_t)std::max(n_spj_tot, 1))))
        return bail(e, "resident SPJ table");
    return PTH_OK;
}

int resident_prepare(pth_ctx *c, pth_resident *&out, int n_walk, const int *n_epi, const int *n_epj,
                     const int *n_spj, long long n_ide, long long n_ids, int n_epj_tot, int n_spj_tot,
                     const char *who)
{
    out = nullptr;
    pth_resident *r = new pth_resident;
    r->ctx = c;
    const int rc = resident_setup(r, n_walk, n_epi, n_epj, n_spj, n_ide, n_ids, n_epj_tot, n_spj_tot, who);
    if (rc != PTH_OK) {
        pth_resident_free(r);
        return rc;
    }
    out = r;
    return PTH_OK;
}

hipError_t resident_finish(pth_resident *r)
{
    const Plan &P = r->plan;
    pth_ctx *c = r->ctx;
    if (!P.mono) {
        const hipError_t e = launch_spj_table(r->spj.as<SpjDev>(), r->n_spj_tot, (float)c->par.eps2,
                                              r->spj_tab.as<SpjTab>(), c->stream);
        if (e != hipSuccess) return e;
    }
    fill_batch(r->b, P, c->par, r->arena, r->A, r->work, r->epj.as<EpjDev>(), r->spj.as<SpjDev>(),
               r->out_accp.as<double4>(), r->out_nnb.as<int>(), r->epj_co.as<EpjCoDev>(),
               r->out_pott.as<double>(), P.mono ? nullptr : r->spj_tab.as<SpjTab>());
    r->b_gen++;
    r->max_ipl = c->par.max_ipl;
    r->chunk_j = c->par.chunk_j;
    r->staged_bytes = (long long)(16 * P.ni_tot + 20 * P.n_e_slots + 4 * PTH_SPJ_STAGED * P.n_s_rec);
    return hipSuccess;
}

}  // namespace pth
