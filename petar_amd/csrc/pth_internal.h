// pth_internal.h -- host-side internals of libpetar_hip.so shared by the
// context / resident code (pth_context.cpp) and the serving loop
// (pth_server.cpp): grow-only device and pinned buffers, the upload arena
// layout of one batch, the batch work buffers, the per-context state.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/petar_hip.h"
#include "pth_device.h"
#include "pth_plan.h"

namespace pth {

constexpr int MAX_TAG = 8;

// Test hook: PTH_FAIL_INJECT=alloc makes every device allocation fail, so
// the error paths can be exercised from the host.
inline bool inject_alloc_failure()
{
    static int v = -1;
    if (v < 0) {
        const char *e = std::getenv("PTH_FAIL_INJECT");
        v = (e && std::strcmp(e, "alloc") == 0) ? 1 : 0;
    }
    return v == 1;
}

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    // grows (never shrinks); the caller has synchronised the stream
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (inject_alloc_failure()) return hipErrorOutOfMemory;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (inject_alloc_failure()) return hipErrorOutOfMemory;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    // grows like ensure, keeping the first `keep` bytes
    hipError_t ensure_keep(size_t bytes, size_t keep)
    {
        if (bytes <= cap) return hipSuccess;
        if (inject_alloc_failure()) return hipErrorOutOfMemory;
        void *q = nullptr;
        const size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
        hipError_t e = hipHostMalloc(&q, want, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        if (p) {
            std::memcpy(q, p, std::min(keep, cap));
            (void)hipHostFree(p);
        }
        p = q;
        cap = want;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Layout of one upload arena (pinned host mirror and device copy).  The
// index lists travel either as they are (ide, ids) or, with n_run >= 0, as
// runs of consecutive indices (pth_dispatch): run_e / run_s hold the first
// index of every run, then every run's position in its walk's list (two
// int32 per run), run_eo / run_so the walks'
// run offsets (n_walk + 1 each), and the lists are expanded on the device.
struct Arena {
    size_t epi, epi_id, walks, es, ss, ide, ids, items, ritems, run_e, run_s, run_eo, run_so, total;
    bool runs = false;
    // the i records (and ids) lead: [0, i_part(ni_tot, co)), i ids at
    // i_id_offset(ni_tot)
    static size_t i_id_offset(long long ni_tot) { return align256(sizeof(EpiDev) * (size_t)ni_tot); }
    static size_t i_part(long long ni_tot, bool co)
    {
        return align256(i_id_offset(ni_tot) + (co ? sizeof(long long) * (size_t)ni_tot : 0));
    }
    void layout(const Plan &P, long long n_ide, long long n_ids, bool co, long long n_run_e = -1,
                long long n_run_s = -1)
    {
        runs = n_run_e >= 0 && n_run_s >= 0;
        epi = 0;
        epi_id = i_id_offset(P.ni_tot);
        size_t o = i_part(P.ni_tot, co);
        walks = o; o = align256(o + sizeof(WalkDev) * P.walks.size());
        es = o; o = align256(o + sizeof(long long) * P.es_off.size());
        ss = o; o = align256(o + sizeof(long long) * P.ss_off.size());
        ide = o; o = align256(o + (runs ? 0 : sizeof(int) * (size_t)n_ide));
        ids = o; o = align256(o + (runs ? 0 : sizeof(int) * (size_t)n_ids));
        items = o; o = align256(o + sizeof(Item) * P.items.size());
        ritems = o; o = align256(o + sizeof(RItem) * P.ritems.size());
        run_e = o; o = align256(o + (runs ? 2 * sizeof(int) * (size_t)n_run_e : 0));
        run_s = o; o = align256(o + (runs ? 2 * sizeof(int) * (size_t)n_run_s : 0));
        run_eo = o; o = align256(o + (runs ? sizeof(int) * (P.walks.size() + 1) : 0));
        run_so = o; o = align256(o + (runs ? sizeof(int) * (P.walks.size() + 1) : 0));
        total = o;
    }
};

// Device work buffers of one batch (staging, partials, outputs).
struct Work {
    DevBuf centre, near_blk, staged_i, staged_e, staged_s, part_accp, part_nnb;
    DevBuf staged_ic, staged_il, staged_ec, part_pott, walk_err;  // changeover mode
    DevBuf staged_el;                                             // linear mode, double-float near pairs
    DevBuf counters;                                              // persistent force launches
    hipError_t ensure(const Plan &P, bool co, bool df)
    {
        hipError_t e;
        if (df) {
            if ((e = staged_il.ensure(sizeof(float4) * std::max<long long>(P.ni_tot, 1)))) return e;
            if ((e = staged_el.ensure(sizeof(float) * (PTH_DF_EXTRA / 8) * std::max<long long>(P.n_e_slots, 8))))
                return e;
        }
        if (co) {
            if ((e = staged_ic.ensure(sizeof(float4) * std::max<long long>(P.ni_tot, 1)))) return e;
            if ((e = staged_il.ensure(sizeof(float4) * std::max<long long>(P.ni_tot, 1)))) return e;
            if ((e = staged_ec.ensure(sizeof(float) * (PTH_CO_EXTRA / 8) * std::max<long long>(P.n_e_slots, 8))))
                return e;
            if ((e = part_pott.ensure(sizeof(double) * 64 * std::max(P.n_part, 1)))) return e;
            if ((e = walk_err.ensure(sizeof(int) * std::max<size_t>(P.walks.size(), 1)))) return e;
        }
        if ((e = centre.ensure(sizeof(double4) * std::max<size_t>(P.walks.size(), 1)))) return e;
        if ((e = near_blk.ensure(sizeof(int) * std::max<size_t>(P.walks.size(), 1)))) return e;
        if ((e = staged_i.ensure(sizeof(float4) * std::max<long long>(P.ni_tot, 1)))) return e;
        if ((e = staged_e.ensure(sizeof(float) * 5 * std::max<long long>(P.n_e_slots, 8)))) return e;
        if ((e = staged_s.ensure(sizeof(float) * PTH_SPJ_STAGED * std::max<long long>(P.n_s_rec, 1)))) return e;
        if ((e = part_accp.ensure(sizeof(double4) * 64 * std::max(P.n_part, 1)))) return e;
        if ((e = part_nnb.ensure(sizeof(int) * 64 * std::max(P.n_part, 1)))) return e;
        if ((e = counters.ensure(sizeof(int) * PTH_MAX_SEGMENTS * PLAN_CLASSES * 8))) return e;
        return hipSuccess;
    }
    void release()
    {
        centre.release(); near_blk.release(); staged_i.release(); staged_e.release(); staged_s.release();
        part_accp.release(); part_nnb.release();
        staged_ic.release(); staged_il.release(); staged_ec.release(); part_pott.release(); walk_err.release();
        staged_el.release();
        counters.release();
    }
};

// Whether [p, p + bytes) lies inside one device allocation of this process
// (host pointers and ranges overrunning their allocation are rejected).
inline bool device_range(const void *p, size_t bytes)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (at.type != hipMemoryTypeDevice) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return (const char *)p + bytes <= (const char *)base + size;
}

// Fewest interactions per pipelined segment of a resident evaluation.
inline long long seg_pairs(const pth_params &par)
{
    return par.min_seg_pairs > 0 ? (long long)par.min_seg_pairs : MIN_SEGMENT_PAIRS;
}

// Linear-cutoff mode with double-float near pairs (fused form; the
// changeover mode has its own double-float neighbour pairs).
inline int df_enabled(const pth_params &par)
{
    return (par.df_near && !par.changeover && par.exec_mode == PTH_EXEC_FUSED) ? 1 : 0;
}

// Copies the plan's segments into the launch descriptor.
inline void set_segments(BatchDev &b, const Plan &P)
{
    b.n_seg = (int)std::min<size_t>(std::max<size_t>(P.segs.size(), 1), PTH_MAX_SEGMENTS);
    for (int k = 0; k < b.n_seg; k++) {
        SegDev &d = b.seg[k];
        d = SegDev();
        if (k >= (int)P.segs.size()) {
            d.n_walk = (int)P.walks.size();
            continue;
        }
        const PlanSeg &g = P.segs[k];
        d.walk_begin = g.walk_begin;
        d.n_walk = g.walk_end - g.walk_begin;
        for (int c = 0; c < PLAN_CLASSES; c++) {
            d.seg_begin[c] = g.seg_begin[c];
            d.seg_count[c] = g.seg_count[c];
        }
        d.ritem_begin = g.ritem_begin;
        d.n_ritems = g.ritem_count;
    }
}

inline void fill_batch(BatchDev &b, const Plan &P, const pth_params &par, const DevBuf &arena,
                const Arena &A, Work &W, const EpjDev *epj, const SpjDev *spj, double4 *out_accp,
                int *out_nnb, const EpjCoDev *epj_co, double *out_pott, const SpjTab *spj_tab)
{
    b.n_walk = (int)P.walks.size();
    b.ni_tot = P.ni_tot;
    b.n_e_slots = P.n_e_slots;
    b.n_s_rec = P.n_s_rec;
    b.n_items = (int)P.items.size();
    b.n_ritems = (int)P.ritems.size();
    set_segments(b, P);
    b.j_vgpr = par.j_vgpr ? 1 : 0;
    b.plain = (par.exec_mode == PTH_EXEC_PLAIN);
    b.nnb_le = (par.nnb_mode == 1);
    b.eps2 = (float)par.eps2;
    b.rout2 = (float)par.rcut2;
    b.G = par.G;
    b.co = par.changeover ? 1 : 0;
    b.df = df_enabled(par);
    b.mono = P.mono ? 1 : 0;
    {
        const double f = par.ngb_factor > 0.0 ? par.ngb_factor : 0.99;
        b.ngb2 = (float)(f * f);
    }
    char *base = arena.as<char>();
    b.epi = reinterpret_cast<const EpiDev *>(base + A.epi);
    b.walks = reinterpret_cast<const WalkDev *>(base + A.walks);
    b.es_off = reinterpret_cast<const long long *>(base + A.es);
    b.ss_off = reinterpret_cast<const long long *>(base + A.ss);
    b.id_epj = reinterpret_cast<const int *>(base + A.ide);
    b.id_spj = reinterpret_cast<const int *>(base + A.ids);
    b.items = reinterpret_cast<const Item *>(base + A.items);
    b.ritems = reinterpret_cast<const RItem *>(base + A.ritems);
    b.epj = epj;
    b.spj = spj;
    b.spj_tab = spj_tab;
    b.centre = W.centre.as<double4>();
    b.near_blk = W.near_blk.as<int>();
    b.staged_i = W.staged_i.as<float4>();
    b.staged_e = W.staged_e.as<float>();
    b.staged_s = W.staged_s.as<float>();
    b.part_accp = W.part_accp.as<double4>();
    b.part_nnb = W.part_nnb.as<int>();
    b.out_accp = out_accp;
    b.out_nnb = out_nnb;
    b.counters = W.counters.as<int>();
    b.persistent = (par.persistent && b.counters) ? 1 : 0;
    if (b.df) {
        b.staged_il = W.staged_il.as<float4>();
        b.staged_el = W.staged_el.as<float>();
    }
    if (b.co) {
        b.epi_id = reinterpret_cast<const long long *>(base + A.epi_id);
        b.epj_co = epj_co;
        b.staged_ic = W.staged_ic.as<float4>();
        b.staged_il = W.staged_il.as<float4>();
        b.staged_ec = W.staged_ec.as<float>();
        b.walk_err = W.walk_err.as<int>();
        b.part_pott = W.part_pott.as<double>();
        b.out_pott = out_pott;
    }
}

// Changeover data of one EPJ from its radii, id and group data (mass_backup,
// status; null: a single); types of src/artificial_particles.hpp:60-102, a
// member's c.m. address unset (src/ptcl.hpp:215-222) when its status is
// -PS::LARGE_FLOAT, taken as any status below -1e30 (an address is an
// array index).
inline EpjCoDev co_record(double rin, double rout, long long id, const double *group)
{
    EpjCoDev d;
    d.rin = (float)rin;
    d.rout = (float)rout;
    d.id = id;
    d.mback = 0.f;
    d.type = PTH_CO_SINGLE;
    if (group) {
        const double mb = group[0], st = group[1];
        if (st == 0.0 && mb == 0.0) d.type = PTH_CO_SINGLE;
        else if (st < 0.0) {
            d.type = st < -1e30 ? PTH_CO_MEMBER_NO_CM : PTH_CO_MEMBER;
            d.mback = (float)mb;
        } else d.type = PTH_CO_ARTIFICIAL;
    }
    return d;
}

// Checks of the changeover-mode preconditions on a plan.
inline const char *co_plan_error(const Plan &P)
{
    for (const WalkDev &w : P.walks)
        if (w.n_i > PTH_CO_MAX_NI) return "changeover mode: a walk has more than 1024 i-particles";
    if (P.n_e_slots >= (1ll << 31)) return "changeover mode: staged EPJ slots exceed the int32 range";
    return nullptr;
}

inline void copy_plan_to_arena(char *h, const Plan &P, const Arena &A)
{
    std::memcpy(h + A.walks, P.walks.data(), sizeof(WalkDev) * P.walks.size());
    std::memcpy(h + A.es, P.es_off.data(), sizeof(long long) * P.es_off.size());
    std::memcpy(h + A.ss, P.ss_off.data(), sizeof(long long) * P.ss_off.size());
    if (!P.items.empty()) std::memcpy(h + A.items, P.items.data(), sizeof(Item) * P.items.size());
    if (!P.ritems.empty()) std::memcpy(h + A.ritems, P.ritems.data(), sizeof(RItem) * P.ritems.size());
}

}  // namespace pth

using namespace pth;

struct TagState {
    DevBuf out_accp, out_nnb, out_pott;
    PinBuf h_accp, h_nnb, h_pott, h_err;
    bool co = false;
    std::vector<int> n_i;
    long long ni_tot = 0;
    bool pending = false;
    hipEvent_t done = nullptr;
    // profile (timing events): before the upload, after it, after the
    // kernels, after the result copy; timed says the batch recorded them
    hipEvent_t prof[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timed = false;
};

struct pth_ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    pth_params par{};
    std::string err;
    // resident j arrays of the current force call
    DevBuf d_epj, d_spj, d_epj_co;
    PinBuf h_epj, h_spj, h_epj_co;
    int n_epj = 0, n_spj = 0;
    bool j_sent = false;
    bool co_sent = false;  // the last send carried the changeover data
    bool mono_sent = false;  // the last send held monopole superparticles
    hipEvent_t ev_send = nullptr;
    // batch upload
    PinBuf h_arena;
    DevBuf d_arena;
    // index lists as runs: host encoding space (reused), the walks' run
    // counts, and the device lists they expand into (shared by all tags, as
    // the work buffers)
    std::vector<int> h_rid, h_rpos;
    std::vector<int> run_ne, run_ns;
    std::vector<long long> w_off;  // pth_dispatch: prefix sums of the walks' i, EPJ, SPJ counts
    std::vector<int> ident;  // 0, 1, 2, ...: the lists of pth_dispatch_copies
    DevBuf d_ids;
    // SPJ table of the sent superparticles, built at the first dispatch
    // after a send (again when eps2 changes)
    DevBuf d_spj_tab;
    bool tab_valid = false;
    float tab_eps2 = 0.f;
    int list_runs = -1;  // PTH_LIST_RUNS: -1 per batch, 0 never, 1 always
    // pth_dispatch runs the 1-i items on a second stream beside the longer
    // classes (PTH_CLASS_STREAMS=0: all on one stream)
    bool class_streams = true;
    // pth_dispatch's kernels write the results straight into the tag's pinned
    // host buffers (PTH_ZERO_COPY=0: device buffers and a copy)
    bool zero_copy = true;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_upload = nullptr;
    Work work;
    TagState tag[MAX_TAG];
    // profile of the hooks (pth_profile_*)
    bool prof_on = false;
    pth_profile prof{};
    hipEvent_t prof_send[2] = {nullptr, nullptr};  // around the last send's copies
    bool prof_send_pending = false;
    // pth_resident_eval_noalloc: plan and launch descriptor kept between
    // calls (the plan's vectors keep their capacity)
    Plan na_plan;
    BatchDev na_batch;
    // pth_dispatch's plan, kept between batches: fresh vectors of a few
    // hundred kB would be fresh pages, whose faults cost more than the planning
    Plan d_plan;

    int fail(int code, const std::string &msg)
    {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char *what)
    {
        err = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
        return PTH_ERR_HIP;
    }
};

struct pth_resident {
    pth_ctx *ctx = nullptr;
    Plan plan;
    Arena A{};
    DevBuf arena, epj, spj, out_accp, out_nnb;
    DevBuf epj_co, out_pott;  // changeover mode
    bool co = false, co_ready = false;
    int n_epj_tot = 0, n_spj_tot = 0;
    DevBuf spj_tab;  // SPJ table (quadrupole superparticles)
    DevBuf plan_dev;  // work items and reduction items after a re-plan
    int max_ipl = 0, chunk_j = 0;
    int n_seg = 1;                       // walk segments of the plan
    std::vector<hipStream_t> side;       // streams the segments run on
    std::vector<hipEvent_t> side_done;   // one per side stream
    std::vector<hipEvent_t> staged;      // one per segment: its staging is done
    bool pipelined = false;              // the last run used the side streams
    // graph mode: the evaluation captured on its own stream (fork and join
    // events over the side streams) and replayed; b_gen counts rebuilds of
    // the launch descriptor, g_gen the one the graph holds
    hipStream_t cap = nullptr;
    hipEvent_t fork = nullptr;
    hipGraphExec_t gexec = nullptr;
    long long b_gen = 0, g_gen = -1;
    bool graph_run = false;              // the last run replayed the graph
    Work work;
    BatchDev b;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ran = false;
    long long staged_bytes = 0;
};

namespace pth {
// A resident evaluation with its plan made and its device memory allocated
// (contents not yet uploaded); on failure the context holds the error and
// out is null.
int resident_prepare(pth_ctx *c, pth_resident *&out, int n_walk, const int *n_epi, const int *n_epj,
                     const int *n_spj, long long n_ide, long long n_ids, int n_epj_tot, int n_spj_tot,
                     const char *who);
// Re-plans an existing resident for new sizes, growing its device memory as
// needed (contents not yet uploaded; the changeover data is invalidated).
// With pre, that plan (made for these sizes and the context's parameters;
// pre_ok its make_plan result) is taken over instead of planning again.
int resident_setup(pth_resident *r, int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                   long long n_ide, long long n_ids, int n_epj_tot, int n_spj_tot, const char *who,
                   Plan *pre = nullptr, bool pre_ok = true);
// Builds the SPJ table and the launch descriptors of a prepared resident
// whose contents are uploaded (or being uploaded on the context's stream).
hipError_t resident_finish(pth_resident *r);
// Looks up this device's address of the allocation-free entry point's
// scratch array (once per device; the device must be current).
void noalloc_setup(int device);
}  // namespace pth

#define HIPCHK(ctx, call, what)                      \
    do {                                             \
        hipError_t e_ = (call);                      \
        if (e_ != hipSuccess) return (ctx)->hip_fail(e_, what); \
    } while (0)

