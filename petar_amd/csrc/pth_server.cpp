// pth_server.cpp -- the serving loop of include/petar_hip.h (pth_server_*):
// many independent soft-force evaluations ("clusters") behind one queue.
//
// Fused mode: one device block of hbm_budget bytes is the pool.  A cluster is
// admitted by giving it one contiguous range of the pool, [upload image:
// batch arena, EPJ, SPJ | outputs], into which its pinned host image is
// copied; it stays resident until evicted.  What a request only needs while
// it runs (the staged lists, the partial sums of split tiles) lives in a
// scratch range per stream, also from the pool, sized for the largest
// cluster the stream has run, so a resident cluster holds only its inputs
// and outputs.  Each request runs as one hipGraph launch on one of
// n_streams streams, so requests overlap on the device; a cluster keeps one
// instantiated graph per stream, captured the first time it runs there and
// updated in place (hipGraphExecUpdate) when its range or the stream's
// scratch has moved, so re-admission captures a few launches but
// instantiates nothing.  Idle clusters are evicted least recently used first
// when an admission or a scratch range needs room: their results are copied
// to pinned host memory on a transfer stream, and every later upload or
// launch waits for those copies on the device, so the worker never blocks on
// them.  A worker thread does admission, launches and completion
// (hipEventQuery polling) under one mutex.
//
// Plain mode: each request replays PeTar's multi-walk call order on a
// PTH_EXEC_PLAIN context (send all EPJ/SPJ, then <= 200-walk batches
// dispatched and retrieved with synchronous copies), one request at a time.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "pth_internal.h"

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator of ranges of one device block (sizes are multiples of
// 256 bytes).
struct PoolAlloc {
    std::map<size_t, size_t> free_;  // offset -> length
    size_t used = 0, peak = 0;
    void init(size_t n)
    {
        free_.clear();
        if (n) free_[0] = n;
        used = peak = 0;
    }
    long long alloc(size_t n)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                const size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[off + n] = len - n;
                used += n;
                peak = std::max(peak, used);
                return (long long)off;
            }
        return -1;
    }
    void release(size_t off, size_t n)
    {
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

constexpr int MAX_STREAMS = 16;

// Byte offsets of a cluster's resident range: the upload image (batch
// arena, EPJ, SPJ) and the outputs.
struct Layout {
    Arena A{};
    size_t epj = 0, spj = 0, upload = 0, out_accp = 0, out_nnb = 0, total = 0;
};

// Byte offsets of what one request of a cluster needs while it runs, inside
// a stream's scratch range.
struct Scratch {
    size_t centre = 0, near_blk = 0, staged_i = 0, staged_e = 0, staged_s = 0, part_accp = 0, part_nnb = 0,
           staged_il = 0, staged_el = 0, spj_tab = 0, total = 0;
};

// Host records of the plain mode (FDPS-like array-of-structs).
struct PEpj {
    double pos[3], mass, rs;
};
struct PSpj {
    double pos[3], mass, q[6];
};
struct PEpi {
    double pos[3], rs;
};
struct PForce {
    double acc[3], pot;
    long long nnb;
};

// A cluster's instantiated graph on one stream and the placements it was
// made for.
struct SGraph {
    hipGraphExec_t exec = nullptr;
    long long cl_gen = -1, sc_gen = -1;
};

struct SCluster {
    Plan plan;
    Layout L;
    Scratch S;
    int n_walk = 0, n_epj_tot = 0, n_spj_tot = 0;
    long long ni = 0, n_int = 0;
    // fused mode
    PinBuf image;          // pinned host copy of the upload image
    long long off = -1;    // pool offset while resident
    long long gen = 0;     // admissions so far: a new one may place it elsewhere
    SGraph g[MAX_STREAMS];
    PinBuf h_accp, h_nnb;  // results copied out at eviction
    hipEvent_t saved = nullptr;  // that copy (on the transfer stream)
    bool busy = false;
    long long last_use = 0;
    bool ran = false;
    bool saved_valid = false;  // h_accp / h_nnb hold the last results (after `saved`)
    // plain mode
    std::vector<int> n_epi, n_epj, n_spj;
    std::vector<PEpj> p_epj;
    std::vector<PSpj> p_spj;
    std::vector<PEpi> p_epi;
    std::vector<PForce> p_force;
    std::vector<int> ide, ids;
    std::vector<const void *> epi_ptr;
    std::vector<const int *> ide_ptr, ids_ptr;
    std::vector<void *> force_ptr;
    bool host_valid = false;
};

struct Req {
    int cluster = -1;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int admitted = 0;  // the request uploaded its cluster
    double t_submit = 0.0, t_done = 0.0;
};

// One stream of the fused server: the request running on it and its scratch
// range.
struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;  // completion of the running request
    long long req = -1;
    long long sc_off = -1;
    size_t sc_size = 0;
    long long sc_gen = -1;       // placement of the scratch range
    long long evict_seen = 0;    // evictions this stream has waited for
};

}  // namespace

struct pth_server {
    pth_ctx *ctx = nullptr;
    int mode = PTH_SERVE_FUSED;
    size_t budget = 0;
    DevBuf pool;
    PoolAlloc pa;
    pth_ctx *pctx = nullptr;  // plain mode context
    std::vector<std::unique_ptr<SCluster>> cl;
    std::deque<Req> req;
    std::deque<long long> queue;
    int n_streams = 1;
    std::vector<Slot> slot;
    hipStream_t cap = nullptr;   // graph capture
    hipStream_t xfer = nullptr;  // result copies (evictions, pth_server_results)
    hipEvent_t ev_evicted = nullptr;  // the latest eviction copy
    long long n_evict_copies = 0;
    long long gen_clock = 0;     // placements of cluster ranges and scratch ranges
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread th;
    bool stop = false;
    bool failed = false;
    std::string err;
    long long n_done = 0, n_admit = 0, n_evict = 0, ints_done = 0, use_clock = 0;
    long long n_capture = 0, n_instantiate = 0;
};

namespace {

void layout(SCluster &c, long long n_ide, long long n_ids, bool df)
{
    const Plan &P = c.plan;
    Layout &L = c.L;
    L.A.layout(P, n_ide, n_ids, false);
    const size_t nw = std::max<size_t>(P.walks.size(), 1), ni = (size_t)std::max<long long>(P.ni_tot, 1);
    const size_t np = (size_t)std::max(P.n_part, 1);
    size_t o = L.A.total;
    L.epj = o; o = align256(o + sizeof(EpjDev) * (size_t)std::max(c.n_epj_tot, 1));
    L.spj = o; o = align256(o + sizeof(SpjDev) * (size_t)std::max(c.n_spj_tot, 1));
    L.upload = o;
    L.out_accp = o; o = align256(o + sizeof(double4) * ni);
    L.out_nnb = o; o = align256(o + sizeof(int) * ni);
    L.total = o;
    Scratch &S = c.S;
    o = 0;
    S.centre = o; o = align256(o + sizeof(double4) * nw);
    S.near_blk = o; o = align256(o + sizeof(int) * nw);
    S.staged_i = o; o = align256(o + sizeof(float4) * ni);
    S.staged_e = o; o = align256(o + sizeof(float) * 5 * (size_t)std::max<long long>(P.n_e_slots, 8));
    S.staged_s = o; o = align256(o + sizeof(float) * PTH_SPJ_STAGED * (size_t)std::max<long long>(P.n_s_rec, 1));
    S.spj_tab = o; o = align256(o + sizeof(SpjTab) * (size_t)std::max(c.n_spj_tot, 1));
    S.part_accp = o; o = align256(o + sizeof(double4) * 64 * np);
    S.part_nnb = o; o = align256(o + sizeof(int) * 64 * np);
    if (df) {
        S.staged_il = o; o = align256(o + sizeof(float4) * ni);
        S.staged_el = o;
        o = align256(o + sizeof(float) * (PTH_DF_EXTRA / 8) * (size_t)std::max<long long>(P.n_e_slots, 8));
    }
    S.total = o;
}

// Launch descriptor of cluster c with its range at base and a stream's
// scratch range at sc.
void bind(SCluster &c, BatchDev &b, char *base, char *sc, const pth_params &par)
{
    const Plan &P = c.plan;
    const Layout &L = c.L;
    const Scratch &S = c.S;
    b = BatchDev();
    b.n_walk = (int)P.walks.size();
    b.ni_tot = P.ni_tot;
    b.n_e_slots = P.n_e_slots;
    b.n_s_rec = P.n_s_rec;
    b.n_items = (int)P.items.size();
    b.n_ritems = (int)P.ritems.size();
    set_segments(b, P);
    b.j_vgpr = par.j_vgpr ? 1 : 0;
    b.nnb_le = (par.nnb_mode == 1);
    b.eps2 = (float)par.eps2;
    b.rout2 = (float)par.rcut2;
    b.G = par.G;
    b.epi = reinterpret_cast<const EpiDev *>(base + L.A.epi);
    b.walks = reinterpret_cast<const WalkDev *>(base + L.A.walks);
    b.es_off = reinterpret_cast<const long long *>(base + L.A.es);
    b.ss_off = reinterpret_cast<const long long *>(base + L.A.ss);
    b.id_epj = reinterpret_cast<const int *>(base + L.A.ide);
    b.id_spj = reinterpret_cast<const int *>(base + L.A.ids);
    b.items = reinterpret_cast<const Item *>(base + L.A.items);
    b.ritems = reinterpret_cast<const RItem *>(base + L.A.ritems);
    b.epj = reinterpret_cast<const EpjDev *>(base + L.epj);
    b.spj = reinterpret_cast<const SpjDev *>(base + L.spj);
    b.spj_tab = reinterpret_cast<const SpjTab *>(sc + S.spj_tab);
    b.out_accp = reinterpret_cast<double4 *>(base + L.out_accp);
    b.out_nnb = reinterpret_cast<int *>(base + L.out_nnb);
    b.centre = reinterpret_cast<double4 *>(sc + S.centre);
    b.near_blk = reinterpret_cast<int *>(sc + S.near_blk);
    b.staged_i = reinterpret_cast<float4 *>(sc + S.staged_i);
    b.staged_e = reinterpret_cast<float *>(sc + S.staged_e);
    b.staged_s = reinterpret_cast<float *>(sc + S.staged_s);
    b.part_accp = reinterpret_cast<double4 *>(sc + S.part_accp);
    b.part_nnb = reinterpret_cast<int *>(sc + S.part_nnb);
    b.df = df_enabled(par);
    if (b.df) {
        b.staged_il = reinterpret_cast<float4 *>(sc + S.staged_il);
        b.staged_el = reinterpret_cast<float *>(sc + S.staged_el);
    }
}

// Evicts an idle resident cluster: its results go to pinned host memory on
// the transfer stream (asynchronously; later uploads and launches wait for
// the copy on the device) and its range returns to the pool.
hipError_t evict(pth_server *s, SCluster &c)
{
    hipError_t e;
    if (c.ran && c.ni > 0) {
        char *base = s->pool.as<char>() + c.off;
        if ((e = hipMemcpyAsync(c.h_accp.p, base + c.L.out_accp, sizeof(double4) * c.ni, hipMemcpyDeviceToHost,
                                s->xfer)))
            return e;
        if ((e = hipMemcpyAsync(c.h_nnb.p, base + c.L.out_nnb, sizeof(int) * c.ni, hipMemcpyDeviceToHost,
                                s->xfer)))
            return e;
        if ((e = hipEventRecord(c.saved, s->xfer))) return e;
        if ((e = hipEventRecord(s->ev_evicted, s->xfer))) return e;
        c.saved_valid = true;
        s->n_evict_copies++;
    }
    s->pa.release((size_t)c.off, c.L.total);
    c.off = -1;
    s->n_evict++;
    return hipSuccess;
}

// A pool range of n bytes; idle clusters (least recently used first, never
// cluster `keep`) and then the scratch of idle streams (never stream
// `keep_slot`) give way.  Returns -1 while the memory is held by running
// requests (e stays hipSuccess) or on an error.
long long pool_alloc(pth_server *s, size_t n, int keep, int keep_slot, hipError_t &e)
{
    e = hipSuccess;
    for (;;) {
        const long long off = s->pa.alloc(n);
        if (off >= 0) return off;
        int victim = -1;
        for (int j = 0; j < (int)s->cl.size(); j++) {
            const SCluster &v = *s->cl[j];
            if (j != keep && v.off >= 0 && !v.busy && (victim < 0 || v.last_use < s->cl[victim]->last_use))
                victim = j;
        }
        if (victim >= 0) {
            if ((e = evict(s, *s->cl[victim]))) return -1;
            continue;
        }
        int idle = -1;
        for (int k = 0; k < s->n_streams; k++)
            if (k != keep_slot && s->slot[k].req < 0 && s->slot[k].sc_off >= 0) idle = k;
        if (idle < 0) return -1;
        Slot &q = s->slot[idle];
        s->pa.release((size_t)q.sc_off, q.sc_size);
        q.sc_off = -1;
        q.sc_size = 0;
    }
}

// Makes cluster k resident with stream `k_slot`'s scratch large enough for
// it, and its graph on that stream current; false while memory is held by
// running requests (e stays hipSuccess) or on an error.  `admitted` tells
// whether the cluster's image was uploaded.
bool admit(pth_server *s, int k, int k_slot, hipError_t &e, bool &admitted)
{
    SCluster &c = *s->cl[k];
    Slot &q = s->slot[k_slot];
    e = hipSuccess;
    admitted = false;
    if (q.sc_size < c.S.total) {
        if (q.sc_off >= 0) s->pa.release((size_t)q.sc_off, q.sc_size);
        q.sc_off = -1;
        q.sc_size = 0;
        const long long off = pool_alloc(s, c.S.total, k, k_slot, e);
        if (off < 0) return false;
        q.sc_off = off;
        q.sc_size = c.S.total;
        q.sc_gen = ++s->gen_clock;
    }
    if (c.off < 0) {
        const long long off = pool_alloc(s, c.L.total, k, k_slot, e);
        if (off < 0) return false;
        c.off = off;
        c.gen = ++s->gen_clock;
        admitted = true;
    }
    // the ranges just handed out may still be read by an eviction copy
    if (q.evict_seen != s->n_evict_copies) {
        if ((e = hipStreamWaitEvent(q.st, s->ev_evicted, 0))) return false;
        q.evict_seen = s->n_evict_copies;
    }
    char *base = s->pool.as<char>() + c.off;
    if (admitted) {
        if ((e = hipMemcpyAsync(base, c.image.p, c.L.upload, hipMemcpyHostToDevice, q.st))) return false;
        s->n_admit++;
    }
    SGraph &g = c.g[k_slot];
    if (!g.exec || g.cl_gen != c.gen || g.sc_gen != q.sc_gen) {
        BatchDev b;
        bind(c, b, base, s->pool.as<char>() + q.sc_off, s->ctx->par);
        hipGraph_t graph = nullptr;
        if ((e = hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal))) return false;
        // the SPJ table in the stream's scratch, then the evaluation
        hipError_t el = launch_spj_table(b.spj, c.n_spj_tot, b.eps2, const_cast<SpjTab *>(b.spj_tab), s->cap);
        if (el == hipSuccess) el = launch_all(b, s->cap);
        e = hipStreamEndCapture(s->cap, &graph);
        if (el != hipSuccess) e = el;
        if (e) {
            if (graph) (void)hipGraphDestroy(graph);
            return false;
        }
        s->n_capture++;
        bool updated = false;
        if (g.exec) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(g.exec, graph, &err_node, &res) == hipSuccess &&
                      res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(g.exec);
                g.exec = nullptr;
            }
        }
        if (!updated) {
            e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            if (e) {
                g.exec = nullptr;
                (void)hipGraphDestroy(graph);
                return false;
            }
            s->n_instantiate++;
        }
        (void)hipGraphDestroy(graph);
        g.cl_gen = c.gen;
        g.sc_gen = q.sc_gen;
    }
    return true;
}

void fail_all(pth_server *s, hipError_t e, const char *what)
{
    s->failed = true;
    s->err = std::string("pth_server: ") + what + ": " + hipGetErrorName(e);
    for (Req &r : s->req)
        if (r.state < 2) r.state = 3;
    s->queue.clear();
    s->cv_done.notify_all();
}

void complete(pth_server *s, long long t)
{
    Req &r = s->req[t];
    SCluster &c = *s->cl[r.cluster];
    r.t_done = now_s();
    r.state = 2;
    c.busy = false;
    c.ran = true;
    c.saved_valid = false;
    c.host_valid = false;
    c.last_use = ++s->use_clock;
    s->n_done++;
    s->ints_done += c.n_int;
}

// One plain-mode request: PeTar's call order on the PTH_EXEC_PLAIN context.
int run_plain(pth_server *s, SCluster &c)
{
    static const pth_epj_view ev{sizeof(PEpj), 0, 24, 32, 0, 0, 0, -1};
    static const pth_spj_view sv{sizeof(PSpj), 0, 24, 32, 40, 48, 56, 64, 72};
    static const pth_epi_view iv{sizeof(PEpi), 0, 24, 0};
    static const pth_force_view fv{sizeof(PForce), 0, 24, 32, 8, -1};
    int rc = pth_send_j(s->pctx, c.p_epj.data(), c.n_epj_tot, &ev, c.p_spj.data(), c.n_spj_tot, &sv);
    if (rc) return rc;
    for (int b0 = 0; b0 < c.n_walk; b0 += 200) {
        const int nw = std::min(200, c.n_walk - b0);
        rc = pth_dispatch(s->pctx, 0, nw, c.epi_ptr.data() + b0, c.n_epi.data() + b0, &iv, c.ide_ptr.data() + b0,
                          c.n_epj.data() + b0, c.ids_ptr.data() + b0, c.n_spj.data() + b0);
        if (rc) return rc;
        rc = pth_retrieve(s->pctx, 0, nw, c.n_epi.data() + b0, c.force_ptr.data() + b0, &fv);
        if (rc) return rc;
    }
    return PTH_OK;
}

void worker_plain(pth_server *s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        if (s->queue.empty() || s->failed) {
            s->cv_work.wait(lk);
            continue;
        }
        const long long t = s->queue.front();
        s->queue.pop_front();
        SCluster &c = *s->cl[s->req[t].cluster];
        s->req[t].state = 1;
        c.busy = true;
        lk.unlock();
        const int rc = run_plain(s, c);
        lk.lock();
        if (rc) {
            s->failed = true;
            s->err = std::string("pth_server (plain): ") + pth_last_error(s->pctx);
            s->req[t].state = 3;
            c.busy = false;
            for (Req &r : s->req)
                if (r.state < 2) r.state = 3;
            s->queue.clear();
        } else {
            complete(s, t);
            c.host_valid = true;
        }
        s->cv_done.notify_all();
    }
}

void worker_fused(pth_server *s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e) fail_all(s, e, "hipSetDevice");
    bool compacted = false;  // everything idle was released for a blocked admission
    while (!s->stop) {
        bool progress = false, running = false;
        // completions
        for (int k = 0; k < s->n_streams && !s->failed; k++) {
            Slot &q = s->slot[k];
            if (q.req < 0) continue;
            e = hipEventQuery(q.ev);
            if (e == hipErrorNotReady) {
                running = true;
                continue;
            }
            if (e != hipSuccess) {
                fail_all(s, e, "request execution");
                break;
            }
            complete(s, q.req);
            q.req = -1;
            progress = true;
            s->cv_done.notify_all();
        }
        // starts: queued requests in arrival order, on free streams
        for (int k = 0; k < s->n_streams && !s->failed && !s->queue.empty(); k++) {
            Slot &q = s->slot[k];
            if (q.req >= 0) continue;
            auto it = s->queue.begin();
            bool blocked = false, admitted = false;
            for (; it != s->queue.end(); ++it) {
                const int ck = s->req[*it].cluster;
                if (s->cl[ck]->busy) continue;
                if (!admit(s, ck, k, e, admitted)) {
                    if (e) fail_all(s, e, "admission");
                    blocked = true;
                }
                break;
            }
            if (s->failed) break;
            if (blocked) {
                bool any = false;
                for (const Slot &x : s->slot) any |= x.req >= 0;
                if (!any) {
                    // nothing runs, yet the request does not fit: free every
                    // stream's scratch and evict every cluster once (the pool
                    // is then empty and holds any cluster with its scratch,
                    // which registration checked); failing after that is an
                    // error
                    if (compacted) {
                        fail_all(s, hipErrorOutOfMemory, "admission: the pool cannot hold the cluster");
                        break;
                    }
                    for (Slot &x : s->slot)
                        if (x.sc_off >= 0) {
                            s->pa.release((size_t)x.sc_off, x.sc_size);
                            x.sc_off = -1;
                            x.sc_size = 0;
                        }
                    for (auto &c : s->cl)
                        if (c->off >= 0 && (e = evict(s, *c))) {
                            fail_all(s, e, "eviction");
                            break;
                        }
                    compacted = true;
                    progress = true;
                }
                break;
            }
            if (it == s->queue.end()) break;
            const long long t = *it;
            SCluster &c = *s->cl[s->req[t].cluster];
            if ((e = hipGraphLaunch(c.g[k].exec, q.st)) || (e = hipEventRecord(q.ev, q.st))) {
                fail_all(s, e, "graph launch");
                break;
            }
            s->queue.erase(it);
            s->req[t].state = 1;
            s->req[t].admitted = admitted ? 1 : 0;
            c.busy = true;
            q.req = t;
            running = true;
            progress = true;
            compacted = false;
        }
        if (progress) continue;
        if (running) {
            // requests on the device: poll again soon (a sleep would add tens
            // of microseconds of timer slack to every latency)
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            s->cv_work.wait(lk);
        }
    }
}

void release_cluster(SCluster &c)
{
    for (SGraph &g : c.g)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (c.saved) (void)hipEventDestroy(c.saved);
    c.image.release();
    c.h_accp.release();
    c.h_nnb.release();
}

}  // namespace

extern "C" {

int pth_server_create(pth_ctx *ctx, pth_server **out, int mode, size_t hbm_budget, int n_streams)
{
    if (!ctx) return PTH_ERR_ARG;
    if (!out) return ctx->fail(PTH_ERR_ARG, "pth_server_create: null output");
    *out = nullptr;
    if (mode != PTH_SERVE_FUSED && mode != PTH_SERVE_PLAIN)
        return ctx->fail(PTH_ERR_ARG, "pth_server_create: unknown mode");
    if (ctx->par.changeover) return ctx->fail(PTH_ERR_ARG, "pth_server_create: the changeover mode is not served");
    if (ctx->par.spj_monopole)
        return ctx->fail(PTH_ERR_ARG, "pth_server_create: monopole superparticles are not served");
    if (n_streams < 1 || n_streams > MAX_STREAMS)
        return ctx->fail(PTH_ERR_ARG, "pth_server_create: n_streams outside [1, 16]");
    HIPCHK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    pth_server *s = new pth_server;
    s->ctx = ctx;
    s->mode = mode;
    // on failure: release what was made (no worker thread runs yet)
    auto bail = [&](int code, const std::string &msg) {
        pth_server_free(s);
        return ctx->fail(code, msg);
    };
    auto hip_bail = [&](hipError_t e, const char *what) {
        pth_server_free(s);
        return ctx->hip_fail(e, what);
    };
    hipError_t e;
    if (mode == PTH_SERVE_FUSED) {
        size_t fr = 0, tot = 0;
        if ((e = hipMemGetInfo(&fr, &tot))) return hip_bail(e, "hipMemGetInfo");
        const size_t cap = fr / 10 * 9;
        size_t want = hbm_budget ? hbm_budget : cap / 2;
        if (want > cap) return bail(PTH_ERR_CAPACITY, "pth_server_create: hbm_budget exceeds free device memory");
        want = want & ~size_t(255);
        if (want < 4096) return bail(PTH_ERR_ARG, "pth_server_create: hbm_budget too small");
        if ((e = s->pool.ensure(want))) return hip_bail(e, "server pool allocation");
        s->budget = want;
        s->pa.init(want);
        s->n_streams = n_streams;
        s->slot.assign(n_streams, Slot());
        for (Slot &q : s->slot) {
            if ((e = hipStreamCreateWithFlags(&q.st, hipStreamNonBlocking))) return hip_bail(e, "hipStreamCreate");
            if ((e = hipEventCreateWithFlags(&q.ev, hipEventDisableTiming))) return hip_bail(e, "hipEventCreate");
        }
        if ((e = hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking))) return hip_bail(e, "hipStreamCreate");
        if ((e = hipStreamCreateWithFlags(&s->xfer, hipStreamNonBlocking))) return hip_bail(e, "hipStreamCreate");
        if ((e = hipEventCreateWithFlags(&s->ev_evicted, hipEventDisableTiming))) return hip_bail(e, "hipEventCreate");
        s->th = std::thread(worker_fused, s);
    } else {
        pth_params p = ctx->par;
        p.exec_mode = PTH_EXEC_PLAIN;
        if (pth_create(&s->pctx, ctx->device, &p) != PTH_OK)
            return bail(PTH_ERR_HIP, std::string("pth_server_create: ") + pth_create_error());
        s->n_streams = 1;
        s->th = std::thread(worker_plain, s);
    }
    *out = s;
    return PTH_OK;
}

int pth_server_add_cluster(pth_server *s, int *cluster, int n_walk, const int *n_epi, const int *n_epj,
                           const int *n_spj, const double *epi_pos, const double *epi_rs, const int *id_epj,
                           const int *id_spj, int n_epj_tot, const double *epj_pos, const double *epj_mass,
                           const double *epj_rs, int n_spj_tot, const double *spj_pos, const double *spj_mass,
                           const double *spj_quad)
{
    if (!s) return PTH_ERR_ARG;
    pth_ctx *ctx = s->ctx;
    if (!cluster) return ctx->fail(PTH_ERR_ARG, "pth_server_add_cluster: null output");
    if (n_walk < 0 || n_epj_tot < 0 || n_spj_tot < 0)
        return ctx->fail(PTH_ERR_ARG, "pth_server_add_cluster: negative count");
    long long n_ide = 0, n_ids = 0, ni = 0, n_int = 0;
    for (int w = 0; w < n_walk; w++) {
        if (n_epi[w] < 0 || n_epj[w] < 0 || n_spj[w] < 0)
            return ctx->fail(PTH_ERR_ARG, "pth_server_add_cluster: negative size in walk " + std::to_string(w));
        n_ide += n_epj[w];
        n_ids += n_spj[w];
        ni += n_epi[w];
        n_int += (long long)n_epi[w] * (n_epj[w] + n_spj[w]);
    }
    for (long long k = 0; k < n_ide; k++)
        if (id_epj[k] < 0 || id_epj[k] >= n_epj_tot)
            return ctx->fail(PTH_ERR_ARG, "pth_server_add_cluster: EPJ index out of range");
    for (long long k = 0; k < n_ids; k++)
        if (id_spj[k] < 0 || id_spj[k] >= n_spj_tot)
            return ctx->fail(PTH_ERR_ARG, "pth_server_add_cluster: SPJ index out of range");
    std::unique_ptr<SCluster> c(new SCluster);
    c->n_walk = n_walk;
    c->n_epj_tot = n_epj_tot;
    c->n_spj_tot = n_spj_tot;
    c->ni = ni;
    c->n_int = n_int;
    if (s->mode == PTH_SERVE_FUSED) {
        pth_params par = ctx->par;
        if (!make_plan(n_walk, n_epi, n_epj, n_spj, PTH_EXEC_FUSED, par.chunk_j, par.max_ipl, c->plan))
            return ctx->fail(PTH_ERR_CAPACITY, "pth_server_add_cluster: sizes exceed the int32 index range");
        layout(*c, n_ide, n_ids, df_enabled(par) != 0);
        if (c->L.total + c->S.total > s->budget)
            return ctx->fail(PTH_ERR_CAPACITY, "pth_server_add_cluster: the cluster needs " +
                                                   std::to_string(c->L.total + c->S.total) +
                                                   " bytes with a stream's scratch, more than the budget");
        hipError_t e;
        if ((e = hipSetDevice(ctx->device)) || (e = c->image.ensure(c->L.upload)) ||
            (e = c->h_accp.ensure(sizeof(double4) * (size_t)std::max<long long>(ni, 1))) ||
            (e = c->h_nnb.ensure(sizeof(int) * (size_t)std::max<long long>(ni, 1))) ||
            (e = hipEventCreateWithFlags(&c->saved, hipEventDisableTiming))) {
            release_cluster(*c);
            return ctx->hip_fail(e, "pth_server_add_cluster: pinned image");
        }
        char *h = c->image.as<char>();
        std::memset(h, 0, c->L.upload);
        const Layout &L = c->L;
        EpiDev *he = reinterpret_cast<EpiDev *>(h + L.A.epi);
        for (long long i = 0; i < ni; i++) {
            he[i].x = epi_pos[3 * i];
            he[i].y = epi_pos[3 * i + 1];
            he[i].z = epi_pos[3 * i + 2];
            he[i].rs = (float)epi_rs[i];
            he[i].pad = 0;
        }
        if (n_ide) std::memcpy(h + L.A.ide, id_epj, sizeof(int) * n_ide);
        if (n_ids) std::memcpy(h + L.A.ids, id_spj, sizeof(int) * n_ids);
        copy_plan_to_arena(h, c->plan, L.A);
        EpjDev *ej = reinterpret_cast<EpjDev *>(h + L.epj);
        for (int i = 0; i < n_epj_tot; i++) {
            ej[i].x = epj_pos[3 * (size_t)i];
            ej[i].y = epj_pos[3 * (size_t)i + 1];
            ej[i].z = epj_pos[3 * (size_t)i + 2];
            ej[i].m = (float)epj_mass[i];
            ej[i].rs = (float)epj_rs[i];
        }
        SpjDev *sj = reinterpret_cast<SpjDev *>(h + L.spj);
        for (int i = 0; i < n_spj_tot; i++) {
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
    } else {
        c->n_epi.assign(n_epi, n_epi + n_walk);
        c->n_epj.assign(n_epj, n_epj + n_walk);
        c->n_spj.assign(n_spj, n_spj + n_walk);
        c->p_epj.resize((size_t)std::max(n_epj_tot, 1));
        for (int i = 0; i < n_epj_tot; i++)
            c->p_epj[i] = PEpj{{epj_pos[3 * (size_t)i], epj_pos[3 * (size_t)i + 1], epj_pos[3 * (size_t)i + 2]},
                               epj_mass[i], epj_rs[i]};
        c->p_spj.resize((size_t)std::max(n_spj_tot, 1));
        for (int i = 0; i < n_spj_tot; i++) {
            PSpj &p = c->p_spj[i];
            for (int d = 0; d < 3; d++) p.pos[d] = spj_pos[3 * (size_t)i + d];
            p.mass = spj_mass[i];
            for (int d = 0; d < 6; d++) p.q[d] = spj_quad[6 * (size_t)i + d];
        }
        c->p_epi.resize((size_t)std::max<long long>(ni, 1));
        for (long long i = 0; i < ni; i++)
            c->p_epi[i] = PEpi{{epi_pos[3 * i], epi_pos[3 * i + 1], epi_pos[3 * i + 2]}, epi_rs[i]};
        c->p_force.resize((size_t)std::max<long long>(ni, 1));
        c->ide.assign(id_epj, id_epj + n_ide);
        c->ids.assign(id_spj, id_spj + n_ids);
        long long io = 0, eo = 0, so = 0;
        for (int w = 0; w < n_walk; w++) {
            c->epi_ptr.push_back(c->p_epi.data() + io);
            c->force_ptr.push_back(c->p_force.data() + io);
            c->ide_ptr.push_back(c->ide.data() + eo);
            c->ids_ptr.push_back(c->ids.data() + so);
            io += n_epi[w];
            eo += n_epj[w];
            so += n_spj[w];
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    *cluster = (int)s->cl.size();
    s->cl.push_back(std::move(c));
    return PTH_OK;
}

long long pth_server_footprint(const pth_server *s, int k)
{
    if (!s || k < 0 || k >= (int)s->cl.size()) return PTH_ERR_ARG;
    return s->mode == PTH_SERVE_FUSED ? (long long)s->cl[k]->L.total : 0;
}

int pth_server_submit(pth_server *s, int k, long long *ticket)
{
    if (!s) return PTH_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (k < 0 || k >= (int)s->cl.size()) return s->ctx->fail(PTH_ERR_ARG, "pth_server_submit: unknown cluster");
    if (s->failed) return s->ctx->fail(PTH_ERR_HIP, s->err);
    Req r;
    r.cluster = k;
    r.t_submit = now_s();
    const long long t = (long long)s->req.size();
    s->req.push_back(r);
    s->queue.push_back(t);
    if (ticket) *ticket = t;
    s->cv_work.notify_one();
    return PTH_OK;
}

int pth_server_wait(pth_server *s, long long t, double *latency_ms)
{
    if (!s) return PTH_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (t < 0 || t >= (long long)s->req.size()) return s->ctx->fail(PTH_ERR_ARG, "pth_server_wait: unknown ticket");
    s->cv_done.wait(lk, [&] { return s->req[t].state >= 2; });
    if (s->req[t].state == 3) return s->ctx->fail(PTH_ERR_HIP, s->err);
    if (latency_ms) *latency_ms = 1e3 * (s->req[t].t_done - s->req[t].t_submit);
    return PTH_OK;
}

int pth_server_replay(pth_server *s, int n_req, const int *cluster, const double *arrival_s, double *latency_ms,
                      double *wall_s, int *admitted)
{
    if (!s) return PTH_ERR_ARG;
    if (n_req < 0 || (n_req > 0 && (!cluster || !arrival_s)))
        return s->ctx->fail(PTH_ERR_ARG, "pth_server_replay: bad arguments");
    for (int k = 1; k < n_req; k++)
        if (arrival_s[k] < arrival_s[k - 1])
            return s->ctx->fail(PTH_ERR_ARG, "pth_server_replay: arrival times must not decrease");
    std::vector<long long> t(n_req);
    const double t0 = now_s();
    for (int k = 0; k < n_req; k++) {
        for (;;) {
            const double left = t0 + arrival_s[k] - now_s();
            if (left <= 0.0) break;
            if (left > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(left - 1e-4));
        }
        const int rc = pth_server_submit(s, cluster[k], &t[k]);
        if (rc) return rc;
    }
    double last = t0;
    for (int k = 0; k < n_req; k++) {
        double l = 0.0;
        const int rc = pth_server_wait(s, t[k], &l);
        if (rc) return rc;
        if (latency_ms) latency_ms[k] = l;
        std::lock_guard<std::mutex> lk(s->mu);
        last = std::max(last, s->req[t[k]].t_done);
        if (admitted) admitted[k] = s->req[t[k]].admitted;
    }
    if (wall_s) *wall_s = n_req ? last - (t0 + arrival_s[0]) : 0.0;
    return PTH_OK;
}

int pth_server_results(pth_server *s, int k, double *acc, double *pot, long long *nnb)
{
    if (!s) return PTH_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (k < 0 || k >= (int)s->cl.size()) return s->ctx->fail(PTH_ERR_ARG, "pth_server_results: unknown cluster");
    SCluster &c = *s->cl[k];
    s->cv_done.wait(lk, [&] { return !c.busy || s->failed; });
    if (s->failed) return s->ctx->fail(PTH_ERR_HIP, s->err);
    if (!c.ran) return s->ctx->fail(PTH_ERR_STATE, "pth_server_results: no completed request for this cluster");
    const long long n = c.ni;
    if (s->mode == PTH_SERVE_PLAIN) {
        for (long long i = 0; i < n; i++) {
            const PForce &f = c.p_force[i];
            if (acc) for (int d = 0; d < 3; d++) acc[3 * i + d] = f.acc[d];
            if (pot) pot[i] = f.pot;
            if (nnb) nnb[i] = f.nnb;
        }
        return PTH_OK;
    }
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device), "hipSetDevice");
    if (c.off >= 0) {
        // resident: its last request has completed; copy the outputs now
        // (the worker is held off by the lock meanwhile)
        if (n > 0) {
            const char *base = s->pool.as<char>() + c.off;
            HIPCHK(s->ctx, hipMemcpyAsync(c.h_accp.p, base + c.L.out_accp, sizeof(double4) * n,
                                          hipMemcpyDeviceToHost, s->xfer), "result download");
            HIPCHK(s->ctx, hipMemcpyAsync(c.h_nnb.p, base + c.L.out_nnb, sizeof(int) * n, hipMemcpyDeviceToHost,
                                          s->xfer), "result download");
            HIPCHK(s->ctx, hipStreamSynchronize(s->xfer), "result download");
        }
    } else {
        if (!c.saved_valid) return s->ctx->fail(PTH_ERR_STATE, "pth_server_results: results not kept");
        HIPCHK(s->ctx, hipEventSynchronize(c.saved), "result download");
    }
    const double4 *a = c.h_accp.as<double4>();
    const int *nb = c.h_nnb.as<int>();
    for (long long i = 0; i < n; i++) {
        const double4 v = a[i];
        if (acc) {
            acc[3 * i] = v.x;
            acc[3 * i + 1] = v.y;
            acc[3 * i + 2] = v.z;
        }
        if (pot) pot[i] = v.w;
        if (nnb) nnb[i] = nb[i];
    }
    return PTH_OK;
}

int pth_server_stats(const pth_server *s, long long *n_done, long long *n_admit, long long *n_evict,
                     long long *interactions, long long *pool_peak)
{
    if (!s) return PTH_ERR_ARG;
    std::lock_guard<std::mutex> lk(const_cast<pth_server *>(s)->mu);
    if (n_done) *n_done = s->n_done;
    if (n_admit) *n_admit = s->n_admit;
    if (n_evict) *n_evict = s->n_evict;
    if (interactions) *interactions = s->ints_done;
    if (pool_peak) *pool_peak = (long long)s->pa.peak;
    return PTH_OK;
}

int pth_server_memory(const pth_server *s, long long *resident_total, long long *scratch_max,
                      long long *pinned_host, long long *n_capture, long long *n_instantiate)
{
    if (!s) return PTH_ERR_ARG;
    std::lock_guard<std::mutex> lk(const_cast<pth_server *>(s)->mu);
    long long res = 0, sc = 0, pin = 0;
    for (const auto &c : s->cl) {
        if (s->mode != PTH_SERVE_FUSED) break;
        res += (long long)c->L.total;
        sc = std::max(sc, (long long)c->S.total);
        pin += (long long)(c->image.cap + c->h_accp.cap + c->h_nnb.cap);
    }
    if (resident_total) *resident_total = res;
    if (scratch_max) *scratch_max = sc;
    if (pinned_host) *pinned_host = pin;
    if (n_capture) *n_capture = s->n_capture;
    if (n_instantiate) *n_instantiate = s->n_instantiate;
    return PTH_OK;
}

void pth_server_free(pth_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    if (s->ctx->device >= 0) (void)hipSetDevice(s->ctx->device);
    for (Slot &q : s->slot)
        if (q.st) (void)hipStreamSynchronize(q.st);
    if (s->xfer) (void)hipStreamSynchronize(s->xfer);
    for (auto &c : s->cl) release_cluster(*c);
    for (Slot &q : s->slot) {
        if (q.ev) (void)hipEventDestroy(q.ev);
        if (q.st) (void)hipStreamDestroy(q.st);
    }
    if (s->ev_evicted) (void)hipEventDestroy(s->ev_evicted);
    if (s->cap) (void)hipStreamDestroy(s->cap);
    if (s->xfer) (void)hipStreamDestroy(s->xfer);
    s->pool.release();
    if (s->pctx) pth_destroy(s->pctx);
    delete s;
}

}  // extern "C"
