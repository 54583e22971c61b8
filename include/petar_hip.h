/*
 * petar_hip.h -- C ABI of libpetar_hip.so, the MI355X (gfx950) soft-force
 * path of PeTar: FDPS multi-walk tree force with linear cutoff, i.e.
 * particle-particle EPJ terms, quadrupole SPJ terms and the neighbour count
 * within r_search.
 *
 * Entry points and the reference interface each one replaces:
 *
 *   pth_create / pth_destroy
 *       first-call initialisation of CalcForceWithLinearCutoffCUDAMultiWalk
 *       (src/force_gpu_cuda.cu:547-571: device = rank % n_devices, fixed
 *       buffers) and the file-static buffers of src/force_gpu_cuda.cu:513-533.
 *   pth_send_j
 *       CalcForceWithLinearCutoffCUDAMultiWalk::operator() with
 *       send_flag == true (src/force_gpu_cuda.cu:577-620; declared at
 *       src/force_gpu_cuda.hpp:120-132): all EPJ and SPJ of the force call.
 *   pth_dispatch
 *       the same operator() with send_flag == false
 *       (src/force_gpu_cuda.cu:621-699): one batch of walks with their
 *       index lists; returns before the kernels finish.
 *   pth_dispatch_copies
 *       CalcForceWithLinearCutoffCUDA::operator() of the build without
 *       PARTICLE_SIMULATOR_GPU_MULIT_WALK_INDEX (src/force_gpu_cuda.cu:704-816):
 *       a batch of walks with per-walk copies of their EPJ and SPJ records.
 *   pth_retrieve
 *       RetrieveForceCUDA (src/force_gpu_cuda.cu:831-880; declared at
 *       src/force_gpu_cuda.hpp:165-168): waits, then ASSIGNS acc, pot and
 *       n_ngb of every i of the batch in walk order.
 *   pth_last_error
 *       no counterpart (the reference asserts and ignores CUDA errors,
 *       src/cuda_pointer.h:7); text of the last failure of a context.
 *   pth_resident_*, pth_server_*, pth_plan_debug
 *       no counterpart: a force call kept resident in HBM, the serving loop
 *       over many independent evaluations, and a host-only view of the work
 *       decomposition.
 *
 * pth_params.changeover = 1 also folds PeTar's changeover correction
 * (src/hard.hpp:1404-1476,1655-1691, src/changeover.hpp:294-375) into the
 * kernel; see below.
 *
 * Records are read and written in place through "views" (stride and byte
 * offsets of the members FDPS's EPISoft / EPJSoft / SPJSoft / ForceSoft carry,
 * src/soft_ptcl.hpp:4-24, 271-376), so the caller's arrays are used as they
 * are; host pointers are only read during the call (as FDPS guarantees).
 *
 * Results follow the reference GPU kernels (src/force_gpu_cuda.cu:50-77,
 * 276-323): acc = G sum m_j (x_j - x_i) / max(r^2+eps^2, r_out^2)^(3/2) over
 * EPJ plus the quadrupole SPJ terms, pot likewise (including the self term
 * -G m_i / r_out), n_ngb counts j (self included) with
 * r^2 + eps^2 < max(rs_i, rs_j)^2 (nnb_mode 0) or <= (nnb_mode 1, the x86
 * kernel's convention).  Pair arithmetic is FP32 on coordinates relative to
 * the walk's centre (computed in FP64); sums are FP64.
 *
 * Every function returns PTH_OK (0) or a negative PTH_ERR_* code.
 */
#ifndef PETAR_HIP_H
#define PETAR_HIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PTH_OK 0
#define PTH_ERR_ARG (-1)
#define PTH_ERR_CAPACITY (-2)
#define PTH_ERR_HIP (-3)
#define PTH_ERR_STATE (-4)
#define PTH_ERR_NOMEM (-5)

/* execution forms */
#define PTH_EXEC_FUSED 0 /* work items over all walks, long lists split, XCD-aware order */
#define PTH_EXEC_PLAIN 1 /* one workgroup per walk, one launch per batch, synchronous copies */

typedef struct pth_ctx pth_ctx;

typedef struct pth_params {
    double eps2;   /* softening^2 (EPISoft::eps^2) */
    double rcut2;  /* r_out^2 of the linear cutoff (EPISoft::r_out^2) */
    double G;      /* ForceSoft::grav_const */
    int nnb_mode;  /* 0: r2+eps2 < rs^2 (GPU reference), 1: <= (x86 reference) */
    int exec_mode; /* PTH_EXEC_FUSED or PTH_EXEC_PLAIN */
    int chunk_j;   /* fused form: max j-entries per work item (0 = default, 1024) */
    int max_walk;  /* largest n_walk accepted by pth_dispatch (0 = 1000, the
                      reference's N_WALK_LIMIT, src/force_gpu_cuda.cu:14) */
    int max_ipl;   /* fused form: most i-particles per lane, 1 to 4 (tiles of 64, 128,
                      192, 256 i; other values: 4) */
    int j_vgpr;    /* fused form: 1 copies each j record into VGPRs before use,
                      0 uses it as SGPR operands */
    int changeover; /* 0: linear cutoff only (the reference GPU kernels);
                       1: changeover mode (fused form only), below */
    double ngb_factor; /* changeover mode: neighbour radius factor on r_search
                          (0 = 0.99, SAFTY_FACTOR_FOR_SEARCH of src/ptcl.hpp:8) */
    int segments;  /* resident evaluations: walks split into this many segments
                      of equal cost (1..32; the first of half the cost) and
                      pipelined: one stream stages
                      the segments in turn, the others run each segment's
                      force and reduce kernels once it is staged, so staging
                      overlaps the force kernels; results do not depend on
                      it.  pth_dispatch always uses one. */
    int streams;   /* resident evaluations: streams of the pipeline (1..8):
                      one staging, streams - 1 force streams */
    int df_near;   /* fused form, linear cutoff: 1 computes the separation of
                      every EPJ pair of a walk's near region (pairs that may
                      lie within max(r_search, r_out) of each other) from
                      double-float walk-relative coordinates (the FP32
                      coordinates plus their staged low-order parts), so a
                      tight pair keeps FP32 relative precision however far it
                      lies from the walk centre; 0 uses the FP32 coordinates
                      alone.  Changeover mode always does this for its
                      neighbour pairs. */
    int spj_monopole; /* 0: quadrupole superparticles (SPJSoft =
                      PS::SPJQuadrupoleInAndOut, PeTar built with USE_QUAD);
                      1: monopole superparticles (PS::SPJMonopoleInAndOut, a
                      build configured with --disable-quad): mass and centre
                      only, the pair of src/soft_force.hpp:125-158 (no cutoff,
                      no neighbour count); the quadrupole members of the SPJ
                      view and arrays are not read */
    int graph;     /* resident evaluations: 1 captures an evaluation's launches
                      (all segments and streams) once as a hipGraph and
                      replays it on later runs; 0 launches them one by one.
                      Results do not depend on it. */
    double min_seg_pairs; /* resident evaluations: fewest pair interactions a
                      pipelined segment gets (0 = 1e9), so a small evaluation
                      runs in fewer segments, down to one on one stream
                      (launches and cross-stream events cost more there than
                      the overlap gains).  Results do not depend on it. */
    int persistent; /* fused form (pth_dispatch and resident evaluations): 1
                      launches each i-per-lane class as one wave per free
                      wave slot, the waves of each XCD taking the plan's
                      items of that XCD from a counter in plan order; 0
                      launches one wave per item.  Results do not depend on
                      it. */
} pth_params;

/* Changeover mode folds PeTar's changeover correction of the soft force
 * (src/hard.hpp:1404-1476 per neighbour, 1655-1691 per particle, with the
 * changeover functions of src/changeover.hpp:294-375) into the force kernel:
 * for every EPJ pair that is a neighbour pair, r^2 < (ngb_factor *
 * max(rs_i, rs_j))^2 (FDPS's symmetric neighbour search on getRSearch(),
 * src/soft_ptcl.hpp:97-101,349-353), j != i, the linear-cutoff terms are
 * replaced by G m (1 - W_acc(r)) dx / r^3 and -G m (1 - W_pot(r)) / r (pot)
 * and -G m / r (pot_tot), W taken from the pair member with the larger r_out;
 * the self term -G m_i / r_out is dropped.  With the EPJ group data
 * (GroupDataDeliver::artificial) the particle types follow
 * src/hard.hpp:1453-1472 and 1669: a group member j (status < 0, zero tree
 * mass) adds its backup mass's -G m_b (1 - W_pot)/r and -G m_b/r to the
 * potentials, an artificial or c.m. particle j (status > 0) leaves no
 * potential (acc as a single in every case), and only singles and members
 * without a c.m. address (status -LARGE_FLOAT, taken as below -1e30) lose
 * their self term.  Requires EPJ ids and r_in/r_out (epj view), EPI ids
 * (epi view), each i's own EPJ in its walk's list, and at most 1024 i per
 * walk.  pot receives pot_soft; pot_tot goes to force view member
 * off_pot_tot when that is >= 0. */

/* Fills the defaults: eps2 0, rcut2 0, G 1, nnb_mode 0, fused, chunk 0,
 * max_walk 1000, max_ipl 2, j_vgpr 0, changeover 0, ngb_factor 0.99,
 * segments 8, streams 3, df_near 1, spj_monopole 0, graph 0, min_seg_pairs 0
 * (1e9), persistent 0. */
void pth_params_default(pth_params *p);

/* Creates a context on device (rank % number of visible devices). */
int pth_create(pth_ctx **ctx, int rank, const pth_params *params);
/* Replaces every parameter of a context (the reference functor receives its
 * parameters at every construction, src/force_gpu_cuda.hpp:108-118).  Values
 * are read when a batch is dispatched, so batches already in flight keep the
 * values they were dispatched with; turning changeover on takes effect from
 * the next pth_send_j.  To change a few fields, read the others with
 * pth_get_params first. */
int pth_set_params(pth_ctx *ctx, const pth_params *params);
/* The context's current parameters (after the clamping pth_create and
 * pth_set_params apply). */
int pth_get_params(const pth_ctx *ctx, pth_params *params);
/* ---- profile of the FDPS hooks -------------------------------------------
 * What a reference GPU build compiled with GPU_PROFILE collects
 * (GPUProfile / GPUCounter, src/force_gpu_cuda.hpp:6-93, filled at
 * src/force_gpu_cuda.cu:578-697,836-863 and written to the profile output by
 * src/petar.hpp:2081-2155): times of the host packing, the copies and the
 * kernels of pth_send_j / pth_dispatch / pth_retrieve, and the batch
 * counts. */
typedef struct pth_profile {
    double copy_s;  /* host wall time packing records (send and dispatch) and
                       assigning results (retrieve) */
    double send_s;  /* host-to-device copies of send and dispatch (HIP events) */
    double calc_s;  /* the kernels of each batch (HIP events) */
    double recv_s;  /* the device-to-host copy of each batch's results */
    long long n_walk, n_epi, n_epj, n_spj, n_call;  /* dispatched walks, i,
                       EPJ and SPJ list entries, dispatch calls */
} pth_profile;
/* Turns the profile on (on != 0: timing events are recorded from the next
 * call) or off. */
int pth_profile_enable(pth_ctx *ctx, int on);
/* Returns what was accumulated since the last take (device times of
 * batches count once they have been retrieved) and starts again from
 * zero. */
int pth_profile_take(pth_ctx *ctx, pth_profile *out);

/* Text of the last pth_create failure of the calling thread. */
const char *pth_create_error(void);
/* Frees the context.  Its resident evaluations, trees and servers use its
 * device and stream: free them first (pth_resident_free, pth_gtree_free,
 * pth_server_free). */
void pth_destroy(pth_ctx *ctx);
const char *pth_last_error(const pth_ctx *ctx);
/* Device index the context runs on, or -1. */
int pth_device(const pth_ctx *ctx);

/* ---- record views (byte offsets into the caller's structs) ---------------- */
typedef struct pth_epj_view { /* EPJSoft */
    size_t stride, off_pos, off_mass, off_rsearch;
    size_t off_id, off_rin, off_rout; /* changeover mode only */
    long long off_group; /* changeover mode: EPJSoft::group_data, whose
                            artificial member holds mass_backup then status
                            (two doubles, src/artificial_particles.hpp:20-23);
                            -1: every particle is a single */
} pth_epj_view;
typedef struct pth_spj_view { /* SPJQuadrupoleInAndOut */
    size_t stride, off_pos, off_mass, off_qxx, off_qyy, off_qzz, off_qxy, off_qxz, off_qyz;
} pth_spj_view;
typedef struct pth_epi_view { /* EPISoft */
    size_t stride, off_pos, off_rsearch;
    size_t off_id; /* changeover mode only */
} pth_epi_view;
typedef struct pth_force_view { /* ForceSoft */
    size_t stride, off_acc, off_pot, off_nngb;
    int nngb_bytes;       /* 4 or 8 */
    long long off_pot_tot; /* changeover mode: pot_tot member, or -1 */
} pth_force_view;
/* pos members are three consecutive doubles (x, y, z), mass/rsearch/r_in/
 * r_out/q* are doubles, ids are 64-bit integers. */

/* ---- FDPS multi-walk hooks ------------------------------------------------- */
int pth_send_j(pth_ctx *ctx, const void *epj, int n_epj_tot, const pth_epj_view *ev,
               const void *spj, int n_spj_tot, const pth_spj_view *sv);

int pth_dispatch(pth_ctx *ctx, int tag, int n_walk, const void *const *epi, const int *n_epi,
                 const pth_epi_view *iv, const int *const *id_epj, const int *n_epj,
                 const int *const *id_spj, const int *n_spj);

int pth_retrieve(pth_ctx *ctx, int tag, int n_walk, const int *ni, void *const *force,
                 const pth_force_view *fv);

/* The non-index multi-walk functor CalcForceWithLinearCutoffCUDA::operator()
 * (src/force_gpu_cuda.cu:704-816; declared at src/force_gpu_cuda.hpp:137-162;
 * used by src/petar.hpp:901 when PARTICLE_SIMULATOR_GPU_MULIT_WALK_INDEX is
 * not defined): one batch of walks, each with its own copies of its EPJ and
 * SPJ records (epj[w][0 .. n_epj[w]), spj[w][0 .. n_spj[w])).  The records
 * become the context's j arrays for this batch (no pth_send_j), then the
 * batch runs as pth_dispatch runs it; pth_retrieve returns its results. */
int pth_dispatch_copies(pth_ctx *ctx, int tag, int n_walk, const void *const *epi, const int *n_epi,
                        const pth_epi_view *iv, const void *const *epj, const int *n_epj, const pth_epj_view *ev,
                        const void *const *spj, const int *n_spj, const pth_spj_view *sv);

/* ---- device-resident evaluation (inputs kept in HBM across calls) --------- */
typedef struct pth_resident pth_resident;

/* Uploads one complete soft-force evaluation: EPI (concatenated per walk),
 * walk sizes, concatenated index lists and the EPJ/SPJ arrays they index.
 * Arrays are plain FP64 ([n][3] positions, [n][6] quadrupoles xx yy zz xy
 * xz yz). */
int pth_resident_create(pth_ctx *ctx, pth_resident **res, int n_walk, const int *n_epi,
                        const int *n_epj, const int *n_spj, const double *epi_pos,
                        const double *epi_rs, const int *id_epj, const int *id_spj,
                        int n_epj_tot, const double *epj_pos, const double *epj_mass,
                        const double *epj_rs, int n_spj_tot, const double *spj_pos,
                        const double *spj_mass, const double *spj_quad);
/* Enqueues one evaluation on the context's stream (asynchronous). */
int pth_resident_run(pth_resident *res);
/* Waits for the last evaluation. */
int pth_resident_sync(pth_resident *res);
/* Changeover mode: the ids of the EPI (concatenated per walk) and the ids,
 * changeover radii and group data ([n][2]: mass_backup, status; NULL: every
 * particle a single) of the EPJ; must be given before the first run. */
int pth_resident_changeover(pth_resident *res, const long long *epi_id, const long long *epj_id,
                            const double *epj_rin, const double *epj_rout, const double *epj_group);
/* Copies results (waits first): acc [ni][3], pot [ni], nnb [ni]; pot_tot
 * [ni] in changeover mode (else left alone).  Any pointer may be NULL. */
int pth_resident_results(pth_resident *res, double *acc, double *pot, long long *nnb, double *pot_tot);
/* Per-kernel GPU times of the last evaluation (ms), from HIP events on the
 * context's stream: stage (centres + gather), force, reduce.  With more than
 * one segment the phases overlap, and a graph replay (pth_params.graph) is
 * timed as a whole: ms_force is then the whole evaluation and ms_stage =
 * ms_reduce = 0. */
int pth_resident_timing(pth_resident *res, float *ms_stage, float *ms_force, float *ms_reduce);
/* Re-plans the work decomposition (fused form) with another i-per-lane
 * bound and chunk size; the uploaded inputs stay. */
int pth_resident_replan(pth_resident *res, int max_ipl, int chunk_j);
/* Autotuner: times (median of reps evaluations, HIP events) every pair of
 * candidate max_ipl x chunk_j on this evaluation and keeps the fastest plan;
 * returns the choice and its time. */
int pth_resident_autotune(pth_resident *res, const int *ipl, int n_ipl, const int *chunk, int n_chunk,
                          int reps, int *best_ipl, int *best_chunk, float *best_ms);
/* Sizes: interactions N_e, N_s; work items; staged bytes. */
int pth_resident_stats(const pth_resident *res, long long *n_epep, long long *n_epsp,
                       long long *n_items, long long *staged_bytes);
void pth_resident_free(pth_resident *res);

/* ---- one evaluation without allocating ----------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the work of
 * pth_resident_create + pth_resident_run + pth_resident_results (same
 * arguments, same kernels and work decomposition, the same results bit for
 * bit) with every device buffer it needs -- packed records, plan, staged
 * lists, partial sums, outputs -- taken from one device array of
 * PTH_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * from module load to the end of the process.  The call allocates no device
 * or pinned memory; it waits for work already on the context's stream,
 * enqueues the uploads (from the caller's host arrays), the kernels and the
 * downloads there, and returns when the results are in acc [ni][3], pot [ni]
 * and nnb [ni] (any of them may be NULL).  Linear-cutoff mode and the fused
 * form only (PTH_ERR_STATE otherwise).  An evaluation whose buffers exceed
 * the array fails with PTH_ERR_CAPACITY and a message giving the size it
 * needs (a 16,384-particle Plummer cluster at theta 0.3, r_out 0.01 needs
 * about 38 MiB, most of it staged lists and partial sums; clusters of up
 * to about 38,000 particles fit).  Every call on
 * one GPU uses the same array: calls that overlap on one GPU (from several
 * threads or contexts) race and must be serialised by the caller. */
#define PTH_NOALLOC_SCRATCH_BYTES (128ll << 20)
int pth_resident_eval_noalloc(pth_ctx *ctx, int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                              const double *epi_pos, const double *epi_rs, const int *id_epj, const int *id_spj,
                              int n_epj_tot, const double *epj_pos, const double *epj_mass, const double *epj_rs,
                              int n_spj_tot, const double *spj_pos, const double *spj_mass, const double *spj_quad,
                              double *acc, double *pot, long long *nnb);

/* ---- serving: many independent evaluations ("clusters") -------------------
 * No counterpart in the reference (one simulation per process there).  A
 * server holds any number of clusters, each one complete soft-force
 * evaluation (the inputs of pth_resident_create); a request asks for one
 * evaluation of one cluster.
 *
 * PTH_SERVE_FUSED: clusters are admitted into one HBM pool of hbm_budget
 *   bytes when their footprint (inputs and outputs) fits, copied from pinned
 *   host images, and stay resident; idle clusters are evicted, least recently
 *   used first, to make room (their results are copied to pinned host memory
 *   asynchronously).  What a request needs only while it runs (staged lists,
 *   partial sums) is a per-stream scratch range from the same pool, sized
 *   for the largest cluster the stream has run.  A worker thread starts each
 *   request as one hipGraph launch (stage + force + reduce) on one of
 *   n_streams streams, so requests that arrive while others run overlap on
 *   the device; a cluster keeps one graph per stream and updates it in place
 *   when its range or the scratch moved.  A cluster runs one request at a
 *   time; requests start in arrival order as streams free up.  Each
 *   cluster's results do not depend on which other requests run beside it
 *   (bit-identical).
 * PTH_SERVE_PLAIN: the plain per-walk form as PeTar would drive it per
 *   request: all EPJ/SPJ sent, then batches of at most 200 walks dispatched
 *   and retrieved with synchronous copies (PTH_EXEC_PLAIN), one request at a
 *   time.  The comparison baseline.
 * Latency is measured from submission to results ready in device memory
 * (fused) or in host memory (plain). */
typedef struct pth_server pth_server;
#define PTH_SERVE_FUSED 0
#define PTH_SERVE_PLAIN 1

int pth_server_create(pth_ctx *ctx, pth_server **srv, int mode, size_t hbm_budget, int n_streams);
/* Registers a cluster (arguments as pth_resident_create); its image goes to
 * pinned host memory, nothing is sent to the device until its first request
 * is admitted.  Fails with PTH_ERR_CAPACITY if its footprint and its scratch
 * exceed the whole budget. */
int pth_server_add_cluster(pth_server *srv, int *cluster, int n_walk, const int *n_epi, const int *n_epj,
                           const int *n_spj, const double *epi_pos, const double *epi_rs, const int *id_epj,
                           const int *id_spj, int n_epj_tot, const double *epj_pos, const double *epj_mass,
                           const double *epj_rs, int n_spj_tot, const double *spj_pos, const double *spj_mass,
                           const double *spj_quad);
/* Device bytes a cluster takes while resident (inputs and outputs; its
 * requests also need a stream's scratch, see pth_server_memory). */
long long pth_server_footprint(const pth_server *srv, int cluster);
/* Fused server memory: the sum of the registered clusters' footprints, the
 * largest per-stream scratch a cluster needs, pinned host bytes (images and
 * saved results), and the request graphs captured and instantiated so far.
 * Any pointer may be NULL. */
int pth_server_memory(const pth_server *srv, long long *resident_total, long long *scratch_max,
                      long long *pinned_host, long long *n_capture, long long *n_instantiate);
/* Queues one evaluation of a cluster (returns at once). */
int pth_server_submit(pth_server *srv, int cluster, long long *ticket);
/* Waits for a request; latency_ms (may be NULL) from submission to done. */
int pth_server_wait(pth_server *srv, long long ticket, double *latency_ms);
/* Load generator: submits request k for cluster[k] at arrival_s[k] seconds
 * after the call (non-decreasing), waits for all, and returns each latency,
 * whether the request had to upload its cluster (admitted[k], may be NULL)
 * and the wall time from the first arrival to the last completion. */
int pth_server_replay(pth_server *srv, int n_req, const int *cluster, const double *arrival_s,
                      double *latency_ms, double *wall_s, int *admitted);
/* Results of the cluster's last completed request (waits for one in
 * flight): acc [ni][3], pot [ni], nnb [ni]; any pointer may be NULL. */
int pth_server_results(pth_server *srv, int cluster, double *acc, double *pot, long long *nnb);
/* Counters: completed requests, admissions, evictions, interactions
 * completed, peak pool bytes in use. */
int pth_server_stats(const pth_server *srv, long long *n_done, long long *n_admit, long long *n_evict,
                     long long *interactions, long long *pool_peak);
void pth_server_free(pth_server *srv);

/* ---- tree and interaction lists on the device -------------------------------
 * No counterpart in the reference GPU code (FDPS builds the tree and the lists
 * on the CPU, src/petar.hpp:619-623,2741,894-899).  Builds the octree
 * (Morton order, cells split above n_leaf_limit particles), its moments
 * (mass, centre of mass, raw second moment, particle box, box widened by
 * r_search), the i-groups (largest cells with at most n_group_limit
 * particles) and every group's EPJ / SPJ lists (theta opening with the
 * symmetric r_search test) with the semantics of pth_tree_build +
 * pth_tree_walk of include/petar_host.h: the same nodes, moments, groups
 * and, per group, the same list entries.  With canonical != 0 every list is
 * put in the host walk's order (EPJ entries ascending, SPJ entries by the
 * cell's first particle: a depth-first walk's order), so the lists equal the
 * host lists entry for entry; with 0 the order inside a list is the walk's
 * batch order.  Positions, masses and r_search (may be NULL: 0) are host
 * arrays [n][3], [n], [n]. */
typedef struct pth_gtree pth_gtree;
int pth_gtree_build(pth_ctx *ctx, pth_gtree **tree, long long n, const double *pos, const double *mass,
                    const double *r_search, int n_leaf_limit, int n_group_limit, double theta, int canonical);
/* Rebuilds the tree for new particles (same limits, theta and ordering),
 * reusing its device memory (grown when n or the lists grow).  With
 * on_device != 0 the arrays are device (HBM) arrays of this process holding
 * n entries, so a step whose particles are already resident needs no host
 * copy; every range read is checked to lie inside one device allocation
 * (host pointers are rejected) before any launch. */
int pth_gtree_update(pth_gtree *tree, long long n, const double *pos, const double *mass, const double *r_search,
                     int on_device);
int pth_gtree_sizes(const pth_gtree *tree, long long *n_node, long long *n_group, long long *n_epj_list,
                    long long *n_spj_list);
/* Copies the tree to the host (any pointer may be NULL): order [n] (sorted
 * position -> input index), nodes [n_node] (begin, end, level, mass, centre
 * of mass [3], raw second moment [6]: xx yy zz xy xz yz), groups [n_group]
 * (first particle, size, list lengths) and the concatenated lists (EPJ
 * entries are sorted positions, SPJ entries node indices). */
int pth_gtree_download(const pth_gtree *tree, long long *order, long long *node_begin, long long *node_end,
                       int *node_level, double *node_mass, double *node_com, double *node_quad,
                       long long *epi_begin, int *n_epi, int *n_epj, int *n_spj, int *id_epj, int *id_spj);
/* Host wall time of the last build (ms): tree (upload, sort, nodes, moments,
 * groups), walk (first walk pass, scan, list sizes) and the rest (writing
 * pass or the move of the lists written by the first pass, canonical
 * ordering when asked for, and meanwhile on the host the plan of a resident
 * evaluation under the context's parameters, which pth_resident_from_gtree /
 * pth_resident_refill then take over). */
int pth_gtree_timing(const pth_gtree *tree, float *ms_build, float *ms_walk, float *ms_order);
/* Walk passes of the last build (0 without groups): 1 when the lists were
 * written in one pass into regions sized from the previous build's list
 * lengths (a rebuild with the same number of groups, every list fitting),
 * 2 when a counting or padded pass was followed by an exact writing pass;
 * tile_walk_passes: the same for the candidate walk of the last neighbour
 * search (regions from the previous search's candidate counts, when it had
 * as many tiles); filter_passes: the same for its exact filter (regions
 * from the previous completed search's per-particle counts, same N).  No
 * counterpart in the reference (FDPS walks on the host); for callers that
 * monitor rebuild costs.  Any pointer may be null. */
int pth_gtree_stats(const pth_gtree *tree, int *walk_passes, int *tile_walk_passes, int *filter_passes);
void pth_gtree_free(pth_gtree *tree);
/* Neighbour search on the tree (replaces the FDPS short-range tree search
 * PeTar runs with SearchNeighborEpEpNoSimd / SearchNeighborEpEpSimd,
 * src/petar.hpp:767-788, src/soft_force.hpp:11-34,239-283, and the
 * getNeighborListOneParticle lists its cluster search reads,
 * src/cluster_list.hpp:461-493): for every particle i the particles j with
 * |x_i - x_j|^2 < max(r_search_i, r_search_j)^2 in FP64, self included.
 * Candidates come from one symmetric r_search walk per tile (the largest
 * cells with at most 64 particles, boxes widened beyond any rounding), so
 * the rows are every such j, in ascending tree order.  Runs on the device
 * and keeps the lists there; n_total (may be NULL) receives the number of
 * entries. */
int pth_gtree_neighbors(pth_gtree *tree, long long *n_total);
/* Copies the last search's lists, indexed by input particle index: count
 * [n], offset [n + 1] (row i is ngb[offset[i] .. offset[i + 1])), ngb
 * [n_total] (input indices, in ascending tree order); ms (may be NULL)
 * receives the search's host wall time.  Any pointer may be NULL. */
int pth_gtree_neighbors_download(const pth_gtree *tree, int *count, long long *offset, int *ngb, float *ms);
/* A resident evaluation (as pth_resident_create) fed from the tree without a
 * host round trip: walk w = group w, EPI = the particles in tree order, EPJ
 * records = the particles in tree order (FP64 position, mass, r_search),
 * SPJ records = the nodes (centre of mass, mass, raw second moment).  The
 * tree may be freed afterwards.  In changeover mode call
 * pth_resident_changeover with arrays in tree order (pth_gtree_download's
 * order maps them).  With a canonical tree the results equal those of
 * pth_resident_create on the host tree's lists bit for bit. */
int pth_resident_from_gtree(pth_ctx *ctx, pth_resident **res, const pth_gtree *tree);
/* Refills an existing resident evaluation from a (re)built tree, reusing its
 * device memory: the per-step form of pth_resident_from_gtree.  In
 * changeover mode pth_resident_changeover must be called again. */
int pth_resident_refill(pth_resident *res, const pth_gtree *tree);

/* ---- diagnostics (host only, no device needed) ----------------------------- */
/* The work decomposition pth_dispatch / pth_resident_create use for these walk
 * sizes: returns the number of work items in launch order (or a negative
 * code).  items (8 ints each: walk, i_begin, n_i | ipl << 16, e_blk_begin,
 * e_blk_end, s_begin, s_end, part; walk < 0 marks an empty slot), the staged
 * offsets es_off / ss_off (n_walk + 1 each), the reduction items (8 ints
 * each: out, n_i, part_first, n_part, ipl, 0, 0, 0) and the segments (12
 * ints for each of 32: walk_begin, walk_end, item begin of ipl 1/2/3/4, item
 * count of ipl 1/2/3/4, ritem_begin, ritem_count; -1 past the last) are
 * written when the pointers are not NULL. */
long long pth_plan_debug(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj,
                         int exec_mode, int chunk_j, int max_ipl, int n_seg, int *items,
                         long long max_items, long long *es_off, long long *ss_off, int *ritems,
                         long long max_ritems, long long *n_ritems, int *segs);

#ifdef __cplusplus
}
#endif
#endif
