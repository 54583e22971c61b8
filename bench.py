"""Benchmark of the soft-force path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5]

A step is one complete soft-force evaluation (all walks of the force call:
staging of the lists + force kernel + reduction of split walks) over inputs
already resident in HBM.  The default workload is the largest single-GPU
configuration: Plummer N = 8,388,608 with 100% primordial binaries (Kroupa
masses, log-uniform a in [1e-5, 1e-2], thermal e), theta = 0.3,
n_leaf_limit 20, n_group_limit 512, quadrupole SPJ, r_out = 0.003,
r_in = 0.0003, r_search = 0.0045, G = 1, eps = 0.  The serving object
always uses the 256-cluster stream (r_out = 0.01).

Metric: pair interactions per second, I = N_e + N_s = sum over walks of
n_i (n_epj + n_spj).  Flop convention (reference formulas, FMA = 2, rsqrt =
max = compare = 1): F = 26 N_e + 65 N_s.

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P32_PEAK_TFLOPS = 157.3   # MI355X FP32 peak (vector = matrix)
HBM_PEAK_GBS = 8000.0
FLOP_EPJ, FLOP_SPJ = 26, 65

# walk segments, pipeline streams and GPU_MAX_HW_QUEUES per workload (0:
# HIP's default): the large evaluations gain from a third force stream given
# queues enough for it (8.4M: 89.0 vs 89.9 ms per step, 4.2M: 37.9 vs 38.9);
# the 1M one runs best on the library's defaults (with 8 queues 10.86-11.18
# ms, with 4 10.84-10.90 ms, either setting)
PIPELINE = {"c3": (16, 4, 8), "c5": (16, 4, 8)}

WORKLOADS = {
    # name: (description, generator kwargs, r_out)
    "c1": ("Plummer N=16384, Kroupa IMF, no binaries, theta=0.3, quad SPJ, r_out=0.01",
           dict(n=16384, model="plummer", f_bin=0.0, seed=1), 0.01),
    "c2": ("Plummer N=1048576, Kroupa IMF, 50% primordial binaries, theta=0.3, quad SPJ, r_out=0.01",
           dict(n=1048576, model="plummer", f_bin=0.5, seed=2), 0.01),
    "c3": ("King W0=7 N=4194304, Kroupa IMF, mass-segregated, theta=0.3, quad SPJ, r_out=0.01",
           dict(n=4194304, model="king", f_bin=0.0, seed=3, segregate=True), 0.01),
    "c5": ("Plummer N=8388608, Kroupa IMF, 100% primordial binaries, theta=0.3, quad SPJ, r_out=0.003",
           dict(n=8388608, model="plummer", f_bin=1.0, seed=5), 0.003),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name, nthreads=0):
    from petar_amd import host
    desc, kw, r_out = WORKLOADS[name]
    t0 = time.time()
    kw = dict(kw)
    n = kw.pop("n")
    model = kw.pop("model")
    m, x, _ = host.cluster(n, model, kroupa=True, **kw)
    t1 = time.time()
    L = host.build_lists(x, m, 1.5 * r_out, theta=0.3, n_leaf_limit=20, n_group_limit=512, nthreads=nthreads)
    t2 = time.time()
    log(f"[bench] workload {name}: {desc}; ic {t1 - t0:.1f}s, tree+walk {t2 - t1:.1f}s, {L.n_walk} walks")
    return L, r_out, desc, dict(ic_s=round(t1 - t0, 2), tree_walk_s=round(t2 - t1, 2)), (m, x)


def device_tree_step(ctx, particles, r_out, L, ref, reps=5):
    """The whole soft-force step with the tree on the device, in its
    steady-state form (tree and resident evaluation rebuilt in place each
    step): octree + moments + groups, group walk, canonical list order,
    resident records and plan, force evaluation.  Timed with the particles
    already in HBM (torch tensors) and, separately, uploaded from host
    arrays; the forces are checked bit for bit against the evaluation fed
    from the host lists."""
    import torch
    from petar_amd import soft_force as sf
    m, x = particles
    rs = np.full(m.size, 1.5 * r_out)
    dev = torch.device("cuda:0")
    tx, tm, tr = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (x, m, rs))
    torch.cuda.synchronize()
    g = sf.GpuTree(ctx, x, m, rs, theta=0.3, n_leaf_limit=20, n_group_limit=512, canonical=True)
    r = sf.Resident.from_gtree(ctx, g)
    r.run()
    r.sync()
    z = g.sizes()
    assert z["n_group"] == L.n_walk and z["n_epj_list"] == L.id_epj.size and z["n_spj_list"] == L.id_spj.size
    acc, pot, nnb = r.results()
    same = bool(np.array_equal(acc, ref[0]) and np.array_equal(pot, ref[1]) and np.array_equal(nnb, ref[2]))
    res = {}
    passes = []
    for label, inputs in (("device_inputs", (tx, tm, tr)), ("host_inputs", (x, m, rs))):
        phases, walls = [], []
        for k in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.update(*inputs)
            t1 = time.perf_counter()
            r.refill(g)
            t2 = time.perf_counter()
            r.run()
            r.sync()
            t3 = time.perf_counter()
            if k > 0:
                tb, tw, to = g.timing()
                phases.append((tb, tw, to, (t2 - t1) * 1e3, (t3 - t2) * 1e3))
                walls.append((t3 - t0) * 1e3)
                passes.append(g.walk_passes())
        ph = np.median(np.array(phases), axis=0)
        res[label] = dict(step_ms=round(float(np.median(walls)), 3), tree_build_ms=round(float(ph[0]), 3),
                          walk_first_pass_ms=round(float(ph[1]), 3), write_order_plan_ms=round(float(ph[2]), 3),
                          refill_ms=round(float(ph[3]), 3), force_ms=round(float(ph[4]), 3))
    g.update(tx, tm, tr)
    ngb_ms = []
    for _ in range(reps):
        cnt, _, _, ms = g.neighbors()
        ngb_ms.append(ms)
    res["neighbor_search"] = dict(ms=round(float(np.median(ngb_ms)), 3), entries=int(cnt.sum()),
                                  particles_with_neighbours=int((cnt > 1).sum()),
                                  tile_walk_passes_last=g.tile_walk_passes(), filter_passes_last=g.filter_passes(),
                                  note="FP64 lists of |dx|^2 < max(rs_i, rs_j)^2: per-tile r_search walks, "
                                       "exact FP64 filter on the device; a repeated search writes both in one "
                                       "pass into regions sized from the last search")
    acc2, pot2, nnb2 = r.results()
    same = same and bool(np.array_equal(acc2, ref[0]) and np.array_equal(pot2, ref[1]) and np.array_equal(nnb2, ref[2]))
    r.close()
    g.close()
    # rebuilds of the same particles write their lists in one walk pass
    # (regions sized from the previous build's list lengths)
    out = dict(res, forces_bit_identical_to_host_lists=same, reps=reps,
               walk_passes_per_rebuild=sorted(set(passes)))
    log(f"[bench] device tree step: {out}")
    return out


def tree_error(res_acc, res_pot, L, r_out, n_sample, seed=1):
    """Error of the tree force (theta truncation + kernel) against the FP64
    all-pairs direct sum with the same linear cutoff, on sampled i
    (informational: the tree's own approximation error)."""
    from oracle import oracle_py as O
    rng = np.random.default_rng(seed)
    ni = res_acc.shape[0]
    rows = np.sort(rng.choice(ni, min(n_sample, ni), replace=False))
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(L.epi_begin, L.n_epi)])[rows]
    t0 = time.time()
    dacc, dpot, _ = O.direct_linear_cutoff(L.epj_pos, L.epj_mass, L.epj_rs, sample=idx, eps2=0.0, rout2=r_out ** 2)
    ea = np.linalg.norm(res_acc[rows] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
    ep = np.abs(res_pot[rows] - dpot) / np.abs(dpot)
    return dict(sample_i=int(rows.size), acc_rel_median=float(np.median(ea)), acc_rel_p99=float(np.percentile(ea, 99)),
                pot_rel_median=float(np.median(ep)), direct_sum_s=round(time.time() - t0, 1))


def build_serving(n_clusters=256, seed=4, f_bin=0.3, r_out=0.01, nthreads=0):
    """The serving stream: n_clusters independent Plummer clusters, N
    log-uniform in [1e3, 2e5], Kroupa masses, 30% primordial binaries, each
    with its own tree (theta 0.3, quadrupole SPJ)."""
    from petar_amd import host
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(2e5), n_clusters)).astype(int)
    t0 = time.time()
    out = []
    for k, n in enumerate(sizes):
        m, x, _ = host.cluster(int(n), "plummer", kroupa=True, f_bin=f_bin, seed=seed * 1000 + k)
        out.append(host.build_lists(x, m, 1.5 * r_out, theta=0.3, nthreads=nthreads))
    log(f"[bench] serving stream: {n_clusters} clusters, N sum {int(sizes.sum())}, built in {time.time() - t0:.1f}s")
    return out, sizes


def _sweep(srv, ids, ints, cap_req_s, loads, rng, label):
    """Poisson arrivals of every cluster twice, at each load (fraction of the
    saturated request rate cap_req_s): p50/p99 latency over all requests and
    over those that had to upload their cluster."""
    out = []
    for f in loads:
        seq = rng.permutation(np.repeat(ids, 2))
        arr = np.cumsum(rng.exponential(1.0 / (f * cap_req_s), seq.size))
        arr -= arr[0]
        st0 = srv.stats()
        lat, wall, adm = srv.replay(seq, arr, admitted=True)
        st1 = srv.stats()
        row = dict(load=f, rate_req_s=round(f * cap_req_s, 1), p50_ms=round(float(np.percentile(lat, 50)), 4),
                   p99_ms=round(float(np.percentile(lat, 99)), 4),
                   gint_s=round(float(ints[seq].sum() / wall / 1e9), 2), admissions=int(adm.sum()),
                   evictions=int(st1["evicted"] - st0["evicted"]))
        if adm.any():
            row["admitted_p50_ms"] = round(float(np.percentile(lat[adm], 50)), 4)
            row["admitted_p99_ms"] = round(float(np.percentile(lat[adm], 99)), 4)
        out.append(row)
        log(f"[bench] serving {label} load {f}: {row}")
    return out


def serving(r_out, n_streams=4, loads=(0.5, 0.8, 0.95), seed=5, accuracy_clusters=12, budget_frac=0.25):
    """Fused server vs plain-form server on the 256-cluster stream: capacity
    and a Poisson load sweep with every cluster resident, then the same sweep
    under a binding HBM budget (every stream's scratch plus budget_frac of the
    clusters' resident bytes), where requests evict and re-admit clusters;
    every cluster's results alone (one request at a time) against its results
    after the mixes (bit-identical), sampled walks of 12 clusters against the
    FP64 oracle."""
    from petar_amd import soft_force as sf
    from oracle import oracle_py as O
    lists, sizes = build_serving(r_out=r_out)
    ctx = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0)
    srv = sf.Server(ctx, sf.SERVE_FUSED, hbm_budget=0, n_streams=n_streams)
    t0 = time.time()
    ids = [srv.add_lists(L) for L in lists]
    add_s = time.time() - t0
    ints = np.array([srv.interactions(k) for k in ids], dtype=np.int64)
    mem = srv.memory()
    rng = np.random.default_rng(seed)
    # alone: one request at a time (the first of each cluster admits it)
    lat_alone, alone = [], {}
    for k in ids:
        lat_alone.append(srv.wait(srv.submit(k)))
        alone[k] = srv.results(k)
    # saturated: every cluster once, all arriving at t = 0
    order = rng.permutation(ids)
    _, wall_sat = srv.replay(order, np.zeros(len(order)))
    cap_req_s = len(order) / wall_sat
    sat_gint = ints.sum() / wall_sat / 1e9
    sweep = _sweep(srv, ids, ints, cap_req_s, loads, rng, "resident")
    same = sum(all(np.array_equal(a, b) for a, b in zip(srv.results(k), alone[k])) for k in ids)
    st = srv.stats()
    mem_after = srv.memory()
    srv.close()
    # a binding budget: clusters are evicted and re-admitted throughout
    budget = int(n_streams * mem["scratch_max"] + budget_frac * mem["resident_total"])
    srvb = sf.Server(ctx, sf.SERVE_FUSED, hbm_budget=budget, n_streams=n_streams)
    idb = [srvb.add_lists(L) for L in lists]
    _, wall_b = srvb.replay(rng.permutation(idb), np.zeros(len(idb)))
    cap_b = len(idb) / wall_b
    sweep_b = _sweep(srvb, idb, ints, cap_b, (0.5, 0.8), rng, "budget")
    same_b = sum(all(np.array_equal(a, b) for a, b in zip(srvb.results(kb), alone[k])) for k, kb in zip(ids, idb))
    stb = srvb.stats()
    memb = srvb.memory()
    srvb.close()
    # accuracy: sampled walks of 12 clusters across the size range
    errs = []
    for k in np.argsort(sizes)[np.linspace(0, len(ids) - 1, accuracy_clusters).astype(int)]:
        L = lists[k]
        w = np.sort(rng.choice(L.n_walk, min(L.n_walk, 24), replace=False))
        a, rows = walk_subset(L, w)
        ref = O.soft_force_walks(**a, eps2=0.0, rout2=r_out ** 2)
        errs.append(np.linalg.norm(alone[ids[k]][0][rows] - ref[0], axis=1) / np.linalg.norm(ref[0], axis=1))
    errs = np.concatenate(errs)
    # plain per-walk form, same stream, saturated
    srvp = sf.Server(ctx, sf.SERVE_PLAIN)
    idp = [srvp.add_lists(L) for L in lists]
    _, wall_p = srvp.replay(rng.permutation(idp), np.zeros(len(idp)))
    srvp.close()
    ctx.close()
    plain_gint = ints.sum() / wall_p / 1e9
    n_part = int(sizes.sum())
    return dict(clusters=len(ids), particles=n_part, interactions_per_pass=int(ints.sum()),
                n_streams=n_streams, saturated_gint_s=round(float(sat_gint), 2),
                saturated_req_s=round(float(cap_req_s), 1), load_sweep=sweep,
                alone_ms_p50=round(float(np.percentile(lat_alone, 50)), 3),
                register_s=round(add_s, 2),
                plain_form_saturated_gint_s=round(float(plain_gint), 2),
                speedup_vs_plain=round(float(sat_gint / plain_gint), 2),
                alone_vs_mix_bit_identical=f"{same}/{len(ids)}",
                acc_rel_err_median=float(np.median(errs)), acc_rel_err_p99=float(np.percentile(errs, 99)),
                acc_rel_err_max=float(errs.max()), accuracy_sample_i=int(errs.size),
                resident_bytes_per_particle=round(mem["resident_total"] / n_part, 1),
                scratch_bytes_per_stream=mem["scratch_max"], pool_peak_bytes=st["pool_peak"],
                evictions=st["evicted"], graph_captures=mem_after["captures"],
                graph_instantiations=mem_after["instantiations"],
                binding_budget=dict(budget_bytes=budget, budget_note=f"{n_streams} streams' scratch + "
                                    f"{budget_frac} of the clusters' resident bytes",
                                    saturated_req_s=round(float(cap_b), 1),
                                    saturated_gint_s=round(float(ints.sum() / wall_b / 1e9), 2),
                                    load_sweep=sweep_b, admissions=stb["admitted"], evictions=stb["evicted"],
                                    pool_peak_bytes=stb["pool_peak"], graph_captures=memb["captures"],
                                    graph_instantiations=memb["instantiations"],
                                    alone_vs_mix_bit_identical=f"{same_b}/{len(ids)}"))


def read_traffic(name):
    """HBM bytes per force-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", f"pmc_force_{name}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                try:
                    return float(line.split(":", 1)[1])
                except ValueError:
                    return None
    return None


def read_list_gather(name):
    """HBM bytes per launch of the list-gather (staging) kernel, from the
    same committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", f"pmc_force_{name}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("list_gather_bytes_per_launch:"):
                try:
                    return float(line.split(":", 1)[1].split()[0])
                except ValueError:
                    return None
    return None


def host_cpu():
    """Model, logical CPUs and SMT of the host (from /proc/cpuinfo)."""
    model, siblings, cores = "unknown", None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "siblings" and siblings is None:
                    siblings = int(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
    except OSError:
        pass
    smt = None if not (siblings and cores) else siblings // cores
    return dict(model=model, logical_cpus=os.cpu_count(), smt_threads_per_core=smt,
                affinity_cpus=len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None)


def cpu_baseline(args_, r_out, target_s):
    """Reference x86 kernel (oracle/_ref, built from the reference source) on
    the host cores, same lists; repeated until target_s of CPU work.  Returns
    the baseline object and the kernel's outputs of the last run."""
    from oracle import oracle_py as O
    if not O.ref_available():
        return None, None
    nt = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    isa = "avx512" if O.cpu_has_avx512() else "avx2"
    n_int = int(np.dot(args_["n_epi"].astype(np.int64), (args_["n_epj"] + args_["n_spj"]).astype(np.int64)))
    O.ref_soft_force_walks(**args_, eps2=0.0, rout2=r_out ** 2, nthreads=nt, isa=isa)  # warm-up
    ts = []
    t_start = time.time()
    outs = None
    while time.time() - t_start < target_s or len(ts) < 3:
        t = time.time()
        outs = O.ref_soft_force_walks(**args_, eps2=0.0, rout2=r_out ** 2, nthreads=nt, isa=isa)
        ts.append(time.time() - t)
    med = float(np.median(ts))
    cpu = host_cpu()
    return dict(value=round(n_int / med / 1e9, 3), unit="Gint/s", cores=nt, kind="reference",
                sample=f"full workload lists ({n_int:.3e} interactions), median of {len(ts)} runs, "
                       f"phantomquad {isa} -DRSQRT_NR_EPJ_X2, {nt} OpenMP threads",
                cpu_model=cpu["model"], logical_cpus=cpu["logical_cpus"], affinity_cpus=cpu["affinity_cpus"],
                smt_threads_per_core=cpu["smt_threads_per_core"], isa=isa), outs


def vs_reference_kernel(res, ref_out, L):
    """The GPU evaluation against the reference x86 kernel's outputs on the
    same lists (every i): relative acc/pot errors and the neighbour counts,
    mismatches classified against the FP32 bands of both (the GPU's
    walk-relative coordinates and the x86 kernel's absolute ones; the x86
    kernel counts r^2 <= rs^2, the GPU r^2 < rs^2)."""
    acc, pot, nnb = res
    racc, rpot, rnnb = ref_out
    ea = np.linalg.norm(acc - racc, axis=1) / np.linalg.norm(racc, axis=1)
    ep = np.abs(pot - rpot) / np.abs(rpot)
    out = dict(i=int(ea.size), acc_rel_median=float(np.median(ea)), acc_rel_p99=float(np.percentile(ea, 99)),
               acc_rel_max=float(ea.max()), pot_rel_median=float(np.median(ep)),
               pot_rel_p99=float(np.percentile(ep, 99)), pot_rel_max=float(ep.max()))
    out.update(nnb_band(nnb, rnnb, L, absolute=True))
    return out


def cpu_baseline_neighbors(L):
    """The neighbour lists on the host cores: the FP64 restatement of
    SearchNeighborEpEpNoSimd (src/soft_force.hpp:11-34) over the same group
    EPJ lists, OpenMP over groups, counting and writing passes (FDPS's own
    short-range tree is not in the reference repository, so this is the
    CPU path timed beside the device neighbour search)."""
    from oracle import oracle_py as O
    nt = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    t = time.time()
    cnt, _, _ = O.ngb_walks(L.n_epi, L.n_epj, L.epj_pos, L.epj_rs, L.id_epj, L.epj_pos, L.epj_rs)
    dt = time.time() - t
    return dict(value=round(dt * 1e3, 1), unit="ms", cores=nt, kind="port",
                sample=f"the full workload: {int(cnt.sum())} entries from {L.id_epj.size} EPJ list entries, "
                       f"one run, {nt} OpenMP threads")


def walk_centres(n_epi, epi_pos):
    """Bounding-box centre of each walk's i (the centre the kernels subtract)."""
    off = np.concatenate([[0], np.cumsum(n_epi)])
    c = np.zeros((len(n_epi), 3))
    for w in range(len(n_epi)):
        if n_epi[w]:
            p = epi_pos[off[w]:off[w + 1]]
            c[w] = 0.5 * (p.min(0) + p.max(0))
    return c


def walk_subset(L, walks):
    """Oracle arguments of a subset of the walks of L, and the rows of their
    i in the evaluation's output (walk order)."""
    eoff, soff = L.walk_offsets()
    ioff = np.concatenate([[0], np.cumsum(L.n_epi)])
    idx = np.concatenate([np.arange(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w]) for w in walks])
    rows = np.concatenate([np.arange(ioff[w], ioff[w + 1]) for w in walks])
    a = dict(n_epi=L.n_epi[walks], n_epj=L.n_epj[walks], n_spj=L.n_spj[walks], epi_pos=L.epj_pos[idx],
             epi_rs=L.epj_rs[idx],
             id_epj=np.concatenate([L.id_epj[eoff[w]:eoff[w + 1]] for w in walks]),
             id_spj=np.concatenate([L.id_spj[soff[w]:soff[w + 1]] for w in walks]),
             epj_pos=L.epj_pos, epj_mass=L.epj_mass, epj_rs=L.epj_rs, spj_pos=L.spj_pos, spj_mass=L.spj_mass,
             spj_quad=L.spj_quad)
    return a, rows


def nnb_band(nnb, nnb_ref, L, rows_all=None, absolute=False):
    """Neighbour counts against a reference, with the pairs that lie within
    FP32 rounding of the search radius counted: the walks holding a
    mismatching i are searched for such pairs (the GPU's walk-relative FP32
    band; with absolute, also the band of FP32 absolute coordinates, the
    reference x86 kernel's), and every mismatch is classified as inside
    (|d| <= pairs in the band) or outside it (a wrong count)."""
    from oracle import oracle_py as O
    d = np.abs(np.asarray(nnb) - np.asarray(nnb_ref))
    bad = np.nonzero(d)[0]
    out = dict(nnb_mismatch=int(bad.size), nnb_mismatch_in_band=0, nnb_mismatch_outside_band=0)
    if bad.size == 0:
        return out
    ioff = np.concatenate([[0], np.cumsum(L.n_epi)])
    rows_all = np.arange(ioff[-1]) if rows_all is None else rows_all
    walks = np.unique(np.searchsorted(ioff, rows_all[bad], side="right") - 1)
    a, rows = walk_subset(L, walks)
    ka = {k: a[k] for k in ("n_epi", "n_epj", "epi_pos", "epi_rs", "id_epj", "epj_pos", "epj_rs")}
    amb = O.nnb_ambiguous(**ka, centres=walk_centres(a["n_epi"], a["epi_pos"]), eps2=0.0)
    if absolute:
        amb = amb + O.nnb_ambiguous(**ka, centres=None, eps2=0.0)
    pos = {r: k for k, r in enumerate(rows)}
    for b in bad:
        k = pos[int(rows_all[b])]
        if d[b] <= amb[k]:
            out["nnb_mismatch_in_band"] += 1
        else:
            out["nnb_mismatch_outside_band"] += 1
    return out


def accuracy(res_acc, res_pot, res_nnb, L, r_out, frac=0.05, seed=0):
    """Error vs the FP64 oracle on the same lists, on a sample of walks;
    neighbour counts compared exactly, mismatches classified against the
    pairs within FP32 rounding of the search radius (counted over the
    sample)."""
    from oracle import oracle_py as O
    rng = np.random.default_rng(seed)
    w_sel = np.sort(rng.choice(L.n_walk, max(1, int(frac * L.n_walk)), replace=False))
    a, rows = walk_subset(L, w_sel)
    racc, rpot, rnnb = O.soft_force_walks(**a, eps2=0.0, rout2=r_out ** 2)
    ea = np.linalg.norm(res_acc[rows] - racc, axis=1) / np.linalg.norm(racc, axis=1)
    ep = np.abs(res_pot[rows] - rpot) / np.abs(rpot)
    ka = {k: a[k] for k in ("n_epi", "n_epj", "epi_pos", "epi_rs", "id_epj", "epj_pos", "epj_rs")}
    amb = O.nnb_ambiguous(**ka, centres=walk_centres(a["n_epi"], a["epi_pos"]), eps2=0.0)
    d = np.abs(res_nnb[rows] - rnnb)
    return dict(sample_i=int(rows.size), sample_walks=int(w_sel.size), acc_rel_median=float(np.median(ea)),
                acc_rel_p99=float(np.percentile(ea, 99)), acc_rel_max=float(ea.max()),
                pot_rel_median=float(np.median(ep)), pot_rel_p99=float(np.percentile(ep, 99)),
                pot_rel_max=float(ep.max()), nnb_mismatch=int((d != 0).sum()),
                nnb_mismatch_outside_band=int((d > amb).sum()), nnb_mismatch_in_band=int(((d != 0) & (d <= amb)).sum()),
                nnb_pairs_in_fp32_band=int(amb.sum()), nnb_i_with_band_pairs=int((amb > 0).sum()))


def changeover_mode(L, r_out, tiles, steps=50, warmup=5, frac=0.02, seed=0, autotune=True, pipeline=(8, 3)):
    """The same lists in changeover mode (the (1 - K) weight and the
    pot_soft / pot_tot rules fused into the EPJ loop): ms per evaluation,
    Gint/s, and the error against the FP64 changeover oracle on a sample of
    walks.  Radii r_in = 0.1 r_out, r_out as the linear cutoff (the lists'
    r_search is 1.5 r_out); ids are the tree positions.  Tiles autotuned for
    this mode (its kernels hold more state per i than the linear-cutoff
    ones), starting from the linear-cutoff evaluation's; walk segments and
    streams as the linear-cutoff evaluation's (8.4M with 8 queues: 94.7-94.9
    ms on 16 segments and 4 streams, 96.3 ms on 8 and 3)."""
    from oracle import oracle_py as O
    from petar_amd import soft_force as sf
    rin, rout = np.full(L.epj_mass.size, 0.1 * r_out), np.full(L.epj_mass.size, r_out)
    ctx = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, changeover=1, max_ipl=tiles[0], chunk_j=tiles[1],
                     segments=pipeline[0], streams=pipeline[1])
    res = sf.Resident.from_lists(ctx, L, rin, rout)
    tuned = None
    if autotune:
        t0 = time.time()
        ipl, cj, tms = res.autotune(ipl=(1, 2, 3, 4), chunk=(512, 1024, 2048), reps=3)
        tiles = (ipl, cj)
        tuned = dict(max_ipl=ipl, chunk_j=cj, ms=round(tms, 4), tuning_s=round(time.time() - t0, 2))
    for _ in range(max(warmup, 1)):
        res.run()
    res.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        res.run()
    res.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    acc, ps, _, pt = res.results()
    res.close()
    ctx.close()
    # the same evaluation unsegmented on one stream (its phases from HIP events)
    ctx1 = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, changeover=1, max_ipl=tiles[0], chunk_j=tiles[1],
                      segments=1)
    r1 = sf.Resident.from_lists(ctx1, L, rin, rout)
    for _ in range(max(warmup, 1)):
        r1.run()
    ph = []
    for _ in range(min(steps, 20)):
        r1.run()
        ph.append(r1.timing())
    r1.close()
    ctx1.close()
    ph = np.median(np.array(ph), axis=0)
    rng = np.random.default_rng(seed)
    w_sel = np.sort(rng.choice(L.n_walk, max(1, int(frac * L.n_walk)), replace=False))
    eoff, soff = L.walk_offsets()
    ioff = np.concatenate([[0], np.cumsum(L.n_epi)])
    idx = np.concatenate([np.arange(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w]) for w in w_sel])
    rows = np.concatenate([np.arange(ioff[w], ioff[w + 1]) for w in w_sel])
    racc, rps, rpt, _, _ = O.soft_force_walks_co(
        L.n_epi[w_sel], L.n_epj[w_sel], L.n_spj[w_sel], L.epj_pos[idx], L.epj_rs[idx], idx,
        np.concatenate([L.id_epj[eoff[w]:eoff[w + 1]] for w in w_sel]),
        np.concatenate([L.id_spj[soff[w]:soff[w + 1]] for w in w_sel]), L.epj_pos, L.epj_mass, L.epj_rs,
        np.arange(L.epj_mass.size), rin, rout, L.spj_pos, L.spj_mass, L.spj_quad, eps2=0.0, rout2=r_out ** 2)
    ea = np.linalg.norm(acc[rows] - racc, axis=1) / np.linalg.norm(racc, axis=1)
    et = np.abs(pt[rows] - rpt) / np.abs(rpt)
    ne, ns = L.interactions()
    out = dict(ms_per_step=round(ms, 4), gint_s=round((ne + ns) / (ms * 1e-3) / 1e9, 2),
               tiles=dict(max_ipl=tiles[0], chunk_j=tiles[1]), autotuned=tuned,
               unsegmented_ms=dict(stage=round(float(ph[0]), 4), force=round(float(ph[1]), 4),
                                   reduce=round(float(ph[2]), 4), step=round(float(ph.sum()), 4)),
               sample_i=int(rows.size),
               acc_rel_median=float(np.median(ea)), acc_rel_p99=float(np.percentile(ea, 99)),
               pot_tot_rel_median=float(np.median(et)), pot_tot_rel_p99=float(np.percentile(et, 99)),
               radii="r_in = 0.1 r_out, r_out = the cutoff radius, r_search = 1.5 r_out")
    log(f"[bench] changeover mode: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--exec", default="fused", choices=["fused", "plain"])
    ap.add_argument("--chunk-j", type=int, default=0)
    ap.add_argument("--max-ipl", type=int, default=2, choices=[1, 2, 3, 4])
    ap.add_argument("--j-vgpr", type=int, default=0, choices=[0, 1])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--segments", type=int, default=0, help="walk segments (0: the workload's setting)")
    ap.add_argument("--streams", type=int, default=0, help="pipeline streams (0: the workload's setting)")
    ap.add_argument("--hw-queues", type=int, default=-1,
                    help="GPU_MAX_HW_QUEUES for this process (0: as the environment has it; -1: the workload's "
                         "setting)")
    ap.add_argument("--direct-sample", type=int, default=65536,
                    help="i sampled for the error against the FP64 all-pairs direct sum (0: skip)")
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--no-device-tree", action="store_true")
    ap.add_argument("--no-changeover", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay comparison")
    ap.add_argument("--persistent", type=int, default=0, choices=[0, 1],
                    help="force launches as one wave per free wave slot taking items from per-XCD counters")
    ap.add_argument("--df-near", type=int, default=1, choices=[0, 1],
                    help="near EPJ pairs in double-float walk-relative coordinates (linear-cutoff mode)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); the soft-force path is "
              f"single-GPU here and multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)
    # hardware queues of the process, set before the first HIP call: a
    # 4-stream pipeline (1 staging + 3 force), the context's stream and the
    # null stream are 6 streams, and beyond HIP's default of 4 queues streams
    # share a queue and run in submission order with its other streams
    # (measured: a process's first 16-segment, 4-stream pipeline at 93.8 ms
    # per 8.4M step with 4 queues, 89.0 ms with 8)
    pipe = PIPELINE.get(a.workload, (8, 3, 0))
    a.segments = a.segments or pipe[0]
    a.streams = a.streams or pipe[1]
    if a.hw_queues < 0:
        a.hw_queues = pipe[2]
    if a.hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)

    from petar_amd import soft_force as sf
    L, r_out, desc, prep, particles = build_workload(a.workload)
    ne, ns = L.interactions()
    n_int = ne + ns
    flops = FLOP_EPJ * ne + FLOP_SPJ * ns

    import torch  # device-wide synchronisation around the timed region
    exec_mode = sf.EXEC_FUSED if a.exec == "fused" else sf.EXEC_PLAIN
    ctx = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, nnb_mode=0, exec_mode=exec_mode, chunk_j=a.chunk_j,
                     max_ipl=a.max_ipl, j_vgpr=a.j_vgpr, segments=a.segments, streams=a.streams, df_near=a.df_near,
                     persistent=a.persistent)
    res = sf.Resident.from_lists(ctx, L)
    tuned = None
    if a.exec == "fused" and not a.no_autotune:
        t0 = time.time()
        ipl, cj, tms = res.autotune(ipl=(1, 2, 3, 4), chunk=(1024, 2048, 4096), reps=3)
        tuned = dict(max_ipl=ipl, chunk_j=cj, ms=round(tms, 4), tuning_s=round(time.time() - t0, 2))
        log(f"[bench] autotuned: {tuned}")
    st = res.stats()
    log(f"[bench] resident: N_e {ne:.3e} N_s {ns:.3e} items {st['n_items']} staged {st['staged_bytes'] / 2**20:.0f} MiB")
    for _ in range(a.warmup):
        res.run()
    res.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res.run()
    res.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    log(f"[bench] timed {a.steps} steps: {ms_step:.3f} ms/step")

    # whole-evaluation latency from HIP events (segments overlap on several
    # streams, the events bracket all of them)
    lat = []
    for _ in range(min(a.steps, 20)):
        res.run()
        s, f, r = res.timing()
        lat.append(s + f + r)
    # per-kernel times: the same evaluation and tiles unsegmented on one
    # stream, HIP events around the stage, force and reduce launches
    ipl_used = tuned["max_ipl"] if tuned else a.max_ipl
    cj_used = tuned["chunk_j"] if tuned else a.chunk_j
    ctx1 = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, exec_mode=exec_mode, chunk_j=cj_used, max_ipl=ipl_used,
                      j_vgpr=a.j_vgpr, segments=1, df_near=a.df_near, persistent=a.persistent)
    r1 = sf.Resident.from_lists(ctx1, L)
    r1.run()
    k_force, k_stage, k_red = [], [], []
    for _ in range(min(a.steps, 20)):
        r1.run()
        s, f, r = r1.timing()
        k_stage.append(s)
        k_force.append(f)
        k_red.append(r)
    r1.close()
    ctx1.close()
    force_ms = float(np.mean(k_force))
    unseg_ms = float(np.mean(np.array(k_stage) + np.array(k_force) + np.array(k_red)))
    achieved = flops / (force_ms * 1e-3) / 1e12
    traffic = read_traffic(a.workload)
    # the list-gather phase (k_stage_walk: every walk's EPJ/SPJ lists gathered
    # into walk-relative records) against HBM: its bytes per launch from the
    # committed pmc summary over its time in the unsegmented evaluation above
    stage_ms = float(np.mean(k_stage))
    lg_bytes = read_list_gather(a.workload)
    list_gather = None
    if lg_bytes is not None and stage_ms > 0:
        gbs = lg_bytes / (stage_ms * 1e-3) / 1e9
        list_gather = {"kernel": "k_stage_walk", "ms": round(stage_ms, 4), "hbm_bytes": lg_bytes,
                       "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(gbs / HBM_PEAK_GBS, 4),
                       "source": f"profiles/pmc_force_{a.workload}.txt (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"}

    out = {
        "metric": "pair interactions/s (Gint/s, % FP32 peak) + p50/p99 latency at fixed force error",
        "value": round(n_int / (ms_step * 1e-3) / 1e9, 2),
        "unit": "Gint/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 accumulation",
        "data": "synthetic",
        "config": {"workload": a.workload, "description": desc, "exec": a.exec,
                   "interactions_per_step": n_int, "n_epep": ne, "n_epsp": ns, "walks": L.n_walk,
                   "flops_per_step": flops, "flop_convention": "26/EPJ pair + 65/SPJ pair"},
        "roofline": {"bound": "valu-fp32", "achieved": round(achieved, 3), "peak": P32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / P32_PEAK_TFLOPS, 4), "traffic": traffic,
                     "kernel": "k_force (FP32 vector ALU, no MFMA); all class launches of one evaluation "
                               "(the 1-i class on a second stream beside the others), unsegmented; HIP events "
                               "before the first and after the last launch, on the evaluation's stream",
                     "kernel_ms": round(force_ms, 4)},
        "list_gather": list_gather,
        "pct_fp32_peak_step": round(flops / (ms_step * 1e-3) / 1e12 / P32_PEAK_TFLOPS * 100, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                       "stage_ms": round(float(np.mean(k_stage)), 4), "reduce_ms": round(float(np.mean(k_red)), 4),
                       "unsegmented_step_ms": round(unseg_ms, 4)},
        "segments": {"segments": a.segments, "streams": a.streams,
                     "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "HIP default (4)")},
        "df_near": a.df_near,
        "prep": prep,
        "autotuned": tuned,
    }
    acc, pot, nnb = res.results()
    if a.exec == "fused" and not a.no_graph:
        # the same evaluation replayed as one captured hipGraph (same tiles,
        # segments and streams, same warm-up and step count)
        ctxg = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, nnb_mode=0, exec_mode=exec_mode, chunk_j=cj_used,
                          max_ipl=ipl_used, j_vgpr=a.j_vgpr, segments=a.segments, streams=a.streams,
                          df_near=a.df_near, graph=1, persistent=a.persistent)
        rg = sf.Resident.from_lists(ctxg, L)
        for _ in range(a.warmup):
            rg.run()
        rg.sync()
        torch.cuda.synchronize()
        tg0 = time.perf_counter()
        for _ in range(a.steps):
            rg.run()
        rg.sync()
        torch.cuda.synchronize()
        ms_graph = (time.perf_counter() - tg0) * 1e3 / a.steps
        ga, gp, gn = rg.results()
        rg.close()
        ctxg.close()
        out["graph_replay"] = dict(ms_per_step=round(ms_graph, 4), gint_s=round(n_int / (ms_graph * 1e-3) / 1e9, 2),
                                   launched_ms_per_step=round(ms_step, 4),
                                   bit_identical=bool(np.array_equal(ga, acc) and np.array_equal(gp, pot)
                                                      and np.array_equal(gn, nnb)))
        log(f"[bench] graph replay: {out['graph_replay']}")
    out["config"]["flop_convention_note"] = (
        "EPJ pair 26 (src/force_gpu_cuda.cu:57-74): dx 3 sub, r2 3 FMA (6), search radius max + square + "
        "compare (3), cutoff max (1), rsqrt (1), m rinv (1), G rinv rinv pij (3), acc 3 FMA (6), pot G pij + sub "
        "(2); SPJ pair 65 (quadrupole, :283-312), FMA = 2")
    if not a.no_accuracy:
        out["accuracy_vs_fp64_same_lists"] = accuracy(acc, pot, nnb, L, r_out)
        if a.direct_sample > 0:
            out["error_vs_fp64_direct_sum"] = tree_error(acc, pot, L, r_out, a.direct_sample)
    res.close()
    if not a.no_changeover and a.exec == "fused":
        # same warm-up and step count as the timed evaluation above
        out["changeover_mode"] = changeover_mode(L, r_out, (ipl_used, cj_used), steps=a.steps, warmup=a.warmup,
                                                 autotune=a.exec == "fused" and not a.no_autotune,
                                                 pipeline=(a.segments, a.streams))
    if not a.no_device_tree and a.exec == "fused":
        # same tiles and segments as the timed evaluation, so the forces can
        # be compared bit for bit
        ctxd = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, nnb_mode=0, exec_mode=exec_mode, chunk_j=cj_used,
                          max_ipl=ipl_used, j_vgpr=a.j_vgpr, segments=a.segments, streams=a.streams,
                          df_near=a.df_near, persistent=a.persistent)
        out["device_tree_step"] = device_tree_step(ctxd, particles, r_out, L, (acc, pot, nnb))
        ctxd.close()
        out["device_tree_step"]["host_tree_walk_s"] = prep["tree_walk_s"]
    if not a.no_plain and a.exec == "fused":
        ctxp = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, exec_mode=sf.EXEC_PLAIN)
        rp = sf.Resident.from_lists(ctxp, L)
        rp.run()
        rp.sync()
        tp = []
        for _ in range(5):
            rp.run()
            s, f, r = rp.timing()
            tp.append(s + f + r)
        out["plain_form_gint_s"] = round(n_int / (np.median(tp) * 1e-3) / 1e9, 2)
        rp.close()
        ctxp.close()
    ctx.close()
    if not a.no_serving:
        # the 256-cluster stream has its own definition (r_out = 0.01 as in
        # the precision configuration), whatever the headline workload
        out["serving"] = serving(0.01, n_streams=a.serve_streams)
        sv = out["serving"]
        p80 = [x for x in sv["load_sweep"] if x["load"] == 0.8]
        if p80:
            out["latency_ms"]["serving_p50_at_80pct_load"] = p80[0]["p50_ms"]
            out["latency_ms"]["serving_p99_at_80pct_load"] = p80[0]["p99_ms"]
    if not a.no_cpu:
        from oracle import oracle_py as O
        out["cpu_baseline"], ref_out = cpu_baseline(O.lists_args(L), r_out, a.cpu_seconds)
        if ref_out is not None:
            out["error_vs_reference_cpu_kernel"] = vs_reference_kernel((acc, pot, nnb), ref_out, L)
            log(f"[bench] vs reference x86 kernel: {out['error_vs_reference_cpu_kernel']}")
        if "device_tree_step" in out:
            out["device_tree_step"]["neighbor_search"]["cpu_baseline"] = cpu_baseline_neighbors(L)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
