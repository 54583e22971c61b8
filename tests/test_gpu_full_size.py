"""The precision workload and the two largest workloads at their full sizes,
checked through properties that do not need an all-pairs reference: the
device tree equals the host tree entry for entry; the pipelined, the
unsegmented and the device-fed evaluations agree bit for bit; sampled walks
agree with the FP64 oracle on the same lists; the neighbour lists equal the
FP64 restatement over all walks; in changeover mode the pipelined and
unsegmented evaluations agree bit for bit and sampled walks agree with the
FP64 changeover oracle.

Workloads (theta 0.3, n_leaf 20, n_group 512, r_search = 1.5 r_out):
precision, Plummer N = 1,048,576 with 50% primordial binaries, Kroupa
masses, r_out 0.01; batching, King W0 = 7 N = 4,194,304, Kroupa,
mass-segregated, r_out 0.01 (132M EPJ + 230M SPJ list entries); large,
Plummer N = 8,388,608 with 100% binaries, r_out 0.003 (251M EPJ + 585M SPJ
list entries)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import host, soft_force as sf
from tests._common import nnb_check, rel_err, walk_centres

pytestmark = pytest.mark.gpu

WORKLOADS = {
    "precision": (dict(n=1 << 20, model="plummer", f_bin=0.5, seed=2), 0.01),
    "batching": (dict(n=1 << 22, model="king", f_bin=0.0, seed=3, segregate=True), 0.01),
    "large": (dict(n=1 << 23, model="plummer", f_bin=1.0, seed=5), 0.003),
}


@pytest.fixture(scope="module", params=list(WORKLOADS))
def workload(request):
    kw, r_out = WORKLOADS[request.param]
    kw = dict(kw)
    m, x, _ = host.cluster(kw.pop("n"), kw.pop("model"), kroupa=True, **kw)
    rs = 1.5 * r_out
    L = host.build_lists(x, m, rs, theta=0.3, n_leaf_limit=20, n_group_limit=512)
    print(f"[full size {request.param}] {L.n_walk} walks, {L.id_epj.size} EPJ + {L.id_spj.size} SPJ entries",
          flush=True)
    return m, x, L, r_out, rs


def _run(ctx, res):
    res.run()
    out = res.results()
    res.close()
    return out


def test_full_size_device_tree_equals_host_tree(workload):
    m, x, L, r_out, rs = workload
    ctx = sf.Context(0)
    g = sf.GpuTree(ctx, x, m, np.full(m.size, rs), theta=0.3)
    d = g.download()
    g.close()
    ctx.close()
    assert np.array_equal(d["order"], L.order)
    assert np.array_equal(d["node_begin"], L.node_begin) and np.array_equal(d["node_end"], L.node_end)
    assert np.array_equal(d["node_mass"], L.spj_mass) and np.array_equal(d["node_quad"], L.spj_quad)
    assert np.array_equal(d["node_com"], L.spj_pos)
    assert np.array_equal(d["n_epi"], L.n_epi) and np.array_equal(d["n_epj"], L.n_epj)
    assert np.array_equal(d["n_spj"], L.n_spj)
    assert np.array_equal(d["id_epj"], L.id_epj) and np.array_equal(d["id_spj"], L.id_spj)


def test_full_size_evaluations_agree(workload):
    m, x, L, r_out, rs = workload
    kw = dict(eps2=0.0, rcut2=r_out ** 2, G=1.0)
    ctx = sf.Context(0, **kw)  # pipelined over 8 segments on 3 streams
    ref = _run(ctx, sf.Resident.from_lists(ctx, L))
    ctx1 = sf.Context(0, segments=1, **kw)
    one = _run(ctx1, sf.Resident.from_lists(ctx1, L))
    g = sf.GpuTree(ctx, x, m, np.full(m.size, rs), theta=0.3)
    dev = _run(ctx, sf.Resident.from_gtree(ctx, g))
    g.close()
    for a, b, c in zip(ref, one, dev):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    assert np.isfinite(ref[0]).all() and np.isfinite(ref[1]).all()
    # sampled walks against the FP64 oracle on the same lists
    rng = np.random.default_rng(5)
    walks = np.sort(rng.choice(L.n_walk, 48, replace=False))
    eo, so = L.walk_offsets()
    rows, ide, ids = [], [], []
    for w in walks:
        rows.append(np.arange(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w]))
        ide.append(L.id_epj[eo[w]:eo[w + 1]])
        ids.append(L.id_spj[so[w]:so[w + 1]])
    rows = np.concatenate(rows)
    a = dict(n_epi=L.n_epi[walks], n_epj=L.n_epj[walks], n_spj=L.n_spj[walks], epi_pos=L.epj_pos[rows],
             epi_rs=L.epj_rs[rows], id_epj=np.concatenate(ide), id_spj=np.concatenate(ids), epj_pos=L.epj_pos,
             epj_mass=L.epj_mass, epj_rs=L.epj_rs, spj_pos=L.spj_pos, spj_mass=L.spj_mass, spj_quad=L.spj_quad)
    racc, rpot, rnnb = O.soft_force_walks(**a, eps2=0.0, rout2=r_out ** 2)
    # the evaluation's i are the walks' EPI in walk order, which here is the tree order
    ea, ep = rel_err(ref[0][rows], racc), rel_err(ref[1][rows], rpot)
    print(f"[full size] {rows.size} sampled i: acc median {np.median(ea):.2e} p99 {np.percentile(ea, 99):.2e}",
          flush=True)
    assert np.median(ea) <= 1e-6 and np.median(ep) <= 1e-6 and np.percentile(ea, 99) <= 1e-4
    # neighbour counts exact wherever no pair lies within FP32 rounding of the
    # search radius, and off by at most the number of such pairs elsewhere
    amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
    bad_exact, bad_band, n_amb = nnb_check(ref[2][rows], rnnb, amb)
    print(f"[full size] nnb: {int((ref[2][rows] != rnnb).sum())} mismatches, outside the band {bad_exact}, "
          f"beyond it {bad_band}; i with band pairs {n_amb}, band pairs {int(amb.sum())}", flush=True)
    assert bad_exact == 0 and bad_band == 0
    ctx.close()
    ctx1.close()


def test_full_size_neighbour_lists(workload):
    m, x, L, r_out, rs = workload
    ocnt, ooff, ongb = O.ngb_walks(L.n_epi, L.n_epj, L.epj_pos, L.epj_rs, L.id_epj, L.epj_pos, L.epj_rs)
    ctx = sf.Context(0)
    g = sf.GpuTree(ctx, x, m, np.full(m.size, rs), theta=0.3)
    cnt, off, ngb, ms = g.neighbors()
    g.close()
    ctx.close()
    n = m.size
    ecnt = np.zeros(n, np.int64)
    ecnt[L.order] = ocnt
    assert np.array_equal(cnt, ecnt)
    rowp = np.repeat(np.arange(n), ocnt)
    o = np.argsort(L.order[rowp], kind="stable")
    assert np.array_equal(ngb, L.order[ongb[o]])
    print(f"[full size] neighbour lists: {ngb.size} entries, {ms:.2f} ms", flush=True)


def test_full_size_changeover_mode(workload):
    """Changeover mode (r_in = 0.1 r_out) on the full lists: the pipelined
    and the unsegmented evaluations agree bit for bit, and sampled walks
    agree with the FP64 changeover oracle (acc and pot_tot everywhere,
    pot_soft away from pairs within FP32 rounding of the neighbour radius,
    where it jumps)."""
    m, x, L, r_out, rs = workload
    n = m.size
    rin, rout = np.full(n, 0.1 * r_out), np.full(n, r_out)
    out = []
    for seg in (8, 1):
        ctx = sf.Context(0, eps2=0.0, rcut2=r_out ** 2, G=1.0, changeover=1, segments=seg)
        out.append(_run(ctx, sf.Resident.from_lists(ctx, L, rin, rout)))
        ctx.close()
    assert all(np.array_equal(a, b) for a, b in zip(*out))
    acc, ps, nnb, pt = out[0]
    rng = np.random.default_rng(9)
    walks = np.sort(rng.choice(L.n_walk, 48, replace=False))
    eo, so = L.walk_offsets()
    ioff = np.concatenate([[0], np.cumsum(L.n_epi)])
    idx = np.concatenate([np.arange(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w]) for w in walks])
    rows = np.concatenate([np.arange(ioff[w], ioff[w + 1]) for w in walks])
    racc, rps, rpt, _, amb = O.soft_force_walks_co(
        L.n_epi[walks], L.n_epj[walks], L.n_spj[walks], L.epj_pos[idx], L.epj_rs[idx], idx,
        np.concatenate([L.id_epj[eo[w]:eo[w + 1]] for w in walks]),
        np.concatenate([L.id_spj[so[w]:so[w + 1]] for w in walks]), L.epj_pos, L.epj_mass, L.epj_rs,
        np.arange(n), rin, rout, L.spj_pos, L.spj_mass, L.spj_quad, eps2=0.0, rout2=r_out ** 2, amb_rel=1e-4)
    ea, et = rel_err(acc[rows], racc), rel_err(pt[rows], rpt)
    keep = amb == 0
    es = rel_err(ps[rows][keep], rps[keep])
    print(f"[full size] changeover, {rows.size} sampled i: acc median {np.median(ea):.2e} p99 "
          f"{np.percentile(ea, 99):.2e}; pot_tot median {np.median(et):.2e}; pot_soft median {np.median(es):.2e} "
          f"({int((~keep).sum())} i at the neighbour radius left out)", flush=True)
    for e in (ea, et, es):
        assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-4
