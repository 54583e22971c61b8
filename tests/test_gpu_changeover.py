"""GPU parity of the changeover mode (the (1 - K) changeover weight of
src/hard.hpp:1404-1476,1655-1691 folded into the force kernel) against the
FP64 oracle orc_soft_force_walks_co on the same lists.

Tolerances as for the linear-cutoff form: median relative error of acc,
pot_soft and pot_tot <= 1e-6, 99th percentile <= 1e-5, maximum <= 1e-3.
pot_soft jumps by G m (1 + c) / r_out where a pair crosses the neighbour
radius (the neighbour pair's soft potential carries that constant), so i with
a pair within FP32 rounding of the neighbour radius (n_amb > 0 from the
oracle) are left out of the pot_soft comparison and counted; their acc and
pot_tot are still compared (both are continuous there).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import soft_force as sf
from tests._common import R_OUT, co_args, co_lists, nnb_check, rel_err, walk_centres

pytestmark = pytest.mark.gpu

MED, P99, MAX = 1e-6, 1e-5, 1e-3
AMB_REL = 1e-4


def _workload(n, seed, f_bin, r_out=R_OUT, eps2=0.0, G=1.0, shift=(0.0, 0.0, 0.0)):
    L, rin, rout = co_lists(n, seed=seed, f_bin=f_bin, r_out=r_out, r_in=0.1 * r_out, shift=shift)
    a = co_args(L, rin, rout)
    ref = O.soft_force_walks_co(**a, eps2=eps2, rout2=r_out ** 2, G=G, amb_rel=AMB_REL)
    amb_nnb = O.nnb_ambiguous(**{k: a[k] for k in ("n_epi", "n_epj", "epi_pos", "epi_rs", "id_epj", "epj_pos",
                                                  "epj_rs")}, centres=walk_centres(a), eps2=eps2)
    return L, rin, rout, ref, amb_nnb


@pytest.fixture(scope="module")
def single():
    return _workload(16384, 1, 0.0)


@pytest.fixture(scope="module")
def binaries():
    return _workload(16384, 4, 0.3)


@pytest.fixture(scope="module")
def dense():
    # r_out = 0.03: several neighbours per particle in the core
    return _workload(16384, 6, 0.3, r_out=0.03)


def _stats(e):
    return f"med {np.median(e):.3e} p99 {np.percentile(e, 99):.3e} max {e.max():.3e}"


def _check(res, ref, amb_nnb, label):
    acc, ps, nnb, pt = res
    racc, rps, rpt, rnnb, amb = ref
    ea, et = rel_err(acc, racc), rel_err(pt, rpt)
    keep = amb == 0
    es = rel_err(ps[keep], rps[keep])
    print(f"[{label}] acc {_stats(ea)}; pot_soft {_stats(es)} ({int((~keep).sum())} i at the neighbour radius "
          f"left out); pot_tot {_stats(et)}", flush=True)
    for e in (ea, es, et):
        assert np.median(e) <= MED and np.percentile(e, 99) <= P99 and e.max() <= MAX
    bad_exact, bad_band, _ = nnb_check(nnb, rnnb, amb_nnb)
    assert bad_exact == 0 and bad_band == 0


def _resident(L, rin, rout, r_out=R_OUT, eps2=0.0, G=1.0, **kw):
    ctx = sf.Context(0, eps2=eps2, rcut2=r_out ** 2, G=G, changeover=1, **kw)
    r = sf.Resident.from_lists(ctx, L, rin, rout)
    r.run()
    res = r.results()
    r.close()
    ctx.close()
    return res


def test_co_resident_singles(single):
    L, rin, rout, ref, amb = single
    _check(_resident(L, rin, rout), ref, amb, "co singles")


def test_co_resident_binaries(binaries):
    L, rin, rout, ref, amb = binaries
    res = _resident(L, rin, rout)
    _check(res, ref, amb, "co binaries")
    # a second evaluation is bit-identical
    res2 = _resident(L, rin, rout)
    assert all(np.array_equal(x, y) for x, y in zip(res, res2))


def test_co_resident_dense(dense):
    L, rin, rout, ref, amb = dense
    print(f"[co dense] i with neighbours {(ref[1] != ref[2]).sum()} of {ref[1].size}", flush=True)
    _check(_resident(L, rin, rout, r_out=0.03), ref, amb, "co dense")


@pytest.mark.parametrize("eps,G", [(0.05, 1.0), (0.3, 1.7)])
def test_co_softened(eps, G):
    """Softening eps = 0.05 and 0.3 r_out in changeover mode: the neighbour
    correction takes K and the 1/r terms at sqrt(r^2 + eps^2)
    (src/hard.hpp:1408-1433), the linear-cutoff part as without changeover;
    G = 1.7 scales the pair terms and the self term G m_i / r_out."""
    eps2 = (eps * 0.03) ** 2
    L, rin, rout, ref, amb = _workload(16384, 6, 0.3, r_out=0.03, eps2=eps2, G=G)
    _check(_resident(L, rin, rout, r_out=0.03, eps2=eps2, G=G), ref, amb, f"co softened eps={eps} r_out G={G}")


def test_co_translated_binaries():
    """Changeover mode with the binaries cluster moved to (8000, -300, 20):
    neighbour pairs use separations taken in FP64 from the walk centre, so
    the errors stay those of the cluster at the origin."""
    L, rin, rout, ref, amb = _workload(16384, 4, 0.3, shift=(8000.0, -300.0, 20.0))
    _check(_resident(L, rin, rout), ref, amb, "co binaries at x = 8000")


@pytest.mark.parametrize("max_ipl", [1, 2, 3, 4])
def test_co_split_chunks(binaries, max_ipl):
    """Lists split over many items: the self term is removed exactly once
    and pot_tot goes through the reduction."""
    L, rin, rout, ref, amb = binaries
    _check(_resident(L, rin, rout, chunk_j=256, max_ipl=max_ipl), ref, amb, f"co chunk=256 max_ipl={max_ipl}")


def test_co_linear_cutoff_unchanged_far_from_neighbours(single):
    """Changeover mode differs from the linear-cutoff form only through
    neighbour pairs and the self term: with a neighbour factor so small that
    no pair qualifies, acc equals the linear-cutoff acc and pot equals it
    plus G m_i / r_out."""
    L, rin, rout, _, _ = single
    # changeover mode takes the linear-cutoff terms of its pairs from the
    # FP32 walk-relative coordinates (double-float separations only for its
    # neighbour pairs): compare with the linear-cutoff form without
    # double-float near pairs
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, df_near=0)
    r = sf.Resident.from_lists(ctx, L)
    r.run()
    acc_a, pot_a, nnb_a = r.results()
    r.close()
    ctx.close()
    acc, ps, nnb, pt = _resident(L, rin, rout, ngb_factor=1e-30)
    assert np.array_equal(acc, acc_a) and np.array_equal(nnb, nnb_a)
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(L.epi_begin, L.n_epi)])
    m_i = L.epj_mass[idx].astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(ps, pot_a + m_i / np.float32(R_OUT), rtol=1e-6)
    assert np.array_equal(ps, pt)


def _replay(L, rin, rout, n_walk_limit=200):
    """FDPS call order with FDPS-layout records and ForceSoft + pot_tot."""
    n = L.epj_pos.shape[0]
    epj = np.zeros(n, sf.EPJ_SOFT)
    epj["id"] = np.arange(n)
    epj["mass"] = L.epj_mass
    epj["pos"] = L.epj_pos
    epj["r_search"] = L.epj_rs
    epj["r_in"] = rin
    epj["r_out"] = rout
    ns = L.spj_mass.size
    spj = np.zeros(ns, sf.SPJ_SOFT)
    spj["mass"] = L.spj_mass
    spj["pos"] = L.spj_pos
    for k, q in enumerate(("xx", "yy", "zz", "xy", "xz", "yz")):
        spj["quad"][q] = L.spj_quad[:, k]
    eoff, soff = L.walk_offsets()
    fn = sf.make_functor(0, 0.0, R_OUT ** 2, 1.0, changeover=1)
    fn(0, 0, [], [], [], [], [], [], epj, n, spj, ns, True)
    out = []
    for b0 in range(0, L.n_walk, n_walk_limit):
        ws = list(range(b0, min(b0 + n_walk_limit, L.n_walk)))
        epi, force = [], []
        for w in ws:
            sl = slice(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w])
            e = np.zeros(L.n_epi[w], sf.EPI_SOFT)
            e["id"] = np.arange(n)[sl]
            e["pos"] = L.epj_pos[sl]
            e["r_search"] = L.epj_rs[sl]
            e["type"] = 1
            epi.append(e)
            f = np.zeros(L.n_epi[w], sf.FORCE_SOFT_CO)
            f["pot"] = f["pot_tot"] = 123.0  # assigned by the retrieve
            force.append(f)
        fn(0, len(ws), epi, L.n_epi[ws], [L.id_epj[eoff[w]:eoff[w + 1]] for w in ws], L.n_epj[ws],
           [L.id_spj[soff[w]:soff[w + 1]] for w in ws], L.n_spj[ws], epj, n, spj, ns, False)
        sf.RetrieveForce(0, len(ws), L.n_epi[ws], force, fn)
        out.extend(force)
    f = np.concatenate(out)
    return f["acc"], f["pot"], f["n_ngb"], f["pot_tot"]


def test_co_fdps_replay_matches_resident(binaries):
    L, rin, rout, ref, amb = binaries
    res = _replay(L, rin, rout)
    _check(res, ref, amb, "co fdps replay")
    res_r = _resident(L, rin, rout)
    assert all(np.array_equal(x, y) for x, y in zip(res, res_r))


def test_co_missing_own_entry_is_reported(single):
    L, rin, rout, _, _ = single
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, changeover=1)
    r = sf.Resident.from_lists(ctx, L, rin, rout)
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(L.epi_begin, L.n_epi)])
    bad = idx.copy()
    bad[5] = -77  # no EPJ carries this id
    r.changeover(bad, np.arange(len(L.epj_mass)), rin, rout)
    r.run()
    with pytest.raises(sf.PthError, match="no EPJ with its id"):
        r.results()
    r.close()
    ctx.close()


def test_co_argument_checks(single):
    L, rin, rout, _, _ = single
    with pytest.raises(sf.PthError, match="fused form"):
        sf.Context(0, rcut2=R_OUT ** 2, changeover=1, exec_mode=sf.EXEC_PLAIN)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, changeover=1)
    # more than 1024 i in a walk
    with pytest.raises(sf.PthError, match="1024"):
        sf.Resident(ctx, [1100], [1100], [0], np.zeros((1100, 3)), np.full(1100, 0.01), np.arange(1100), [],
                    np.zeros((1100, 3)), np.ones(1100), np.full(1100, 0.01), [], [], [])
    # a run before the changeover data
    r = sf.Resident(ctx, [2], [2], [0], np.zeros((2, 3)), np.full(2, 0.01), np.arange(2), [], np.zeros((2, 3)),
                    np.ones(2), np.full(2, 0.01), [], [], [])
    with pytest.raises(sf.PthError, match="pth_resident_changeover first"):
        r.run()
    r.close()
    ctx.close()


def _group_types(L, seed=5):
    """Group data (mass_backup, status) for the lists' EPJ, in tree order, and
    the tree masses PeTar's soft force then sees: 10% group members whose c.m.
    carries their mass (tree mass 0, backup = mass, status -(c.m. address)),
    4% members of groups without artificial particles (mass kept, status
    -LARGE_FLOAT), 6% artificial particles (status > 0), the rest singles."""
    rng = np.random.default_rng(seed)
    n = L.epj_mass.size
    u = rng.random(n)
    grp = np.zeros((n, 2))
    mass = L.epj_mass.copy()
    mem = u < 0.10
    grp[mem, 0] = mass[mem]
    grp[mem, 1] = -(1.0 + rng.integers(0, n, int(mem.sum())))
    mass[mem] = 0.0
    nocm = (u >= 0.10) & (u < 0.14)
    grp[nocm, 0] = mass[nocm]
    grp[nocm, 1] = -float(np.finfo(np.float32).max) / 16
    art = (u >= 0.14) & (u < 0.20)
    grp[art, 0] = 0.0
    grp[art, 1] = 2.0 + rng.integers(0, 8, int(art.sum()))
    return grp, mass


def test_co_group_members_and_artificial_particles():
    """Particle types of PeTar's groups in changeover mode
    (src/hard.hpp:1453-1472 per neighbour, :1669 for the self term):
    members with zero tree mass add their backup mass's changeover-weighted
    potential, artificial particles leave no potential, and only singles and
    members without a c.m. address lose their self term; against the FP64
    oracle with the same group data, at the changeover tolerances."""
    L, rin, rout = co_lists(16384, seed=9, f_bin=0.5)
    grp, mass = _group_types(L)
    L.epj_mass[:] = mass
    a = co_args(L, rin, rout)
    ref = O.soft_force_walks_co(**a, eps2=0.0, rout2=R_OUT ** 2, amb_rel=AMB_REL, epj_group=grp)
    single = O.soft_force_walks_co(**a, eps2=0.0, rout2=R_OUT ** 2, amb_rel=AMB_REL)
    amb = O.nnb_ambiguous(**{k: a[k] for k in ("n_epi", "n_epj", "epi_pos", "epi_rs", "id_epj", "epj_pos",
                                              "epj_rs")}, centres=walk_centres(a), eps2=0.0)
    # the types change the potentials of many i (not a vacuous comparison)
    assert np.mean(rel_err(ref[2], single[2]) > 1e-6) > 0.1 and np.mean(rel_err(ref[1], single[1]) > 1e-6) > 0.05
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, changeover=1)
    r = sf.Resident.from_lists(ctx, L, rin, rout, grp)
    r.run()
    res = r.results()
    r.close()
    ctx.close()
    _check(res, ref, amb, "co group types")
