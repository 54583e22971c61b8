"""GPU parity of the soft-force path against the FP64 oracle (same lists).

Tolerances (FP32 pair arithmetic on walk-relative coordinates, FP64 sums):
median relative error of acc and pot <= 1e-6, 99th percentile <= 1e-5,
maximum <= 1e-3 (relative errors of tiny net accelerations); neighbour
counts exact except for pairs inside the FP32 rounding band of r_search,
where they may differ by at most the number of such pairs.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import host, soft_force as sf
from tests._common import R_OUT, c1_lists, nnb_check, rel_err, walk_centres

pytestmark = pytest.mark.gpu

MED, P99, MAX = 1e-6, 1e-5, 1e-3


@pytest.fixture(scope="module")
def c1():
    L = c1_lists(16384, seed=1)
    a = O.lists_args(L)
    ref = O.soft_force_walks(**a, eps2=0.0, rout2=R_OUT ** 2, G=1.0, nnb_mode=0)
    amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
    return L, a, ref, amb


def _check(res, ref, amb, label):
    acc, pot, nnb = res
    racc, rpot, rnnb = ref
    ea, ep = rel_err(acc, racc), rel_err(pot, rpot)
    print(f"[{label}] acc err med {np.median(ea):.3e} p99 {np.percentile(ea, 99):.3e} max {ea.max():.3e}; "
          f"pot err med {np.median(ep):.3e} p99 {np.percentile(ep, 99):.3e} max {ep.max():.3e}", flush=True)
    assert np.median(ea) <= MED and np.median(ep) <= MED
    assert np.percentile(ea, 99) <= P99 and np.percentile(ep, 99) <= P99
    assert ea.max() <= MAX and ep.max() <= MAX
    bad_exact, bad_band, n_amb_i = nnb_check(nnb, rnnb, amb)
    print(f"[{label}] nnb: mismatches outside band {bad_exact}, beyond band {bad_band}, i with ambiguous pairs "
          f"{n_amb_i}", flush=True)
    assert bad_exact == 0 and bad_band == 0


@pytest.mark.parametrize("exec_mode", [sf.EXEC_FUSED, sf.EXEC_PLAIN])
def test_resident_c1(c1, exec_mode):
    L, a, ref, amb = c1
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, nnb_mode=0, exec_mode=exec_mode)
    r = sf.Resident.from_lists(ctx, L)
    r.run()
    res = r.results()
    _check(res, ref, amb, f"resident exec={exec_mode}")
    # a second evaluation gives bit-identical results
    r.run()
    res2 = r.results()
    assert np.array_equal(res[0], res2[0]) and np.array_equal(res[1], res2[1]) and np.array_equal(res[2], res2[2])
    r.close()
    ctx.close()


@pytest.mark.parametrize("max_ipl", [1, 2, 3, 4])
def test_fused_split_chunks(c1, max_ipl):
    """Small chunk size forces every walk's list over many items + reduction,
    for each number of i-particles per lane."""
    L, a, ref, amb = c1
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, chunk_j=256, max_ipl=max_ipl)
    r = sf.Resident.from_lists(ctx, L)
    assert r.stats()["n_items"] > L.n_walk * 2
    r.run()
    _check(r.results(), ref, amb, f"fused chunk=256 max_ipl={max_ipl}")
    r.close()
    ctx.close()


@pytest.mark.parametrize("max_ipl", [1, 2, 3, 4])
def test_ipl_variants_agree(c1, max_ipl):
    """Unsplit walks: every tiling gives results within FP32 rounding of each other."""
    L, a, ref, amb = c1
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, chunk_j=1 << 30, max_ipl=max_ipl)
    r = sf.Resident.from_lists(ctx, L)
    r.run()
    _check(r.results(), ref, amb, f"unsplit max_ipl={max_ipl}")
    r.close()
    ctx.close()


@pytest.mark.parametrize("j_vgpr", [1])
@pytest.mark.parametrize("max_ipl", [1, 2, 4])
def test_j_record_modes_are_bit_identical(c1, j_vgpr, max_ipl):
    """The j-record modes (SGPR operands, VGPR copies) compute the same FP32
    operations in the same order: bit-identical results, also with split
    walks."""
    L = c1[0]
    out = []
    for jm in (0, j_vgpr):
        ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, chunk_j=512, max_ipl=max_ipl, j_vgpr=jm)
        r = sf.Resident.from_lists(ctx, L)
        r.run()
        out.append(r.results())
        r.close()
        ctx.close()
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def _epj_records(L):
    """EPJSoft records of a host list set (ids = tree positions)."""
    n = L.epj_pos.shape[0]
    epj = np.zeros(n, sf.EPJ_SOFT)
    epj["id"] = np.arange(n)
    epj["mass"] = L.epj_mass
    epj["pos"] = L.epj_pos
    epj["r_search"] = L.epj_rs
    epj["r_out"] = R_OUT
    return epj


def _spj_records(L):
    """SPJQuadrupoleInAndOut records of a host list set."""
    spj = np.zeros(L.spj_mass.size, sf.SPJ_SOFT)
    spj["mass"] = L.spj_mass
    spj["pos"] = L.spj_pos
    for k, q in enumerate(("xx", "yy", "zz", "xy", "xz", "yz")):
        spj["quad"][q] = L.spj_quad[:, k]
    return spj


def _fdps_replay(L, functor, n_walk_limit=200, send_j=True):
    """Replays FDPS's multi-walk call order (send once, then dispatch/retrieve
    per batch of <= n_walk_limit walks) with FDPS-layout records."""
    epj, spj = _epj_records(L), _spj_records(L)
    n, ns = epj.size, spj.size
    eoff, soff = L.walk_offsets()
    if send_j:
        functor(0, 0, [], [], [], [], [], [], epj, n, spj, ns, True)
    out = []
    for b0 in range(0, L.n_walk, n_walk_limit):
        ws = range(b0, min(b0 + n_walk_limit, L.n_walk))
        epi, force = [], []
        for w in ws:
            e = np.zeros(L.n_epi[w], sf.EPI_SOFT)
            sl = slice(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w])
            e["pos"] = L.epj_pos[sl]
            e["r_search"] = L.epj_rs[sl]
            e["type"] = 1
            epi.append(e)
            f = np.zeros(L.n_epi[w], sf.FORCE_SOFT)
            f["pot"] = 123.0  # must be overwritten (assignment semantics)
            force.append(f)
        ide = [L.id_epj[eoff[w]:eoff[w + 1]] for w in ws]
        ids = [L.id_spj[soff[w]:soff[w + 1]] for w in ws]
        nw = len(ws)
        functor(0, nw, epi, L.n_epi[list(ws)], ide, L.n_epj[list(ws)], ids, L.n_spj[list(ws)], epj, n, spj, ns,
                False)
        sf.RetrieveForce(0, nw, L.n_epi[list(ws)], force, functor)
        out.extend(force)
    f = np.concatenate(out)
    return f["acc"], f["pot"], f["n_ngb"]


@pytest.mark.parametrize("eps", [0.1, 0.5, 2.0])
def test_softened_c1(c1, eps):
    """Plummer softening eps = 0.1, 0.5 and 2 r_out: eps^2 enters every EPJ
    pair, the SPJ monopole and the softened quadrupole term
    (src/soft_force.hpp:44-60,132-140); at eps = 2 r_out every EPJ pair is
    outside the cutoff."""
    L, a, _, _ = c1
    eps2 = (eps * R_OUT) ** 2
    ref = O.soft_force_walks(**a, eps2=eps2, rout2=R_OUT ** 2, G=1.0, nnb_mode=0)
    amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=eps2)
    for mode in (sf.EXEC_FUSED, sf.EXEC_PLAIN):
        ctx = sf.Context(0, eps2=eps2, rcut2=R_OUT ** 2, G=1.0, nnb_mode=0, exec_mode=mode)
        r = sf.Resident.from_lists(ctx, L)
        r.run()
        _check(r.results(), ref, amb, f"softened eps={eps} r_out exec={mode}")
        r.close()
        ctx.close()


def test_translated_cluster():
    """The cluster moved to (8000, -300, 20), a star cluster on a galactic
    orbit in pc with r_out = 0.01 pc: absolute FP32 coordinates (spacing
    4.9e-4 at 8000) could not resolve the cutoff, walk-relative ones (the
    walk centre subtracted in FP64) keep the errors of the untranslated
    cluster."""
    m, x, _ = host.cluster(16384, "plummer", kroupa=True, f_bin=0.0, seed=12)
    L = host.build_lists(x + np.array([8000.0, -300.0, 20.0]), m, 1.5 * R_OUT, theta=0.3)
    a = O.lists_args(L)
    ref = O.soft_force_walks(**a, eps2=0.0, rout2=R_OUT ** 2, G=1.0, nnb_mode=0)
    amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
    for mode in (sf.EXEC_FUSED, sf.EXEC_PLAIN):
        ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, nnb_mode=0, exec_mode=mode)
        r = sf.Resident.from_lists(ctx, L)
        r.run()
        _check(r.results(), ref, amb, f"translated to x = 8000 exec={mode}")
        r.close()
        ctx.close()


@pytest.mark.parametrize("exec_mode", [sf.EXEC_FUSED, sf.EXEC_PLAIN])
def test_fdps_multiwalk_replay(c1, exec_mode):
    L, a, ref, amb = c1
    fn = sf.make_functor(0, 0.0, R_OUT ** 2, 1.0, exec_mode=exec_mode)
    res = _fdps_replay(L, fn)
    _check(res, ref, amb, f"fdps replay exec={exec_mode}")
    # the drop-in path and the resident path compute every walk identically
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, exec_mode=exec_mode)
    r = sf.Resident.from_lists(ctx, L)
    r.run()
    racc, rpot, rnnb = r.results()
    assert np.array_equal(res[0], racc) and np.array_equal(res[1], rpot) and np.array_equal(res[2], rnnb)
    r.close()
    ctx.close()


def test_nnb_mode_le_matches_reference_kernel(c1):
    """nnb_mode 1 is the x86 kernel's r2 <= rs^2: compare with the reference
    phantomquad outputs (built from source) outside the rounding band of
    both FP32 evaluations."""
    L, a, _, _ = c1
    if not O.ref_available():
        pytest.skip("reference kernel not built")
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, nnb_mode=1)
    r = sf.Resident.from_lists(ctx, L)
    r.run()
    acc, pot, nnb = r.results()
    racc, rpot, rnnb = O.ref_soft_force_walks(**a, eps2=0.0, rout2=R_OUT ** 2)
    amb_rel = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
    amb_abs = O.nnb_ambiguous(**a, centres=None, eps2=0.0)
    bad_exact, bad_band, _ = nnb_check(nnb, rnnb, amb_rel + amb_abs)
    assert bad_exact == 0 and bad_band == 0
    # forces agree to the reference kernel's own rsqrt accuracy
    # (SPJ terms use the raw hardware estimate, src/petar.hpp:12-15)
    ea = rel_err(acc, racc)
    assert np.median(ea) < 1e-4
    r.close()
    ctx.close()


def test_edge_cases():
    """Empty walks, walks with empty lists, the self pair, zero-mass j and
    a list longer than one chunk."""
    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.normal(size=(n, 3)) * 0.02
    mass = rng.random(n) / n
    mass[::17] = 0.0
    rs = np.full(n, 0.015)
    n_epi = np.array([0, 1, 5, 64, 65, 200, 3, 0], np.int32)
    n_epj = np.array([10, 1, 0, 2999, 700, 3000, 5, 0], np.int32)
    n_spj = np.array([0, 0, 4, 0, 33, 100, 0, 7], np.int32)
    ni = int(n_epi.sum())
    ipick = rng.integers(0, n, ni)
    # walk 1 is the self pair only
    off = np.concatenate([[0], np.cumsum(n_epi)])
    ide = [rng.permutation(n)[:k] for k in n_epj]
    ide[1] = ipick[off[1]:off[1] + 1]
    id_epj = np.concatenate(ide).astype(np.int32)
    nsp = 50
    spos = rng.normal(size=(nsp, 3)) * 3 + 5
    smass = rng.random(nsp) * 1e-3
    squad = rng.random((nsp, 6)) * 1e-4
    id_spj = np.concatenate([rng.integers(0, nsp, k) for k in n_spj]).astype(np.int32)
    args = dict(n_epi=n_epi, n_epj=n_epj, n_spj=n_spj, epi_pos=pos[ipick], epi_rs=rs[ipick], id_epj=id_epj,
                id_spj=id_spj, epj_pos=pos, epj_mass=mass, epj_rs=rs, spj_pos=spos, spj_mass=smass, spj_quad=squad)
    G = 2.5
    ref = O.soft_force_walks(**args, eps2=1e-8, rout2=R_OUT ** 2, G=G)
    amb = O.nnb_ambiguous(**args, centres=walk_centres(args), eps2=1e-8)
    for mode in (sf.EXEC_FUSED, sf.EXEC_PLAIN):
        for cj in (0, 64):
            ctx = sf.Context(0, eps2=1e-8, rcut2=R_OUT ** 2, G=G, exec_mode=mode, chunk_j=cj)
            r = sf.Resident(ctx, **args)
            r.run()
            acc, pot, nnb = r.results()
            nz = np.linalg.norm(ref[0], axis=1) > 0
            assert np.all(rel_err(acc[nz], ref[0][nz]) < 1e-5)
            assert np.all(acc[~nz] == 0)
            pz = ref[1] != 0
            assert np.all(rel_err(pot[pz], ref[1][pz]) < 1e-5)
            bad_exact, bad_band, _ = nnb_check(nnb, ref[2], amb)
            assert bad_exact == 0 and bad_band == 0
            # self pair: pot = -G m / r_out (softened by eps), acc = 0, nnb = 1
            i1 = off[1]
            m1 = mass[ipick[i1]]
            assert abs(pot[i1] - (-G * m1 / np.sqrt(max(1e-8, R_OUT ** 2)))) <= 1e-6 * abs(pot[i1]) + 1e-300
            assert nnb[i1] == 1 and np.all(acc[i1] == 0)
            r.close()
            ctx.close()


def test_functors_share_one_context(c1):
    """Functors with different parameters use the rank's one context (its
    parameters updated by pth_set_params): j arrays sent through one serve
    the dispatch of another, and the dispatch uses the dispatching functor's
    parameters (here G)."""
    L, a, ref, amb = c1
    sf.CalcForceWithLinearCutoffMultiWalk.close_contexts()
    send = sf.make_functor(0, 0.0, R_OUT ** 2, 1.0)
    send(0, 0, [], [], [], [], [], [], _epj_records(L), L.epj_mass.size, _spj_records(L), L.spj_mass.size, True)
    res = _fdps_replay(L, sf.make_functor(0, 0.0, R_OUT ** 2, 2.0), send_j=False)
    assert send.context() is sf.make_functor(0, 0.0, R_OUT ** 2, 2.0).context()
    _check((res[0] / 2.0, res[1] / 2.0, res[2]), ref, amb, "functor with G = 2 on j arrays sent with G = 1")
    sf.CalcForceWithLinearCutoffMultiWalk.close_contexts()


def test_error_paths():
    sf.CalcForceWithLinearCutoffMultiWalk.close_contexts()
    ctx = sf.Context(0, rcut2=1e-4)
    fn = sf.CalcForceWithLinearCutoffMultiWalk(0, 0.0, 1e-4, 1.0)
    # retrieve before any dispatch
    with pytest.raises(sf.PthError):
        fn.retrieve(0, 1, [1], [np.zeros(1, sf.FORCE_SOFT)])
    # dispatch before send
    fresh = sf.CalcForceWithLinearCutoffMultiWalk(0, 0.0, 2e-4, 1.0)
    with pytest.raises(sf.PthError, match="no j arrays"):
        fresh(0, 1, [np.zeros(1, sf.EPI_SOFT)], [1], [np.zeros(1, np.int32)], [1], [np.zeros(0, np.int32)], [0],
              None, 0, None, 0, False)
    # list index out of range is rejected before any kernel runs
    epj = np.zeros(4, sf.EPJ_SOFT)
    spj = np.zeros(0, sf.SPJ_SOFT)
    fn(0, 0, [], [], [], [], [], [], epj, 4, spj, 0, True)
    with pytest.raises(sf.PthError, match="out of range"):
        fn(0, 1, [np.zeros(2, sf.EPI_SOFT)], [2], [np.array([0, 9], np.int32)], [2], [np.zeros(0, np.int32)], [0],
           epj, 4, spj, 0, False)
    # too many walks for one dispatch
    with pytest.raises(sf.PthError, match="n_walk"):
        fn(0, 1001, [np.zeros(1, sf.EPI_SOFT)] * 1001, [1] * 1001, [np.zeros(1, np.int32)] * 1001, [1] * 1001,
           [np.zeros(0, np.int32)] * 1001, [0] * 1001, epj, 4, spj, 0, False)
    ctx.close()


@pytest.mark.parametrize("list_runs", ["0", "1", None])
def test_dispatch_rejects_bad_lists_in_every_list_form(list_runs, monkeypatch):
    """pth_dispatch validates every index of a batch's lists whether it
    sends them as they are (PTH_LIST_RUNS=0), always as runs (=1) or as
    runs when that halves their bytes (unset): an index outside the sent
    arrays, negative, or in the SPJ list, fails the call before any kernel
    runs, and the context serves the next batch."""
    if list_runs is None:
        monkeypatch.delenv("PTH_LIST_RUNS", raising=False)
    else:
        monkeypatch.setenv("PTH_LIST_RUNS", list_runs)
    sf.CalcForceWithLinearCutoffMultiWalk.close_contexts()
    try:
        fn = sf.CalcForceWithLinearCutoffMultiWalk(0, 0.0, 1e-4, 1.0)
        epj = np.zeros(64, sf.EPJ_SOFT)
        epj["mass"] = 1.0
        epj["pos"][:, 0] = np.arange(64)
        spj = np.zeros(3, sf.SPJ_SOFT)
        spj["mass"] = 1.0
        spj["pos"][:, 1] = 100.0 + np.arange(3)
        fn(0, 0, [], [], [], [], [], [], epj, 64, spj, 3, True)
        epi = [np.zeros(2, sf.EPI_SOFT), np.zeros(3, sf.EPI_SOFT)]
        good_e = [np.arange(0, 40, dtype=np.int32), np.arange(10, 64, dtype=np.int32)]
        good_s = [np.array([0, 2], np.int32), np.array([1], np.int32)]
        ne, nj, ns = [2, 3], [40, 54], [2, 1]
        for bad_e, bad_s in (([good_e[0], np.r_[good_e[1][:-1], 64].astype(np.int32)], good_s),
                             ([np.r_[-1, good_e[0][1:]].astype(np.int32), good_e[1]], good_s),
                             (good_e, [good_s[0], np.array([3], np.int32)])):
            with pytest.raises(sf.PthError, match="out of range"):
                fn(0, 2, epi, ne, bad_e, nj, bad_s, ns, epj, 64, spj, 3, False)
        fn(0, 2, epi, ne, good_e, nj, good_s, ns, epj, 64, spj, 3, False)
        force = [np.zeros(2, sf.FORCE_SOFT), np.zeros(3, sf.FORCE_SOFT)]
        sf.RetrieveForce(0, 2, ne, force, fn)
        pot = np.concatenate([f["pot"] for f in force])
        assert np.all(np.isfinite(pot)) and np.all(pot < 0)
    finally:
        sf.CalcForceWithLinearCutoffMultiWalk.close_contexts()


def test_replan_and_autotune(c1):
    """A re-planned resident evaluation equals one created with the same
    tiling parameters; the autotuner keeps one of the candidates and the
    result stays within the parity bar."""
    L, a, ref, amb = c1
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    r = sf.Resident.from_lists(ctx, L)
    r.replan(1, 512)
    r.run()
    res = r.results()
    ctx2 = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, max_ipl=1, chunk_j=512)
    r2 = sf.Resident.from_lists(ctx2, L)
    r2.run()
    assert all(np.array_equal(x, y) for x, y in zip(res, r2.results()))
    r2.close()
    ctx2.close()
    ipl, cj, ms = r.autotune(ipl=(1, 2, 3, 4), chunk=(512, 2048), reps=2)
    print(f"[autotune c1] max_ipl {ipl} chunk_j {cj}: {ms:.3f} ms", flush=True)
    assert ipl in (1, 2, 3, 4) and cj in (512, 2048) and ms > 0
    r.run()
    _check(r.results(), ref, amb, "autotuned")
    r.close()
    ctx.close()


def test_theta0_is_direct_summation():
    """theta = 0 turns every j into an EPJ, so the GPU result is the direct
    sum with the linear cutoff: compare with the FP64 all-pairs oracle
    (the small plain configuration, 16,384 particles, 2.7e8 pairs)."""
    L = c1_lists(16384, theta=0.0, seed=1)
    assert L.n_spj.sum() == 0 and np.all(L.n_epj == L.epj_mass.size)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    r = sf.Resident.from_lists(ctx, L)
    r.run()
    acc, pot, nnb = r.results()
    r.close()
    ctx.close()
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(L.epi_begin, L.n_epi)])
    dacc, dpot, dnnb = O.direct_linear_cutoff(L.epj_pos, L.epj_mass, L.epj_rs, sample=idx, eps2=0.0,
                                              rout2=R_OUT ** 2)
    ea, ep = rel_err(acc, dacc), rel_err(pot, dpot)
    print(f"[theta=0 direct] acc med {np.median(ea):.3e} p99 {np.percentile(ea, 99):.3e} max {ea.max():.3e}; "
          f"pot med {np.median(ep):.3e}; nnb mismatches {(nnb != dnnb).sum()}", flush=True)
    assert np.median(ea) <= MED and np.percentile(ea, 99) <= P99 and ea.max() <= MAX
    assert np.median(ep) <= MED and np.percentile(ep, 99) <= P99
    # every j is an EPJ: the band of the walk-relative FP32 pair arithmetic
    # covers every pair the direct sum counts
    a = O.lists_args(L)
    amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
    bad_exact, bad_band, _ = nnb_check(nnb, dnnb, amb)
    assert bad_exact == 0 and bad_band == 0


@pytest.mark.parametrize("segments,streams", [(1, 1), (4, 2), (8, 4), (16, 4), (32, 8)])
def test_segments_are_bit_identical(c1, segments, streams):
    """Walk segments run concurrently on several streams; every walk keeps
    its items, so the results equal the one-stream evaluation bit for bit
    (min_seg_pairs = 1: this small evaluation is cut into every segment asked
    for)."""
    L, a, ref, amb = c1
    out = []
    for seg, st in ((1, 1), (segments, streams)):
        ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, segments=seg, streams=st, min_seg_pairs=1)
        r = sf.Resident.from_lists(ctx, L)
        r.run()
        r.run()
        out.append(r.results())
        r.close()
        ctx.close()
    assert all(np.array_equal(x, y) for x, y in zip(out[0], out[1]))
    _check(out[1], ref, amb, f"segments={segments} streams={streams}")


@pytest.mark.parametrize("exec_mode", [sf.EXEC_FUSED, sf.EXEC_PLAIN])
def test_monopole_superparticles(c1, exec_mode):
    """spj_monopole = 1 (a PeTar build without USE_QUAD): the superparticles'
    mass and centre only, the monopole pair of src/soft_force.hpp:125-158,
    against the FP64 oracle's monopole restatement; the evaluation fed from
    the device tree equals the one fed from the host lists bit for bit."""
    L, a, _, amb = c1
    ref = O.soft_force_walks(**a, eps2=0.0, rout2=R_OUT ** 2, mono=True)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, exec_mode=exec_mode, spj_monopole=1)
    r = sf.Resident(ctx, L.n_epi, L.n_epj, L.n_spj, a["epi_pos"], a["epi_rs"], L.id_epj, L.id_spj, L.epj_pos,
                    L.epj_mass, L.epj_rs, L.spj_pos, L.spj_mass, np.zeros(0))
    r.run()
    res = r.results()
    r.close()
    _check(res, ref, amb, f"monopole exec={exec_mode}")
    if exec_mode == sf.EXEC_FUSED:
        m = np.zeros(L.epj_mass.size)
        m[L.order] = L.epj_mass
        x = np.zeros_like(L.epj_pos)
        x[L.order] = L.epj_pos
        g = sf.GpuTree(ctx, x, m, np.full(m.size, 1.5 * R_OUT), theta=0.3)
        rd = sf.Resident.from_gtree(ctx, g)
        rd.run()
        dev = rd.results()
        rd.close()
        g.close()
        assert all(np.array_equal(p, q) for p, q in zip(dev, res))
    ctx.close()


def test_double_float_near_pairs():
    """Linear-cutoff mode with df_near = 1: near EPJ pairs take their
    separation from double-float walk-relative coordinates.  On a cluster of
    tight binaries moved far from the origin the acc error against the FP64
    oracle stays at FP32 rounding (max <= 1e-5), where the FP32-only
    separations (df_near = 0) leave an error of 2^-24 |x - c| in each close
    pair's separation; the neighbour counts stay exact outside the band."""
    m, x, _ = host.cluster(16384, "plummer", kroupa=True, f_bin=0.5, seed=21)
    L = host.build_lists(x + np.array([300.0, -20.0, 5.0]), m, 1.5 * R_OUT, theta=0.3)
    a = O.lists_args(L)
    ref = O.soft_force_walks(**a, eps2=0.0, rout2=R_OUT ** 2)
    amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
    err = {}
    for df in (0, 1):
        ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, df_near=df)
        r = sf.Resident.from_lists(ctx, L)
        r.run()
        acc, pot, nnb = r.results()
        r.close()
        ctx.close()
        ea, ep = rel_err(acc, ref[0]), rel_err(pot, ref[1])
        err[df] = ea
        print(f"[df_near={df}] acc med {np.median(ea):.2e} p99 {np.percentile(ea, 99):.2e} max {ea.max():.2e}; "
              f"pot med {np.median(ep):.2e} max {ep.max():.2e}", flush=True)
        bad_exact, bad_band, _ = nnb_check(nnb, ref[2], amb)
        assert bad_exact == 0 and bad_band == 0
        assert np.median(ea) <= MED and np.median(ep) <= MED
    assert err[1].max() <= 1e-5 and np.percentile(err[1], 99) <= 1e-6
    assert err[1].max() < err[0].max()


@pytest.mark.parametrize("segments,streams", [(8, 3), (16, 4), (1, 1)])
def test_graph_replay_equals_launches(c1, segments, streams):
    """pth_params.graph = 1: the evaluation's launches (all segments and
    streams) captured once as a hipGraph and replayed.  Every replay, with
    eager evaluations of another context and synchronisations between the
    replays, equals the launched evaluation bit for bit; a replan (new work
    items) is captured again and still equals the launched form."""
    L, a, ref, amb = c1
    eager = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, segments=segments, streams=streams, min_seg_pairs=1)
    re = sf.Resident.from_lists(eager, L)
    re.run()
    want = re.results()
    gctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, segments=segments, streams=streams, graph=1,
                      min_seg_pairs=1)
    rg = sf.Resident.from_lists(gctx, L)
    for k in range(4):
        rg.run()
        got = rg.results()
        print(f"[graph seg={segments}] replay {k}: identical {all(np.array_equal(x, y) for x, y in zip(got, want))}",
              flush=True)
        assert all(np.array_equal(x, y) for x, y in zip(got, want))
        re.run()
        re.sync()
    _check(got, ref, amb, f"graph segments={segments}")
    for ipl, cj in ((1, 256), (4, 2048)):
        rg.replan(ipl, cj)
        re.replan(ipl, cj)
        re.run()
        want2 = re.results()
        for k in range(2):
            rg.run()
            assert all(np.array_equal(x, y) for x, y in zip(rg.results(), want2)), f"replan {ipl},{cj} replay {k}"
    ms = rg.timing()
    assert ms[1] > 0 and ms[0] == 0 and ms[2] == 0
    rg.close()
    re.close()
    gctx.close()
    eager.close()


def test_graph_replay_after_refill():
    """Graph mode on an evaluation fed from the device tree: a refill from a
    rebuilt tree (new particles) is captured again and equals the launched
    evaluation of the same tree."""
    outs = []
    for graph in (0, 1):
        ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, graph=graph)
        res = []
        g = None
        r = None
        for seed in (5, 6):
            m, x, _ = host.cluster(12000, "plummer", kroupa=True, seed=seed)
            if g is None:
                g = sf.GpuTree(ctx, x, m, np.full(m.size, 1.5 * R_OUT), theta=0.3)
                r = sf.Resident.from_gtree(ctx, g)
            else:
                g.update(x, m, np.full(m.size, 1.5 * R_OUT))
                r.refill(g)
            r.run()
            r.run()
            res.append(r.results())
        r.close()
        g.close()
        ctx.close()
        outs.append(res)
    for k in range(2):
        assert all(np.array_equal(x, y) for x, y in zip(outs[0][k], outs[1][k])), f"tree {k}"


@pytest.mark.parametrize("exec_mode", [sf.EXEC_FUSED, sf.EXEC_PLAIN])
def test_fdps_copies_functor(c1, exec_mode):
    """The functor of a build without the multi-walk index
    (CalcForceWithLinearCutoffCUDA, src/force_gpu_cuda.hpp:137-162;
    pth_dispatch_copies): every batch hands each walk copies of its EPJ and
    SPJ records.  Against the FP64 oracle, and bit-identical to the
    multi-walk-index path and the resident path (the same lists in the same
    order)."""
    L, a, ref, amb = c1
    epj_all, spj_all = _epj_records(L), _spj_records(L)
    eoff, soff = L.walk_offsets()
    fn = sf.CalcForceWithLinearCutoff(0, 0.0, R_OUT ** 2, 1.0, exec_mode=exec_mode)
    out = []
    for b0 in range(0, L.n_walk, 200):
        ws = list(range(b0, min(b0 + 200, L.n_walk)))
        epi, force, epj, spj = [], [], [], []
        for w in ws:
            e = np.zeros(L.n_epi[w], sf.EPI_SOFT)
            sl = slice(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w])
            e["pos"] = L.epj_pos[sl]
            e["r_search"] = L.epj_rs[sl]
            epi.append(e)
            f = np.zeros(L.n_epi[w], sf.FORCE_SOFT)
            f["pot"] = 123.0
            force.append(f)
            epj.append(epj_all[L.id_epj[eoff[w]:eoff[w + 1]]])
            spj.append(spj_all[L.id_spj[soff[w]:soff[w + 1]]])
        fn(0, len(ws), epi, L.n_epi[ws], epj, L.n_epj[ws], spj, L.n_spj[ws])
        sf.RetrieveForce(0, len(ws), L.n_epi[ws], force, fn)
        out.extend(force)
    f = np.concatenate(out)
    res = (f["acc"], f["pot"], f["n_ngb"])
    _check(res, ref, amb, f"copies functor exec={exec_mode}")
    idx_res = _fdps_replay(L, sf.make_functor(0, 0.0, R_OUT ** 2, 1.0, exec_mode=exec_mode))
    assert all(np.array_equal(x, y) for x, y in zip(res, idx_res))


def test_persistent_launches_are_bit_identical(c1):
    """pth_params.persistent = 1: each i-per-lane class runs as one wave per
    free wave slot, the waves of each XCD taking that XCD's items from a
    counter.  Every item is computed as in the launch of one wave per item,
    so the results are the same bit for bit: resident evaluations over 8
    segments on 3 streams at several tilings (double-float near pairs, plain
    FP32 near pairs, changeover mode), a captured graph replayed, and the
    FDPS multi-walk replay through pth_dispatch (1-i items on the second
    stream)."""
    L, a, ref, amb = c1
    rin, rout = np.full(L.epj_mass.size, 0.1 * R_OUT), np.full(L.epj_mass.size, R_OUT)
    tiles = ((2, 1024), (1, 256), (3, 2048), (4, 512))
    for mode in ("df", "fp32", "changeover"):
        outs = []
        for pers in (0, 1):
            kw = dict(segments=8, streams=3, min_seg_pairs=1, persistent=pers)
            if mode == "changeover":
                ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, changeover=1, **kw)
                r = sf.Resident.from_lists(ctx, L, rin, rout)
            else:
                ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, df_near=int(mode == "df"), **kw)
                r = sf.Resident.from_lists(ctx, L)
            got = []
            for ipl, cj in tiles:
                r.replan(ipl, cj)
                r.run()
                r.run()
                got.append(r.results())
            outs.append(got)
            r.close()
            ctx.close()
        for (ipl, cj), x, y in zip(tiles, outs[0], outs[1]):
            same = all(np.array_equal(u, v) for u, v in zip(x, y))
            print(f"[persistent {mode}] ipl {ipl} chunk {cj}: identical {same}", flush=True)
            assert same, f"{mode} ipl {ipl} chunk {cj}"
        if mode != "changeover":
            _check(outs[1][0], ref, amb, f"persistent {mode}")
    # a captured graph of persistent launches, replayed
    eager = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, segments=8, streams=3, min_seg_pairs=1)
    re = sf.Resident.from_lists(eager, L)
    re.run()
    want = re.results()
    gctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, segments=8, streams=3, min_seg_pairs=1, graph=1,
                      persistent=1)
    rg = sf.Resident.from_lists(gctx, L)
    for k in range(3):
        rg.run()
        assert all(np.array_equal(x, y) for x, y in zip(rg.results(), want)), f"graph replay {k}"
        re.run()
        re.sync()
    rg.close()
    re.close()
    gctx.close()
    eager.close()
    # the drop-in path
    fn = sf.make_functor(0, 0.0, R_OUT ** 2, 1.0)
    base = _fdps_replay(L, fn)
    c = fn.context()
    p = c.get_params()
    p.persistent = 1
    c.set_params(p)
    try:
        got = _fdps_replay(L, fn)
    finally:
        p.persistent = 0
        c.set_params(p)
    assert all(np.array_equal(x, y) for x, y in zip(base, got))


def test_single_segment_class_streams_are_bit_identical(c1, monkeypatch):
    """A resident evaluation of one segment runs its 1-i items on a side
    stream beside the longer classes (PTH_CLASS_STREAMS=0: one stream), in
    launched and in graph mode: the same results bit for bit, at several
    tilings."""
    L, a, ref, amb = c1
    outs = []
    for env, graph in (("0", 0), ("1", 0), ("1", 1)):
        monkeypatch.setenv("PTH_CLASS_STREAMS", env)
        ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, segments=1, graph=graph)
        r = sf.Resident.from_lists(ctx, L)
        got = []
        for ipl, cj in ((2, 1024), (1, 256), (3, 2048)):
            r.replan(ipl, cj)
            r.run()
            r.run()
            got.append(r.results())
        ms = r.timing()
        print(f"[class streams {env} graph {graph}] force ms {ms[1]:.3f}", flush=True)
        outs.append(got)
        r.close()
        ctx.close()
    monkeypatch.delenv("PTH_CLASS_STREAMS", raising=False)
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert all(np.array_equal(u, v) for u, v in zip(x, y))
    _check(outs[1][0], ref, amb, "single segment, class streams")


def test_two_tags_in_flight_are_bit_identical(c1):
    """Batches dispatched on tags 0 and 1 before either is retrieved, then
    retrieved in the other order (each tag has its own result buffers, which
    the kernels write directly, and the batches share the context's stream
    and its second stream): the same forces bit for bit as the one-tag
    replay of the same batches."""
    L, a, ref, amb = c1
    fn = sf.make_functor(0, 0.0, R_OUT ** 2, 1.0)
    want = _fdps_replay(L, fn)
    epj, spj = _epj_records(L), _spj_records(L)
    n, ns = epj.size, spj.size
    eoff, soff = L.walk_offsets()
    fn(0, 0, [], [], [], [], [], [], epj, n, spj, ns, True)

    def batch(ws):
        epi, force = [], []
        for w in ws:
            e = np.zeros(L.n_epi[w], sf.EPI_SOFT)
            sl = slice(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w])
            e["pos"] = L.epj_pos[sl]
            e["r_search"] = L.epj_rs[sl]
            e["type"] = 1
            epi.append(e)
            force.append(np.zeros(L.n_epi[w], sf.FORCE_SOFT))
        ide = [L.id_epj[eoff[w]:eoff[w + 1]] for w in ws]
        ids = [L.id_spj[soff[w]:soff[w + 1]] for w in ws]
        return epi, force, ide, ids

    out = [None] * L.n_walk
    for b0 in range(0, L.n_walk, 400):
        wa = list(range(b0, min(b0 + 200, L.n_walk)))
        wb = list(range(b0 + 200, min(b0 + 400, L.n_walk)))
        ea, fa, ia, sa = batch(wa)
        fn(0, len(wa), ea, L.n_epi[wa], ia, L.n_epj[wa], sa, L.n_spj[wa], epj, n, spj, ns, False)
        if wb:
            eb, fb, ib, sb = batch(wb)
            fn(1, len(wb), eb, L.n_epi[wb], ib, L.n_epj[wb], sb, L.n_spj[wb], epj, n, spj, ns, False)
            sf.RetrieveForce(1, len(wb), L.n_epi[wb], fb, fn)
            for w, f in zip(wb, fb):
                out[w] = f
        sf.RetrieveForce(0, len(wa), L.n_epi[wa], fa, fn)
        for w, f in zip(wa, fa):
            out[w] = f
    f = np.concatenate(out)
    got = (f["acc"], f["pot"], f["n_ngb"])
    assert all(np.array_equal(x, y) for x, y in zip(want, got))
