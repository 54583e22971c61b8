"""The tree built on the device (pth_gtree_*) against the host tree
(petar_amd.host.build_lists, the FDPS-semantics stand-in the lists of every
other test come from): same particle order, same nodes, bit-identical
moments, same i-groups and, per group, the same EPJ and SPJ entries (entry
for entry in canonical order, as sets in batch order); and a force
evaluation fed from the device tree (pth_resident_from_gtree) equals the
one fed from the host lists bit for bit and the FP64 oracle within the
kernel tolerance."""
import ctypes as C
import time

import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import host, soft_force as sf
from tests._common import R_OUT, rel_err

pytestmark = pytest.mark.gpu


def _compare(n, seed, f_bin, theta, leaf=20, group=512, r_search=1.5 * R_OUT, model="plummer", canonical=True,
             keep=False, mx=None):
    m, x = mx if mx is not None else host.cluster(n, model, kroupa=True, f_bin=f_bin, seed=seed)[:2]
    t0 = time.time()
    L = host.build_lists(x, m, r_search, theta=theta, n_leaf_limit=leaf, n_group_limit=group)
    t_host = time.time() - t0
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    g = sf.GpuTree(ctx, x, m, np.full(n, r_search), theta=theta, n_leaf_limit=leaf, n_group_limit=group,
                   canonical=canonical)
    d = g.download()
    tb, tw, to = g.timing()
    if not keep:
        g.close()
        ctx.close()
    print(f"[gtree N={n} theta={theta} canonical={canonical}] host tree+walk {t_host * 1e3:.1f} ms, device build "
          f"{tb:.2f} ms walk count {tw:.2f} ms write+order {to:.2f} ms; {len(d['n_epi'])} groups, "
          f"{d['id_epj'].size} EPJ + {d['id_spj'].size} SPJ entries", flush=True)
    assert np.array_equal(d["order"], L.order)
    assert np.array_equal(d["node_begin"], L.node_begin) and np.array_equal(d["node_end"], L.node_end)
    assert np.array_equal(d["node_level"], L.node_level)
    # moments: the same FP64 sums in the same order
    assert np.array_equal(d["node_mass"], L.spj_mass)
    assert np.array_equal(d["node_com"], L.spj_pos) and np.array_equal(d["node_quad"], L.spj_quad)
    assert np.array_equal(d["epi_begin"], L.epi_begin) and np.array_equal(d["n_epi"], L.n_epi)
    assert np.array_equal(d["n_epj"], L.n_epj) and np.array_equal(d["n_spj"], L.n_spj)
    if canonical:
        assert np.array_equal(d["id_epj"], L.id_epj) and np.array_equal(d["id_spj"], L.id_spj)
    else:
        eo, so = L.walk_offsets()
        for w in range(L.n_walk):
            assert np.array_equal(np.sort(d["id_epj"][eo[w]:eo[w + 1]]), np.sort(L.id_epj[eo[w]:eo[w + 1]]))
            assert np.array_equal(np.sort(d["id_spj"][so[w]:so[w + 1]]), np.sort(L.id_spj[so[w]:so[w + 1]]))
    if keep:
        return L, d, ctx, g, m
    return L, d


@pytest.mark.parametrize("n,seed,f_bin,theta", [(16384, 1, 0.0, 0.3), (60000, 7, 0.3, 0.3), (5000, 3, 0.3, 0.0),
                                                (30000, 9, 0.5, 0.7)])
def test_device_tree_equals_host_tree(n, seed, f_bin, theta):
    _compare(n, seed, f_bin, theta)


def test_device_tree_limits():
    _compare(20000, 5, 0.3, 0.4, leaf=4, group=64)
    _compare(20000, 5, 0.3, 0.4, leaf=64, group=2048)
    _compare(20000, 5, 0.3, 0.4, leaf=1, group=1)


def test_device_tree_batch_order():
    _compare(30000, 4, 0.3, 0.3, canonical=False)


def test_device_tree_tiny():
    _compare(1, 0, 0.0, 0.3, mx=(np.ones(1), np.zeros((1, 3))))
    _compare(2, 0, 0.0, 0.3, mx=(np.ones(2), np.zeros((2, 3))))  # coincident: a zero-size cube
    for n in (2, 3, 21):
        _compare(n, 11, 0.0, 0.3)


def test_device_lists_feed_the_force_path():
    """pth_resident_from_gtree against pth_resident_create on the host lists
    (bit for bit) and against the FP64 oracle on the same lists."""
    L, d, ctx, g, _ = _compare(16384, 2, 0.3, 0.3, keep=True)
    rd = sf.Resident.from_gtree(ctx, g)
    g.close()
    rd.run()
    acc, pot, nnb = rd.results()
    rd.close()
    rh = sf.Resident.from_lists(ctx, L)
    rh.run()
    acc_h, pot_h, nnb_h = rh.results()
    rh.close()
    ctx.close()
    assert np.array_equal(acc, acc_h) and np.array_equal(pot, pot_h) and np.array_equal(nnb, nnb_h)
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(L.epi_begin, L.n_epi)])
    racc, rpot, rnnb = O.soft_force_walks(L.n_epi, L.n_epj, L.n_spj, L.epj_pos[idx], L.epj_rs[idx], L.id_epj,
                                          L.id_spj, L.epj_pos, L.epj_mass, L.epj_rs, L.spj_pos, L.spj_mass,
                                          L.spj_quad, eps2=0.0, rout2=R_OUT ** 2)
    ea, ep = rel_err(acc, racc), rel_err(pot, rpot)
    assert np.median(ea) <= 1e-6 and np.median(ep) <= 1e-6 and np.percentile(ea, 99) <= 1e-4
    assert np.mean(nnb != rnnb) <= 1e-3


def test_device_tree_translated_cluster():
    """The cluster centred at (8000, -300, 20): the device tree still equals
    the host tree (keys from the bounding box, FP64 moments in the same
    order) and its evaluation the host-list one bit for bit."""
    m, x, _ = host.cluster(16384, "plummer", kroupa=True, f_bin=0.3, seed=12)
    L, d, ctx, g, _ = _compare(16384, 12, 0.3, 0.3, keep=True, mx=(m, x + np.array([8000.0, -300.0, 20.0])))
    rd = sf.Resident.from_gtree(ctx, g)
    g.close()
    rd.run()
    res_d = rd.results()
    rd.close()
    rh = sf.Resident.from_lists(ctx, L)
    rh.run()
    res_h = rh.results()
    rh.close()
    ctx.close()
    assert all(np.array_equal(u, v) for u, v in zip(res_d, res_h))


def test_device_tree_changeover_mode():
    """Changeover mode fed from the device tree (radii and ids in tree order)
    equals the host-list evaluation bit for bit."""
    n = 12000
    m, x, _ = host.cluster(n, "plummer", kroupa=True, f_bin=0.5, seed=21)
    rout = R_OUT * np.maximum(np.cbrt(m / m.mean()), 1.0)
    rin = 0.1 * rout
    rs = 1.5 * rout
    L = host.build_lists(x, m, rs, theta=0.3)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, changeover=1)
    g = sf.GpuTree(ctx, x, m, rs, theta=0.3)
    rd = sf.Resident.from_gtree(ctx, g)
    g.close()
    rd.changeover(np.arange(n), np.arange(n), rin[L.order], rout[L.order])
    rd.run()
    out_d = rd.results()
    rd.close()
    rh = sf.Resident.from_lists(ctx, L, rin[L.order], rout[L.order])
    rh.run()
    out_h = rh.results()
    rh.close()
    ctx.close()
    for a, b in zip(out_d, out_h):
        assert np.array_equal(a, b)


def test_device_tree_large():
    """N = 262144 with 50% binaries: the lists equal the host's entry for
    entry."""
    _compare(262144, 8, 0.5, 0.3)


def _hip():
    """The HIP runtime the library is linked against (already loaded), for
    raw device buffers in tests that have no torch in the process."""
    sf.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _to_device(hip, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p.value


def test_device_tree_update_and_refill():
    """A tree rebuilt in place (larger, then smaller N; host and device
    inputs) equals a fresh build, and a refilled resident evaluation equals
    a fresh one bit for bit."""
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    hip = _hip()
    sets = []
    for n, seed in ((8000, 31), (20000, 32), (5000, 33)):
        m, x, _ = host.cluster(n, "plummer", kroupa=True, f_bin=0.3, seed=seed)
        sets.append((m, x, np.full(n, 1.5 * R_OUT)))
    g = sf.GpuTree(ctx, sets[0][1], sets[0][0], sets[0][2])
    r = sf.Resident.from_gtree(ctx, g)
    for k, (m, x, rs) in enumerate(sets[1:] + sets[:1]):
        if k % 2 == 0:
            g.update(x, m, rs)
        else:
            bufs = [_to_device(hip, a) for a in (x, m, rs)]
            assert sf.lib().pth_gtree_update(g._h, len(m), bufs[0], bufs[1], bufs[2], 1) == 0
            g.n = len(m)
            for p in bufs:
                assert hip.hipFree(C.c_void_p(p)) == 0
        d = g.download()
        r.refill(g)
        r.run()
        out = r.results()
        f = sf.GpuTree(ctx, x, m, rs)
        d2 = f.download()
        rf = sf.Resident.from_gtree(ctx, f)
        rf.run()
        out2 = rf.results()
        rf.close()
        f.close()
        for key in d:
            assert np.array_equal(d[key], d2[key]), key
        for a, b in zip(out, out2):
            assert np.array_equal(a, b)
        L = host.build_lists(x, m, rs, theta=0.3)
        assert np.array_equal(d["id_epj"], L.id_epj) and np.array_equal(d["id_spj"], L.id_spj)
    r.close()
    g.close()
    ctx.close()


@pytest.mark.parametrize("canonical", [True, False])
def test_device_tree_one_pass_rebuild(canonical):
    """A rebuild with as many groups as the last build writes its lists in
    one pass into regions sized from the last build's list lengths; lists
    that outgrow their regions (here: r_search 0.5, which keeps the tree
    and groups but puts most of the cluster on every EPJ list) are written
    again by an exact second pass.  Either way the lists and the forces equal a fresh
    build's."""
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    m, x, _ = host.cluster(20000, "plummer", kroupa=True, f_bin=0.3, seed=41)
    rs = np.full(m.size, 1.5 * R_OUT)
    g = sf.GpuTree(ctx, x, m, rs, canonical=canonical)
    assert g.walk_passes() == 2
    r = sf.Resident.from_gtree(ctx, g)
    m2, x2, _ = host.cluster(12000, "plummer", kroupa=True, f_bin=0.3, seed=42)
    rs4 = np.full(m.size, 50.0 * R_OUT)
    steps = [(x, m, rs, 1), (x, m, rs, 1), (x, m, rs4, 2), (x, m, rs4, 1), (0.5 * x, m, rs, None),
             (x2, m2, np.full(m2.size, 1.5 * R_OUT), None), (x, m, rs, None)]
    for xs, ms, rss, passes in steps:
        g.update(xs, ms, rss)
        if passes is not None:
            assert g.walk_passes() == passes
        d = g.download()
        r.refill(g)
        r.run()
        out = r.results()
        f = sf.GpuTree(ctx, xs, ms, rss, canonical=canonical)
        d2 = f.download()
        rf = sf.Resident.from_gtree(ctx, f)
        rf.run()
        out2 = rf.results()
        rf.close()
        f.close()
        for key in d:
            assert np.array_equal(d[key], d2[key]), key
        for a, b in zip(out, out2):
            assert np.array_equal(a, b)
    r.close()
    g.close()
    ctx.close()


def test_device_tree_rejects_host_pointers_as_device_input():
    """on_device inputs are checked on the host before any launch: arrays
    that are not device allocations (here host arrays) are an error."""
    ctx = sf.Context(0)
    m, x, _ = host.cluster(1000, "plummer", seed=1)
    g = sf.GpuTree(ctx, x, m)
    lib = sf.lib()
    xh = np.ascontiguousarray(x)
    mh = np.ascontiguousarray(m)
    assert lib.pth_gtree_update(g._h, 1000, xh.ctypes.data, mh.ctypes.data, None, 1) != 0
    assert "device allocations" in sf.lib().pth_last_error(ctx.handle).decode()
    with pytest.raises(RuntimeError):
        g.sizes()  # the failed update leaves no tree
    g.update(x, m)
    assert g.sizes()["n_group"] >= 1
    g.close()
    ctx.close()


def _expected_ngb(L, ocnt, ooff, ongb):
    """Oracle lists (tree order, EPJ positions) re-indexed by input index."""
    n = L.order.size
    rowp = np.repeat(np.arange(n), ocnt)
    o = np.argsort(L.order[rowp], kind="stable")
    cnt = np.zeros(n, np.int64)
    cnt[L.order] = ocnt
    off = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=off[1:])
    return cnt, off, L.order[ongb[o]]


@pytest.mark.parametrize("n,f_bin,theta,shift", [(30000, 0.5, 0.3, 0.0), (4000, 0.3, 0.0, 0.0),
                                                 (20000, 0.0, 0.8, 0.0), (30000, 0.5, 0.3, 8000.0)])
def test_device_neighbors_equal_oracle(n, f_bin, theta, shift):
    """pth_gtree_neighbors against the FP64 restatement on the host tree's
    lists: counts, offsets and entries equal (order included); shift moves
    the cluster to (shift, -300, 20) when nonzero."""
    m, x, _ = host.cluster(n, "plummer", kroupa=True, f_bin=f_bin, seed=41)
    if shift:
        x = x + np.array([shift, -300.0, 20.0])
    rs = 1.5 * R_OUT * np.maximum(np.cbrt(m / m.mean()), 1.0)
    L = host.build_lists(x, m, rs, theta=theta)
    ocnt, ooff, ongb = O.ngb_walks(L.n_epi, L.n_epj, L.epj_pos, L.epj_rs, L.id_epj, L.epj_pos, L.epj_rs)
    ctx = sf.Context(0)
    g = sf.GpuTree(ctx, x, m, rs, theta=theta)
    cnt, off, ngb, ms = g.neighbors()
    g.close()
    ctx.close()
    ecnt, eoff, engb = _expected_ngb(L, ocnt, ooff, ongb)
    print(f"[ngb N={n} theta={theta} shift={shift}] {ngb.size} entries, {(cnt > 1).sum()} particles with neighbours, {ms:.2f} ms",
          flush=True)
    assert np.array_equal(cnt, ecnt) and np.array_equal(off, eoff) and np.array_equal(ngb, engb)


def test_device_neighbors_one_pass_repeat():
    """A search with as many tiles as the last one writes its candidate lists
    in one pass into regions of the last search's candidate counts, and its
    neighbour lists into regions of the last search's per-particle counts; a
    larger r_search (same tree and tiles, more candidates and neighbours)
    outgrows them and takes the exact second passes.  The lists equal the
    FP64 restatement each time."""
    ctx = sf.Context(0)
    m, x, _ = host.cluster(20000, "plummer", kroupa=True, f_bin=0.5, seed=43)
    rs = np.full(m.size, 1.5 * R_OUT)
    g = sf.GpuTree(ctx, x, m, rs)
    # (tile walk passes, filter passes): a first search counts first; a
    # repeated one writes once; a 4x r_search outgrows both kinds of region
    for rss, rebuild, passes in ((rs, False, (2, 2)), (rs, False, (1, 1)), (4.0 * rs, True, (2, 2)),
                                 (4.0 * rs, False, (1, 1))):
        if rebuild:
            g.update(x, m, rss)
        cnt, off, ngb, _ = g.neighbors()
        assert (g.tile_walk_passes(), g.filter_passes()) == passes
        L = host.build_lists(x, m, rss, theta=0.3)
        ecnt, eoff, engb = _expected_ngb(L, *O.ngb_walks(L.n_epi, L.n_epj, L.epj_pos, L.epj_rs, L.id_epj,
                                                         L.epj_pos, L.epj_rs))
        assert np.array_equal(cnt, ecnt) and np.array_equal(off, eoff) and np.array_equal(ngb, engb)
    g.close()
    ctx.close()


def test_device_neighbors_edge_cases():
    """No r_search (no neighbours, not even self), one particle, and a
    repeated search after an update."""
    ctx = sf.Context(0)
    m, x, _ = host.cluster(3000, "plummer", seed=3)
    g = sf.GpuTree(ctx, x, m)
    cnt, off, ngb, _ = g.neighbors()
    assert cnt.sum() == 0 and off[-1] == 0 and ngb.size == 0
    g.close()
    g = sf.GpuTree(ctx, np.zeros((1, 3)), np.ones(1), np.full(1, 0.1))
    cnt, off, ngb, _ = g.neighbors()
    assert list(cnt) == [1] and list(off) == [0, 1] and list(ngb) == [0]
    m2, x2, _ = host.cluster(5000, "plummer", kroupa=True, f_bin=0.5, seed=4)
    rs2 = np.full(5000, 1.5 * R_OUT)
    g.update(x2, m2, rs2)
    cnt, off, ngb, _ = g.neighbors()
    L = host.build_lists(x2, m2, rs2, theta=0.3)
    ecnt, eoff, engb = _expected_ngb(L, *O.ngb_walks(L.n_epi, L.n_epj, L.epj_pos, L.epj_rs, L.id_epj, L.epj_pos,
                                                     L.epj_rs))
    assert np.array_equal(cnt, ecnt) and np.array_equal(ngb, engb)
    g.close()
    ctx.close()


def test_snapshot_feeds_the_device_step(tmp_path):
    """A PeTar snapshot (ASCII, as PeTar writes by default) read back and fed
    to the device tree and the changeover-mode evaluation with its own
    r_search, r_in, r_out and ids: bit-identical to the same step fed from
    the arrays the snapshot was written from; then the same clusters through
    the serving loop equal their resident evaluations."""
    from petar_amd import snapshot as S
    arrays, paths = [], []
    for k, n in enumerate((6000, 9000, 3000)):
        m, x, v = host.cluster(n, "plummer", kroupa=True, f_bin=0.3, seed=60 + k, with_vel=True)
        rout = R_OUT * np.maximum(np.cbrt(m / m.mean()), 1.0)
        s = S.Snapshot.new(n, file_id=k, time=0.125 * k)
        s["mass"], s["pos"], s["vel"] = m, x, v
        s["r_search"], s["r_in"], s["r_out"], s["id"] = 1.5 * rout, 0.1 * rout, rout, np.arange(1, n + 1) * 3
        p = tmp_path / f"data.{k}"
        s.write(p, "ascii")
        s.close()
        arrays.append((m, x, 1.5 * rout, 0.1 * rout, rout, np.arange(1, n + 1) * 3))
        paths.append(p)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, changeover=1)
    for p, (m, x, rs, rin, rout, ids) in zip(paths, arrays):
        inp = S.soft_inputs(S.Snapshot.read(p, "ascii"))
        g = sf.GpuTree(ctx, inp["pos"], inp["mass"], inp["r_search"])
        order = g.download()["order"]
        rd = sf.Resident.from_gtree(ctx, g)
        g.close()
        rd.changeover(inp["id"][order], inp["id"][order], inp["r_in"][order], inp["r_out"][order])
        rd.run()
        out_d = rd.results()
        rd.close()
        L = host.build_lists(x, m, rs, theta=0.3)
        rh = sf.Resident.from_lists(ctx, L, rin[L.order], rout[L.order])
        rh.changeover(ids[L.order], ids[L.order], rin[L.order], rout[L.order])
        rh.run()
        out_h = rh.results()
        rh.close()
        for a, b in zip(out_d, out_h):
            assert np.array_equal(a, b)
    ctx.close()
    # serving: the snapshot clusters as requests
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    srv = sf.Server(ctx, sf.SERVE_FUSED, hbm_budget=0, n_streams=2)
    lists = []
    for p in paths:
        inp = S.soft_inputs(S.Snapshot.read(p, "ascii"))
        lists.append(host.build_lists(inp["pos"], inp["mass"], inp["r_search"], theta=0.3))
    ids = [srv.add_lists(L) for L in lists]
    srv.replay(np.array(ids * 2), np.zeros(2 * len(ids)))
    for k, L in zip(ids, lists):
        r = sf.Resident.from_lists(ctx, L)
        r.run()
        ref = r.results()
        r.close()
        got = srv.results(k)
        assert all(np.array_equal(a, b) for a, b in zip(got, ref))
    srv.close()
    ctx.close()


def test_device_tree_rejects_bad_arguments():
    """Argument errors come back as error codes with a message, before any
    launch, and leave the context usable."""
    ctx = sf.Context(0)
    m, x, _ = host.cluster(500, "plummer", seed=2)
    with pytest.raises(sf.PthError, match="n outside"):
        sf.GpuTree(ctx, np.zeros((0, 3)), np.zeros(0))
    with pytest.raises(sf.PthError, match="limits must be"):
        sf.GpuTree(ctx, x, m, n_leaf_limit=0)
    with pytest.raises(sf.PthError, match="limits must be"):
        sf.GpuTree(ctx, x, m, n_group_limit=0)
    g = sf.GpuTree(ctx, x, m)
    lib = sf.lib()
    assert lib.pth_gtree_update(g._h, 500, None, None, None, 0) == sf.PTH_ERR_ARG
    assert lib.pth_gtree_update(g._h, -1, x.ctypes.data, m.ctypes.data, None, 0) == sf.PTH_ERR_ARG
    assert g.sizes()["n_group"] >= 1  # an argument error does not discard the tree
    g.close()
    ctx.close()


def test_device_tree_rejects_non_finite_input():
    ctx = sf.Context(0)
    m, x, _ = host.cluster(1000, "plummer", seed=4)
    xb = x.copy()
    xb[17, 1] = np.nan
    with pytest.raises(sf.PthError, match="non-finite"):
        sf.GpuTree(ctx, xb, m)
    rs = np.full(1000, 0.01)
    rs[3] = np.inf
    with pytest.raises(sf.PthError, match="non-finite"):
        sf.GpuTree(ctx, x, m, rs)
    g = sf.GpuTree(ctx, x, m, np.full(1000, 0.01))
    assert g.sizes()["n_group"] >= 1
    g.close()
    ctx.close()


def test_resident_from_a_tree_of_another_context():
    """The plan made during the build is taken over only under the same
    parameters: a resident evaluation in a context with other tiles and
    segments plans again and still equals the host-list evaluation."""
    m, x, _ = host.cluster(12000, "plummer", kroupa=True, f_bin=0.3, seed=71)
    rs = np.full(m.size, 1.5 * R_OUT)
    L = host.build_lists(x, m, rs, theta=0.3)
    kw = dict(eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    ca = sf.Context(0, min_seg_pairs=1, **kw)
    cb = sf.Context(0, segments=1, max_ipl=4, chunk_j=512, **kw)
    g = sf.GpuTree(ca, x, m, rs)
    outs = []
    for ctx, src in ((cb, "tree"), (cb, "lists"), (ca, "tree"), (ca, "tree")):
        r = sf.Resident.from_gtree(ctx, g) if src == "tree" else sf.Resident.from_lists(ctx, L)
        r.run()
        outs.append(r.results())
        r.close()
    g.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    for a, b in zip(outs[2], outs[3]):  # taken-over plan, then a plan made again
        assert np.array_equal(a, b)
    ca.close()
    cb.close()


def test_context_close_frees_its_objects_first():
    """Closing a context first closes the trees, resident evaluations and
    servers made on it (they use its device and stream), in whatever order
    the caller or the garbage collector gets to them."""
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    m, x, _ = host.cluster(3000, "plummer", seed=5)
    g = sf.GpuTree(ctx, x, m, np.full(m.size, 1.5 * R_OUT))
    r = sf.Resident.from_gtree(ctx, g)
    r.run()
    srv = sf.Server(ctx)
    ctx.close()
    assert not g._h and not r._h and not srv._h
    g.close()
    r.close()
    srv.close()
