"""C ABI without a GPU: the libraries load, export every function their
headers declare, the host-only planner covers every pair exactly once, and
the entry points fail cleanly where no device exists."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from petar_amd import host, soft_force as sf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(pth_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


@pytest.mark.parametrize("header,loader", [("petar_hip.h", sf.lib), ("petar_host.h", host.lib)])
def test_every_declared_symbol_is_exported(header, loader):
    L = loader()
    names = _declared(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(_declared("petar_hip.h")) == set(sf.EXPORTED)


def test_params_default():
    p = sf.Params()
    sf.lib().pth_params_default(C.byref(p))
    assert (p.eps2, p.rcut2, p.G, p.nnb_mode, p.exec_mode, p.chunk_j, p.max_walk, p.max_ipl, p.j_vgpr,
            p.changeover, p.ngb_factor, p.segments, p.streams) == (0.0, 0.0, 1.0, 0, 0, 0, 1000, 2, 0, 0, 0.99, 8, 3)
    assert (p.df_near, p.spj_monopole, p.graph, p.min_seg_pairs, p.persistent) == (1, 0, 0, 0.0, 0)


def test_create_without_device_fails_cleanly():
    h = C.c_void_p()
    p = sf.default_params(rcut2=1e-4)
    rc = sf.lib().pth_create(C.byref(h), 0, C.byref(p))
    if rc == sf.PTH_OK:  # a GPU is visible here
        sf.lib().pth_destroy(h)
        pytest.skip("a device is present")
    assert rc == sf.PTH_ERR_HIP
    assert b"no HIP device" in sf.lib().pth_create_error()


def test_create_rejects_bad_params():
    h = C.c_void_p()
    for kw in (dict(nnb_mode=3), dict(exec_mode=7), dict(rcut2=-1.0), dict(changeover=2),
               dict(changeover=1, exec_mode=sf.EXEC_PLAIN), dict(ngb_factor=-1.0)):
        p = sf.default_params(**kw)
        assert sf.lib().pth_create(C.byref(h), 0, C.byref(p)) == sf.PTH_ERR_ARG
    assert sf.lib().pth_create(None, 0, None) == sf.PTH_ERR_ARG


def test_null_context_calls():
    L = sf.lib()
    assert L.pth_send_j(None, None, 0, None, None, 0, None) == sf.PTH_ERR_ARG
    assert L.pth_dispatch(None, 0, 0, None, None, None, None, None, None, None) == sf.PTH_ERR_ARG
    assert L.pth_retrieve(None, 0, 0, None, None, None) == sf.PTH_ERR_ARG
    assert L.pth_resident_run(None) == sf.PTH_ERR_ARG
    assert L.pth_device(None) == -1
    assert L.pth_last_error(None) == b"null context"
    # the device tree's entry points check their handles before any HIP call
    assert L.pth_gtree_build(None, None, 10, None, None, None, 20, 512, 0.3, 1) == sf.PTH_ERR_ARG
    assert L.pth_gtree_update(None, 10, None, None, None, 0) == sf.PTH_ERR_ARG
    assert L.pth_gtree_sizes(None, None, None, None, None) == sf.PTH_ERR_ARG
    assert L.pth_gtree_download(None, *([None] * 13)) == sf.PTH_ERR_ARG
    assert L.pth_gtree_timing(None, None, None, None) == sf.PTH_ERR_ARG
    assert L.pth_gtree_stats(None, None, None, None) == sf.PTH_ERR_ARG
    assert L.pth_gtree_neighbors(None, None) == sf.PTH_ERR_ARG
    assert L.pth_gtree_neighbors_download(None, None, None, None, None) == sf.PTH_ERR_ARG
    assert L.pth_resident_from_gtree(None, None, None) == sf.PTH_ERR_ARG
    assert L.pth_resident_refill(None, None) == sf.PTH_ERR_ARG
    L.pth_gtree_free(None)


def _coverage(ne, nj, ns, mode, cj, max_ipl=4):
    items, es, ss, rit = sf.plan_debug(ne, nj, ns, mode, cj, max_ipl)
    ioff = np.concatenate([[0], np.cumsum(ne)])
    cov_e, cov_s, slots, tiles = {}, {}, [], {}
    for it in items:
        w, ib, nipl, eb0, eb1, s0, s1, part = it
        if w < 0:
            continue
        ni, ipl = nipl & 0xFFFF, nipl >> 16
        assert ipl in (1, 2, 3, 4) and ipl <= max_ipl
        assert 0 < ni <= 64 * ipl and ioff[w] <= ib and ib + ni <= ioff[w + 1]
        assert es[w] // 8 <= eb0 <= eb1 <= es[w + 1] // 8 and ss[w] <= s0 <= s1 <= ss[w + 1]
        assert tiles.setdefault((w, ib), (ni, ipl)) == (ni, ipl)
        cov_e.setdefault((w, ib), []).append((eb0, eb1))
        cov_s.setdefault((w, ib), []).append((s0, s1))
        if part >= 0:
            slots.extend(range(part, part + ipl))
    assert len(slots) == len(set(slots))
    for w in range(len(ne)):
        assert es[w + 1] - es[w] == ((nj[w] + 7) // 8 * 8 + 8 if nj[w] else 0) and ss[w + 1] - ss[w] == ns[w]
        # the walk's tiles cover its i exactly once
        mine = sorted((ib, ni) for (ww, ib), (ni, _) in tiles.items() if ww == w)
        pos = ioff[w]
        for ib, ni in mine:
            assert ib == pos
            pos += ni
        assert pos == ioff[w + 1]
        for ib, _ in mine:
            for cov, lo, hi in ((cov_e, es[w] // 8, es[w + 1] // 8), (cov_s, ss[w], ss[w + 1])):
                r = sorted(cov[(w, ib)])
                assert r[0][0] == lo and r[-1][1] == hi
                assert all(r[k][1] == r[k + 1][0] for k in range(len(r) - 1))
    # reduction items: one per split tile
    for out, ni, pf, npart, ipl, *_ in rit:
        assert npart > 1 and 0 < ni <= 64 * ipl
        assert tiles[(np.searchsorted(ioff, out, side="right") - 1, out)] == (ni, ipl)
    return items


def test_plan_covers_every_pair_once():
    rng = np.random.default_rng(0)
    for trial in range(150):
        nw = int(rng.integers(0, 30))
        ne = rng.integers(0, 300, nw)
        ne[rng.random(nw) < 0.1] = 0
        nj = rng.integers(0, 5000, nw)
        ns = rng.integers(0, 1500, nw)
        cj = int(rng.choice([0, 64, 500, 3000]))
        items = _coverage(ne, nj, ns, sf.EXEC_FUSED, cj, int(rng.choice([1, 2, 4])))
        # launch order: 8 XCD runs interleaved in groups of 4 items
        assert items.shape[0] % 32 == 0 or items.shape[0] == 0


def test_plan_is_independent_of_batch_mates():
    """A walk's items (tiles and j-chunks) depend only on its own sizes, so
    its results are the same alone and inside any batch."""
    rng = np.random.default_rng(3)
    ne, nj, ns = rng.integers(1, 400, 12), rng.integers(0, 20000, 12), rng.integers(0, 9000, 12)

    def walk_items(items, es, ss, w, ioff):
        out = []
        for it in items:
            if it[0] == w:
                out.append((it[1] - ioff[w], it[2], it[3] - es[w] // 8, it[4] - es[w] // 8, it[5] - ss[w],
                            it[6] - ss[w]))
        return sorted(out)

    items, es, ss, _ = sf.plan_debug(ne, nj, ns, sf.EXEC_FUSED, 1024)
    ioff = np.concatenate([[0], np.cumsum(ne)])
    for w in (0, 5, 11):
        alone, es1, ss1, _ = sf.plan_debug(ne[w:w + 1], nj[w:w + 1], ns[w:w + 1], sf.EXEC_FUSED, 1024)
        a = walk_items(alone, es1, ss1, 0, [0, ne[w]])
        b = walk_items(items, es, ss, w, ioff)
        assert a == b


def test_plan_xcd_runs_balanced():
    rng = np.random.default_rng(5)
    ne = rng.integers(32, 512, 300)
    nj = (10 ** rng.uniform(3, 5, 300)).astype(int)
    ns = (10 ** rng.uniform(2, 4, 300)).astype(int)
    items, *_ = sf.plan_debug(ne, nj, ns, sf.EXEC_FUSED, 2048)
    cost = np.zeros(8)
    for k, it in enumerate(items):
        if it[0] >= 0:
            x = (k // 4) % 8
            ipl = it[2] >> 16
            ce, cs = 9.5 + 2.25 / ipl, 18.5 + 5.5 / ipl
            cost[x] += 64 * ipl * (ce * 8 * (it[4] - it[3]) + cs * (it[6] - it[5]) + 50)
    assert cost.max() / cost.mean() < 1.05


def test_plan_plain_has_no_items():
    items, es, ss, rit = sf.plan_debug([10, 70], [5, 9], [1, 0], sf.EXEC_PLAIN, 0)
    assert items.shape[0] == 0 and rit.shape[0] == 0 and list(es) == [0, 16, 40]


def test_plan_rejects_bad_input():
    L = sf.lib()
    assert L.pth_plan_debug(-1, None, None, None, 0, 0, 4, 1, None, 0, None, None, None, 0, None, None) == sf.PTH_ERR_ARG
    assert L.pth_plan_debug(2, None, None, None, 0, 0, 4, 1, None, 0, None, None, None, 0, None, None) == sf.PTH_ERR_ARG


def test_fdps_record_layouts():
    """The numpy record types match the C++ structs of src/soft_ptcl.hpp."""
    assert sf.EPI_SOFT.itemsize == 48 and sf.EPI_SOFT.fields["pos"][1] == 8
    assert sf.EPJ_SOFT.itemsize == 120 and sf.EPJ_SOFT.fields["pos"][1] == 16
    assert sf.EPJ_SOFT.fields["r_search"][1] == 80
    assert sf.FORCE_SOFT.itemsize == 40 and sf.FORCE_SOFT.fields["n_ngb"][1] == 32


def test_plan_segments_keep_every_walks_items():
    """Splitting a batch into segments changes only the launch arrangement:
    every walk keeps exactly its items (so results are bit-identical), each
    segment is a contiguous walk range holding only its own items and its own
    reduction items, and the segments cover all walks."""
    rng = np.random.default_rng(11)
    ne = rng.integers(1, 400, 200)
    nj = (10 ** rng.uniform(2, 4.5, 200)).astype(int)
    ns = (10 ** rng.uniform(1, 4, 200)).astype(int)

    def per_walk(items):
        d = {}
        for it in items:
            if it[0] >= 0:
                d.setdefault(int(it[0]), []).append(tuple(int(x) for x in it[1:]))
        return {k: sorted(v) for k, v in d.items()}

    base, es0, ss0, rit0 = sf.plan_debug(ne, nj, ns, sf.EXEC_FUSED, 1024, 2)
    for n_seg in (2, 5, 8, 32):
        items, es, ss, rit, sg = sf.plan_debug(ne, nj, ns, sf.EXEC_FUSED, 1024, 2, n_seg=n_seg, segments=True)
        assert per_walk(items) == per_walk(base)
        assert np.array_equal(es, es0) and np.array_equal(ss, ss0)
        assert sorted(map(tuple, rit)) == sorted(map(tuple, rit0))
        assert len(sg) == n_seg and sg[0, 0] == 0 and sg[-1, 1] == len(ne)
        assert all(sg[k, 1] == sg[k + 1, 0] for k in range(len(sg) - 1))
        ioff = np.concatenate([[0], np.cumsum(ne)])
        for g in sg:
            wb, we = g[0], g[1]
            for c in range(4):
                for it in items[g[2 + c]:g[2 + c] + g[6 + c]]:
                    assert it[0] < 0 or wb <= it[0] < we
            for r in rit[g[10]:g[10] + g[11]]:
                assert ioff[wb] <= r[0] < ioff[we]
