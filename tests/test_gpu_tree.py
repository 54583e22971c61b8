"""The tree built on the device (pth_gtree_*) against the host tree
(petar_amd.host.build_lists, the FDPS-semantics stand-in the lists of every
other test come from): same particle order, same nodes, bit-identical
moments, same i-groups and, per group, the same EPJ and SPJ entries (the
order inside a list may differ); and a force evaluation fed from the device
lists equals the FP64 oracle on them."""
import time

import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import host, soft_force as sf
from tests._common import R_OUT, rel_err

pytestmark = pytest.mark.gpu


def _compare(n, seed, f_bin, theta, leaf=20, group=512, r_search=1.5 * R_OUT, model="plummer"):
    m, x, _ = host.cluster(n, model, kroupa=True, f_bin=f_bin, seed=seed)
    t0 = time.time()
    L = host.build_lists(x, m, r_search, theta=theta, n_leaf_limit=leaf, n_group_limit=group)
    t_host = time.time() - t0
    ctx = sf.Context(0)
    g = sf.GpuTree(ctx, x, m, np.full(n, r_search), theta=theta, n_leaf_limit=leaf, n_group_limit=group)
    d = g.download()
    tb, tw = g.timing()
    g.close()
    ctx.close()
    print(f"[gtree N={n} theta={theta}] host tree+walk {t_host * 1e3:.1f} ms, device build {tb:.2f} ms walk {tw:.2f} ms; "
          f"{len(d['n_epi'])} groups, {d['id_epj'].size} EPJ + {d['id_spj'].size} SPJ entries", flush=True)
    assert np.array_equal(d["order"], L.order)
    assert np.array_equal(d["node_begin"], L.node_begin) and np.array_equal(d["node_end"], L.node_end)
    assert np.array_equal(d["node_level"], L.node_level)
    # moments: the same FP64 sums in the same order
    assert np.array_equal(d["node_mass"], L.spj_mass)
    assert np.array_equal(d["node_com"], L.spj_pos) and np.array_equal(d["node_quad"], L.spj_quad)
    assert np.array_equal(d["epi_begin"], L.epi_begin) and np.array_equal(d["n_epi"], L.n_epi)
    assert np.array_equal(d["n_epj"], L.n_epj) and np.array_equal(d["n_spj"], L.n_spj)
    eo, so = L.walk_offsets()
    ge = np.concatenate([[0], np.cumsum(d["n_epj"])])
    gs = np.concatenate([[0], np.cumsum(d["n_spj"])])
    for w in range(L.n_walk):
        assert np.array_equal(np.sort(d["id_epj"][ge[w]:ge[w + 1]]), np.sort(L.id_epj[eo[w]:eo[w + 1]]))
        assert np.array_equal(np.sort(d["id_spj"][gs[w]:gs[w + 1]]), np.sort(L.id_spj[so[w]:so[w + 1]]))
    return L, d


@pytest.mark.parametrize("n,seed,f_bin,theta", [(16384, 1, 0.0, 0.3), (60000, 7, 0.3, 0.3), (5000, 3, 0.3, 0.0),
                                                (30000, 9, 0.5, 0.7)])
def test_device_tree_equals_host_tree(n, seed, f_bin, theta):
    _compare(n, seed, f_bin, theta)


def test_device_tree_limits():
    _compare(20000, 5, 0.3, 0.4, leaf=4, group=64)
    _compare(20000, 5, 0.3, 0.4, leaf=64, group=2048)


def test_device_lists_feed_the_force_path():
    """The device lists (their own order) through the resident evaluation
    against the FP64 oracle on the same lists."""
    L, d = _compare(16384, 2, 0.3, 0.3)
    ge = np.concatenate([[0], np.cumsum(d["n_epj"])])
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(d["epi_begin"], d["n_epi"])])
    args = dict(n_epi=d["n_epi"], n_epj=d["n_epj"], n_spj=d["n_spj"], epi_pos=L.epj_pos[idx], epi_rs=L.epj_rs[idx],
                id_epj=d["id_epj"], id_spj=d["id_spj"], epj_pos=L.epj_pos, epj_mass=L.epj_mass, epj_rs=L.epj_rs,
                spj_pos=d["node_com"], spj_mass=d["node_mass"], spj_quad=d["node_quad"])
    racc, rpot, rnnb = O.soft_force_walks(**args, eps2=0.0, rout2=R_OUT ** 2)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    r = sf.Resident(ctx, **args)
    r.run()
    acc, pot, nnb = r.results()
    r.close()
    ctx.close()
    ea, ep = rel_err(acc, racc), rel_err(pot, rpot)
    assert np.median(ea) <= 1e-6 and np.median(ep) <= 1e-6 and np.percentile(ea, 99) <= 1e-4
    assert np.mean(nnb != rnnb) <= 1e-3
    assert ge[-1] == d["id_epj"].size
