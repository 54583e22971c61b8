"""The serving workload at its full size: 256 independent Plummer clusters, N
log-uniform in [1e3, 2e5] (Kroupa masses, 30% primordial binaries as in the
precision workload, theta 0.3, quadrupole SPJ, r_out 0.01, r_search 1.5
r_out; the stream bench.py serves), replayed with Poisson arrivals under an
HBM budget of every stream's scratch plus a quarter of the clusters'
resident bytes, so that clusters are evicted and re-admitted throughout.

Checked: every one of the 256 clusters gives bit-identical results alone
(one request at a time) and in the mix; sampled walks of 12 clusters
spread over the size range agree with the FP64 oracle on the same lists
(median 1e-6, 99th percentile 1e-5 for acc and pot; neighbour counts exact
outside the FP32 band of the search radius)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import host, soft_force as sf
from tests._common import nnb_check, rel_err, walk_centres

pytestmark = pytest.mark.gpu

R_OUT = 0.01
N_CLUSTERS = 256


@pytest.fixture(scope="module")
def stream():
    # the stream of bench.py's serving object (same seeds)
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(2e5), N_CLUSTERS)).astype(int)
    lists = []
    for k, n in enumerate(sizes):
        m, x, _ = host.cluster(int(n), "plummer", kroupa=True, f_bin=0.3, seed=4 * 1000 + k)
        lists.append(host.build_lists(x, m, 1.5 * R_OUT, theta=0.3))
    print(f"[stream] {N_CLUSTERS} clusters, N sum {int(sizes.sum())}", flush=True)
    return sizes, lists


def _sample(L, walks):
    eo, so = L.walk_offsets()
    ioff = np.concatenate([[0], np.cumsum(L.n_epi)])
    idx = np.concatenate([np.arange(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w]) for w in walks])
    rows = np.concatenate([np.arange(ioff[w], ioff[w + 1]) for w in walks])
    a = dict(n_epi=L.n_epi[walks], n_epj=L.n_epj[walks], n_spj=L.n_spj[walks], epi_pos=L.epj_pos[idx],
             epi_rs=L.epj_rs[idx], id_epj=np.concatenate([L.id_epj[eo[w]:eo[w + 1]] for w in walks]),
             id_spj=np.concatenate([L.id_spj[so[w]:so[w + 1]] for w in walks]), epj_pos=L.epj_pos,
             epj_mass=L.epj_mass, epj_rs=L.epj_rs, spj_pos=L.spj_pos, spj_mass=L.spj_mass, spj_quad=L.spj_quad)
    return a, rows


def test_serving_stream_256_clusters(stream):
    sizes, lists = stream
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    # footprints from an unconstrained server: the budget holds every
    # stream's scratch and a quarter of the clusters' resident bytes
    probe = sf.Server(ctx, sf.SERVE_FUSED, hbm_budget=0, n_streams=4)
    fp = np.array([probe.footprint(probe.add_lists(L)) for L in lists], dtype=np.int64)
    scratch = probe.memory()["scratch_max"]
    probe.close()
    budget = int(4 * scratch + max(0.25 * fp.sum(), 1.2 * fp.max()))
    srv = sf.Server(ctx, sf.SERVE_FUSED, hbm_budget=budget, n_streams=4)
    ids = [srv.add_lists(L) for L in lists]
    print(f"[stream] resident {fp.sum() / 2**30:.2f} GiB ({fp.sum() / sizes.sum():.0f} B per particle), scratch "
          f"{scratch / 2**30:.2f} GiB per stream, budget {budget / 2**30:.2f} GiB", flush=True)
    # alone: one request at a time
    alone = {}
    for k in ids:
        srv.wait(srv.submit(k))
        alone[k] = srv.results(k)
    print("[stream] alone pass done", flush=True)
    # the mix: every cluster twice, Poisson arrivals at about the saturated rate
    rng = np.random.default_rng(11)
    seq = rng.permutation(np.repeat(ids, 2))
    _, wall = srv.replay(rng.permutation(ids), np.zeros(len(ids)))
    rate = len(ids) / wall
    arr = np.cumsum(rng.exponential(1.0 / rate, seq.size))
    lat, wall = srv.replay(seq, arr - arr[0])
    st = srv.stats()
    print(f"[stream] mix of {seq.size} requests at {rate:.0f}/s: p50 {np.percentile(lat, 50):.3f} ms, p99 "
          f"{np.percentile(lat, 99):.3f} ms; admitted {st['admitted']}, evicted {st['evicted']}, pool peak "
          f"{st['pool_peak'] / 2**30:.2f} GiB", flush=True)
    assert st["evicted"] > 0 and st["pool_peak"] <= budget
    same = [k for k in ids if all(np.array_equal(a, b) for a, b in zip(srv.results(k), alone[k]))]
    assert len(same) == N_CLUSTERS, f"{N_CLUSTERS - len(same)} clusters differ alone vs in the mix"
    srv.close()
    # sampled walks of 12 clusters across the size range against the FP64 oracle
    order = np.argsort(sizes)
    chosen = order[np.linspace(0, N_CLUSTERS - 1, 12).astype(int)]
    for k in chosen:
        L = lists[k]
        walks = np.sort(rng.choice(L.n_walk, min(L.n_walk, 24), replace=False))
        a, rows = _sample(L, walks)
        racc, rpot, rnnb = O.soft_force_walks(**a, eps2=0.0, rout2=R_OUT ** 2)
        acc, pot, nnb = alone[ids[k]]
        ea, ep = rel_err(acc[rows], racc), rel_err(pot[rows], rpot)
        amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
        bad_exact, bad_band, n_amb = nnb_check(nnb[rows], rnnb, amb)
        print(f"[stream] cluster {k} (N={sizes[k]}): {rows.size} i, acc med {np.median(ea):.2e} p99 "
              f"{np.percentile(ea, 99):.2e} max {ea.max():.2e}; pot med {np.median(ep):.2e}; nnb outside band "
              f"{bad_exact}, beyond band {bad_band}, i with band pairs {n_amb}", flush=True)
        assert np.median(ea) <= 1e-6 and np.median(ep) <= 1e-6
        assert np.percentile(ea, 99) <= 1e-5 and np.percentile(ep, 99) <= 1e-5
        assert bad_exact == 0 and bad_band == 0
    ctx.close()
