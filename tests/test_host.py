"""Host side: initial conditions and the tree walk that produces the lists.

FDPS (the reference's tree) is not in the reference repository, so the
lists are pinned by invariants instead of by FDPS's output:
  - every particle of the system is accounted for exactly once in every
    group's interaction list (as an EPJ, or inside exactly one SPJ cell);
  - the SPJ moments equal brute-force moments of the cell's members;
  - theta = 0 opens everything (every j an EPJ);
  - every j within max(rs_i, rs_j) of an i of the group is an EPJ.
The MT19937 stream is pinned against numpy's legacy RandomState (the same
generator, independent code), and the Plummer generator against a
restatement of makePlummerModel's recipe driven by that numpy stream.
"""
import math

import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import host


def test_mt19937_matches_numpy_legacy():
    for seed in (0, 1, 12345):
        rs = np.random.RandomState(seed)
        m = host.MT19937(seed)
        ref = rs.randint(0, 2 ** 32, size=2000, dtype=np.uint64)
        got = np.array([m.int32() for _ in range(2000)], dtype=np.uint64)
        assert np.array_equal(ref, got)
        rs = np.random.RandomState(seed)
        m = host.MT19937(seed)
        assert np.array_equal(rs.random_sample(500), np.array([m.res53() for _ in range(500)]))


def _plummer_recipe(n, seed):
    """src/particle_distribution_generator.hpp:173-250 with numpy's MT19937."""
    rs = np.random.RandomState(seed)
    real2 = lambda: rs.randint(0, 2 ** 32, dtype=np.uint64) / 4294967296.0  # noqa: E731
    res53 = rs.random_sample
    PI = math.atan(1.0) * 4.0
    r_cutoff = 22.8 / (-3.0 * PI / (64.0 * -0.25))
    pos = np.zeros((n, 3))
    for i in range(n):
        r = 9999.9
        while r > r_cutoff:
            r = 1.0 / math.sqrt(math.pow(res53(), -2.0 / 3.0) - 1.0)
        phi = 2.0 * PI * res53()
        cth = 2.0 * (real2() - 0.5)
        sth = math.sqrt(1.0 - cth * cth)
        pos[i] = (r * sth * math.cos(phi), r * sth * math.sin(phi), r * cth)
        while True:
            v_try = res53()
            v_crit = 0.1 * res53()
            if v_crit < v_try * v_try * math.pow(1.0 - v_try * v_try, 3.5):
                res53(), res53()
                break
    m = np.full(n, 1.0 / n)
    cm = (m[:, None] * pos).sum(0) / m.sum()
    return (pos - cm) * (-3.0 * PI / (64.0 * -0.25))


def test_plummer_equal_follows_reference_recipe():
    m, x, v = host.plummer_equal(300, seed=0)
    ref = _plummer_recipe(300, 0)
    np.testing.assert_allclose(x, ref, rtol=1e-12, atol=1e-14)
    assert np.all(m == 1.0 / 300)


def test_plummer_statistics():
    m, x, v = host.plummer_equal(200000, seed=3)
    np.testing.assert_allclose((m[:, None] * x).sum(0), 0, atol=1e-12)
    r = np.sort(np.linalg.norm(x, axis=1))
    a = 3 * np.pi / 16
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1)
    assert r[100000] == pytest.approx(a / math.sqrt(2 ** (2 / 3) - 1), rel=0.02)
    # kinetic energy 1/4 in N-body units (virial, E = -1/4)
    t = 0.5 * (m * (v ** 2).sum(1)).sum()
    assert t == pytest.approx(0.25, rel=0.03)


def test_kroupa_imf():
    raw = host.kroupa_masses(400000, m_tot=0.0, seed=2)
    # analytic fraction of stars above 0.5 Msun for slopes 1.3 / 2.3 on [0.08, 150]
    a1, a2, lo, b, hi = 1.3, 2.3, 0.08, 0.5, 150.0
    w1 = (b ** (1 - a1) - lo ** (1 - a1)) / (1 - a1)
    w2 = b ** (a2 - a1) * (hi ** (1 - a2) - b ** (1 - a2)) / (1 - a2)
    n1 = (b ** (2 - a1) - lo ** (2 - a1)) / (2 - a1)
    n2 = b ** (a2 - a1) * (hi ** (2 - a2) - b ** (2 - a2)) / (2 - a2)
    mean = (n1 + n2) / (w1 + w2)
    frac_hi = w2 / (w1 + w2)
    assert (raw > 0.5).mean() == pytest.approx(frac_hi, abs=0.004)
    assert raw.mean() == pytest.approx(mean, rel=0.02)
    assert raw.min() >= 0.08 and raw.max() <= 150
    mm = host.kroupa_masses(1000, m_tot=2.0, seed=2)
    assert mm.sum() == pytest.approx(2.0)


def _potential_energy(x, m):
    _, pot = O.direct_newton(x, m)
    return 0.5 * (m * pot).sum()


@pytest.mark.parametrize("model", ["plummer", "king"])
def test_cluster_virial_units(model):
    m, x, v = host.cluster(4096, model, w0=7.0, kroupa=True, seed=4, with_vel=True)
    assert m.sum() == pytest.approx(1.0, rel=1e-12)
    np.testing.assert_allclose((m[:, None] * x).sum(0), 0, atol=1e-12)
    u = _potential_energy(x, m)
    assert u == pytest.approx(-0.5, rel=0.08)
    t = 0.5 * (m * (v ** 2).sum(1)).sum()
    assert t == pytest.approx(0.25, rel=0.08)


def _king_scipy(w0):
    """Independent King (1966) model: scipy ODE for W(r), M(r)."""
    from scipy.integrate import solve_ivp
    from scipy.special import erf

    def rho(w):
        w = max(w, 0.0)
        return math.exp(w) * erf(math.sqrt(w)) - math.sqrt(4 * w / math.pi) * (1 + 2 * w / 3)

    rho0 = rho(w0)

    def f(r, y):
        w, dw, m = y
        return [dw, -9 * rho(w) / rho0 - 2 * dw / r, 4 * math.pi * r * r * rho(w) / rho0]

    ev = lambda r, y: y[0]  # noqa: E731
    ev.terminal = True
    r0 = 1e-6
    sol = solve_ivp(f, [r0, 1e4], [w0 - 1.5 * r0 * r0, -3 * r0, 0.0], events=ev, rtol=1e-10, atol=1e-12,
                    dense_output=True)
    rt = sol.t_events[0][0]
    mt = sol.y_events[0][0][2]
    from scipy.optimize import brentq
    lag = [brentq(lambda r: sol.sol(r)[2] - q * mt, r0 * 10, rt) for q in (0.1, 0.5, 0.9)]
    return rt, lag


def test_king_w7_structure():
    rt, (r10, r50, r90) = _king_scipy(7.0)
    # literature: W0 = 7 has concentration c = log10(r_t / r_0) = 1.53
    assert math.log10(rt) == pytest.approx(1.53, abs=0.02)
    m, x, _ = host.cluster(200000, "king", w0=7.0, kroupa=False, seed=5)
    r = np.linalg.norm(x, axis=1)
    q10, q50, q90 = np.quantile(r, [0.1, 0.5, 0.9])
    assert q10 / q50 == pytest.approx(r10 / r50, rel=0.02)
    assert q90 / q50 == pytest.approx(r90 / r50, rel=0.02)
    assert r.max() / q50 <= rt / r50 * 1.001


def test_mass_segregation():
    m, x, _ = host.cluster(20000, "king", w0=7.0, kroupa=True, segregate=True, seed=6)
    r = np.linalg.norm(x, axis=1)
    # heaviest decile sits well inside the lightest decile
    o = np.argsort(m)
    assert np.median(r[o[-2000:]]) < 0.3 * np.median(r[o[:2000]])


def test_binaries_orbits():
    n = 2000
    m, x, v = host.cluster(n, "plummer", kroupa=True, f_bin=1.0, a_lo=1e-5, a_hi=1e-2, seed=7, with_vel=True)
    m1, m2 = m[0::2], m[1::2]
    dr = x[0::2] - x[1::2]
    dv = v[0::2] - v[1::2]
    r = np.linalg.norm(dr, axis=1)
    e_orb = 0.5 * (dv ** 2).sum(1) - (m1 + m2) / r
    a = -(m1 + m2) / (2 * e_orb)
    assert np.all(a > 1e-5 * 0.999) and np.all(a < 1e-2 * 1.001)
    la = np.log(a)
    assert la.mean() == pytest.approx(0.5 * (np.log(1e-5) + np.log(1e-2)), abs=0.15)
    hvec = np.cross(dr, dv)
    ecc = np.sqrt(np.maximum(0, 1 + 2 * e_orb * (hvec ** 2).sum(1) / (m1 + m2) ** 2))
    assert np.mean(ecc ** 2) == pytest.approx(0.5, abs=0.05)  # thermal: <e^2> = 1/2
    # centre of mass of each pair follows the cluster
    cm = (m1[:, None] * x[0::2] + m2[:, None] * x[1::2]) / (m1 + m2)[:, None]
    assert np.linalg.norm(cm, axis=1).max() < 40


# ---- tree walk invariants -------------------------------------------------------

@pytest.fixture(scope="module")
def small_lists():
    m, x, _ = host.cluster(6000, "plummer", kroupa=True, f_bin=0.2, seed=9)
    rng = np.random.default_rng(0)
    rs = 0.01 + 0.02 * rng.random(6000)
    L = host.build_lists(x, m, rs, theta=0.4, n_leaf_limit=16, n_group_limit=128)
    return L, m, x, rs


def test_lists_cover_every_particle_once(small_lists):
    L, m, x, rs = small_lists
    n = m.size
    eoff, soff = L.walk_offsets()
    for w in range(L.n_walk):
        cnt = np.zeros(n + 1, np.int64)
        ide = L.id_epj[eoff[w]:eoff[w + 1]]
        np.add.at(cnt, ide, 1)
        for s in L.id_spj[soff[w]:soff[w + 1]]:
            cnt[L.node_begin[s]] += 1
            cnt[L.node_end[s]] -= 1
        # ranges -> per-particle counts by prefix sum of the SPJ markers
        spj_cov = np.cumsum(np.bincount(
            np.concatenate([L.node_begin[L.id_spj[soff[w]:soff[w + 1]]], L.node_end[L.id_spj[soff[w]:soff[w + 1]]]]),
            weights=np.concatenate([np.ones(L.n_spj[w]), -np.ones(L.n_spj[w])]), minlength=n + 1))[:n]
        epj_cov = np.bincount(ide, minlength=n)
        assert np.all(epj_cov + spj_cov == 1), w
        mass = L.epj_mass[ide].sum() + L.spj_mass[L.id_spj[soff[w]:soff[w + 1]]].sum()
        assert mass == pytest.approx(m.sum(), rel=1e-12)


def test_group_particles_are_their_own_epj(small_lists):
    L, *_ = small_lists
    eoff, _ = L.walk_offsets()
    for w in range(L.n_walk):
        own = np.arange(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w])
        assert np.all(np.isin(own, L.id_epj[eoff[w]:eoff[w + 1]]))
        assert 0 < L.n_epi[w] <= 128


def test_spj_moments_brute_force(small_lists):
    L, *_ = small_lists
    rng = np.random.default_rng(1)
    for s in rng.choice(L.spj_mass.size, 60, replace=False):
        b, e = L.node_begin[s], L.node_end[s]
        M, c, q = O.moments(L.epj_pos[b:e], L.epj_mass[b:e])
        assert L.spj_mass[s] == pytest.approx(M, rel=1e-12)
        np.testing.assert_allclose(L.spj_pos[s], c, rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(L.spj_quad[s], q, rtol=1e-9, atol=1e-14 * max(M, 1e-300))


def test_symmetric_search_neighbours_are_epj(small_lists):
    L, m, x, rs = small_lists
    eoff, _ = L.walk_offsets()
    pos, r_s = L.epj_pos, L.epj_rs
    for w in range(0, L.n_walk, max(1, L.n_walk // 40)):
        i = np.arange(L.epi_begin[w], L.epi_begin[w] + L.n_epi[w])
        d = np.linalg.norm(pos[i][:, None, :] - pos[None, :, :], axis=2)
        crit = np.maximum(r_s[i][:, None], r_s[None, :])
        need = np.unique(np.nonzero(d < crit)[1])
        assert np.all(np.isin(need, L.id_epj[eoff[w]:eoff[w + 1]])), w


def test_theta_zero_opens_everything():
    m, x, _ = host.cluster(1500, "plummer", seed=3)
    L = host.build_lists(x, m, 0.01, theta=0.0, n_group_limit=64)
    assert L.n_spj.sum() == 0 and np.all(L.n_epj == 1500)


def test_tree_force_close_to_direct_sum():
    """theta = 0.3 quadrupole tree vs FP64 all-pairs sum (linear cutoff):
    the truncation error of the walk, about 1e-5."""
    m, x, _ = host.cluster(12000, "plummer", kroupa=True, seed=11)
    L = host.build_lists(x, m, 0.015, theta=0.3)
    a = O.lists_args(L)
    acc, pot, nnb = O.soft_force_walks(**a, eps2=0.0, rout2=1e-4)
    idx = np.concatenate([np.arange(b, b + n) for b, n in zip(L.epi_begin, L.n_epi)])
    s = np.arange(0, idx.size, 7)
    dacc, dpot, dnnb = O.direct_linear_cutoff(L.epj_pos, L.epj_mass, L.epj_rs, sample=idx[s], rout2=1e-4)
    e = np.linalg.norm(acc[s] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
    assert np.median(e) < 1e-4 and np.percentile(e, 99) < 1e-3
    assert np.array_equal(nnb[s], dnnb)
