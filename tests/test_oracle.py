"""Pins the FP64 oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the oracle (dropped
term, wrong sign or factor, swapped operand, wrong neighbour boundary)
fails it:
  - the reference x86 kernel itself (golden fixture of the simd_test.cxx
    scenario, and live on tree lists when oracle/_ref is built);
  - closed forms of the two-body force with linear cutoff and of the self pair;
  - Newton's third law of the all-pairs sum;
  - the multipole error of the quadrupole SPJ falling as (s/R)^3;
  - the monopole limit;
  - the changeover identities.
"""
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from tests._common import R_OUT, c1_lists, rel_err, walk_centres

GOLD = os.path.join(os.path.dirname(__file__), "golden", "simd_test_scenario.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _gold_walk(d):
    n_epi = int(d["n_epi"])
    n_epj = d["epj_mass"].size
    n_spj = d["spj_mass"].size
    return dict(n_epi=[n_epi], n_epj=[n_epj], n_spj=[n_spj], epi_pos=d["epj_pos"][:n_epi],
                epi_rs=d["epj_rs"][:n_epi], id_epj=np.arange(n_epj), id_spj=np.arange(n_spj),
                epj_pos=d["epj_pos"], epj_mass=d["epj_mass"], epj_rs=d["epj_rs"], spj_pos=d["spj_pos"],
                spj_mass=d["spj_mass"], spj_quad=d["spj_quad"])


# ---- golden fixture: the simd_test.cxx scenario --------------------------------

def test_golden_oracle_reproduces_stored_oracle(gold):
    """The oracle still computes what it computed when the fixture was made."""
    w = _gold_walk(gold)
    for mode in (0, 1, 2):
        a, p, nb = O.soft_force_walks(**w, eps2=float(gold["eps2"]), rout2=float(gold["rout2"]), nnb_mode=mode,
                                      part="epj")
        np.testing.assert_allclose(a, gold[f"oracle_epj_mode{mode}_acc"], rtol=1e-12, atol=0)
        np.testing.assert_allclose(p, gold[f"oracle_epj_mode{mode}_pot"], rtol=1e-12)
        assert np.array_equal(nb, gold[f"oracle_epj_mode{mode}_nnb"])
    a, p, _ = O.soft_force_walks(**w, eps2=float(gold["eps2"]), rout2=float(gold["rout2"]), part="spj")
    np.testing.assert_allclose(a, gold["oracle_spj_acc"], rtol=1e-12)


@pytest.mark.parametrize("isa", ["avx512", "avx2"])
def test_golden_oracle_vs_reference_kernel(gold, isa):
    """Oracle vs the reference phantomquad kernel on the reference's own test
    scenario.  Tolerances follow from the kernel's rsqrt: one Newton step on
    EPJ (src/petar.hpp:13), the raw estimate on SPJ (rsqrt14, 2^-14, on
    AVX-512; rsqrt, 1.5*2^-12, on AVX2) times the r^-7..r^-1 powers, widened
    by cancellation between monopole and quadrupole terms; simd_test.cxx
    prints differences above DF_MAX = 7e-3."""
    ea = rel_err(gold[f"ref_{isa}_epj_acc"], gold["oracle_epj_mode1_acc"])
    ep = rel_err(gold[f"ref_{isa}_epj_pot"], gold["oracle_epj_mode1_pot"])
    assert np.median(ea) < 2e-6 and ea.max() < 2e-5
    assert np.median(ep) < 2e-6 and ep.max() < 2e-5
    # the x86 kernel counts r2 + eps2 <= rs^2 (src/phantomquad_for_p3t_x86.hpp:516-525)
    assert np.array_equal(gold[f"ref_{isa}_epj_nnb"], gold["oracle_epj_mode1_nnb"])
    sa = rel_err(gold[f"ref_{isa}_spj_acc"], gold["oracle_spj_acc"])
    sp = rel_err(gold[f"ref_{isa}_spj_pot"], gold["oracle_spj_pot"])
    lim = {"avx512": (2e-4, 1e-3, 5e-3), "avx2": (5e-4, 5e-3, 1e-2)}[isa]
    assert np.median(sa) < lim[0] and sa.max() < lim[1]
    assert np.median(sp) < lim[1] and sp.max() < lim[2]


def test_golden_wrong_quadrupole_sign_is_caught(gold):
    """A planted mistake in the SPJ restatement (quadrupole with flipped sign)
    must fail the comparison with the reference kernel."""
    w = _gold_walk(gold)
    w["spj_quad"] = -gold["spj_quad"]
    a, _, _ = O.soft_force_walks(**w, eps2=float(gold["eps2"]), rout2=float(gold["rout2"]), part="spj")
    assert np.median(rel_err(gold["ref_avx512_spj_acc"], a)) > 1e-2


# ---- live comparison with the reference kernel on tree lists -------------------

@pytest.mark.skipif(not O.ref_available(), reason="reference kernel not built (needs /root/reference)")
@pytest.mark.parametrize("eps", [0.0, 0.1, 0.5, 2.0])
def test_oracle_vs_reference_kernel_tree_lists(eps):
    """eps is the Plummer softening length in units of r_out: eps2 enters
    every EPJ and SPJ pair as r^2 + eps^2 (src/soft_force.hpp:44-60,132-140)."""
    eps2 = (eps * R_OUT) ** 2
    L = c1_lists(8192, seed=4)
    a = O.lists_args(L)
    acc, pot, nnb = O.soft_force_walks(**a, eps2=eps2, rout2=R_OUT ** 2, nnb_mode=1)
    racc, rpot, rnnb = O.ref_soft_force_walks(**a, eps2=eps2, rout2=R_OUT ** 2)
    avx512 = O.cpu_has_avx512()
    ea, ep = rel_err(racc, acc), rel_err(rpot, pot)
    assert np.median(ea) < (5e-5 if avx512 else 2e-4) and ea.max() < (1e-3 if avx512 else 5e-3)
    assert np.median(ep) < (5e-5 if avx512 else 2e-4) and ep.max() < (1e-3 if avx512 else 5e-3)
    amb = O.nnb_ambiguous(**a, centres=None, eps2=eps2)
    d = np.abs(rnnb - nnb)
    assert np.all(d[amb == 0] == 0) and np.all(d <= amb)


# ---- monopole superparticles (builds without USE_QUAD) ----------------------------

def test_monopole_spj_is_the_quadrupole_pair_without_moments():
    """The monopole pair (src/soft_force.hpp:125-158) equals the quadrupole
    pair of src/soft_force.hpp:160-200 with a zero second moment, term for
    term, on tree lists with softening."""
    L = c1_lists(4096, seed=6)
    a = O.lists_args(L)
    mono = O.soft_force_walks(**a, eps2=1e-6, rout2=R_OUT ** 2, part="spj", mono=True)
    a0 = dict(a, spj_quad=np.zeros_like(L.spj_quad))
    quad0 = O.soft_force_walks(**a0, eps2=1e-6, rout2=R_OUT ** 2, part="spj")
    assert rel_err(mono[0], quad0[0]).max() < 1e-13 and rel_err(mono[1], quad0[1]).max() < 1e-13
    # and differs from the quadrupole evaluation by the quadrupole terms
    quad = O.soft_force_walks(**a, eps2=1e-6, rout2=R_OUT ** 2, part="spj")
    assert np.median(rel_err(mono[0], quad[0])) > 1e-6


def test_monopole_spj_two_body():
    """One superparticle of mass m at separation d: acc = G m d / (d^2 +
    eps^2)^(3/2) towards it, pot = -G m / sqrt(d^2 + eps^2), no cutoff."""
    xi, xs, m, G, eps2 = np.array([0.1, -0.2, 0.3]), np.array([0.1003, -0.2004, 0.3]), 2.5e-3, 1.7, 1e-8
    w = dict(n_epi=[1], n_epj=[0], n_spj=[1], epi_pos=np.array([xi]), epi_rs=[0.0], id_epj=[], id_spj=[0],
             epj_pos=[], epj_mass=[], epj_rs=[], spj_pos=np.array([xs]), spj_mass=[m], spj_quad=np.ones((1, 6)))
    a, p, nb = O.soft_force_walks(**w, eps2=eps2, rout2=R_OUT ** 2, G=G, mono=True)
    d = xs - xi
    r2 = d @ d + eps2
    np.testing.assert_allclose(a[0], G * m * d / r2 ** 1.5, rtol=1e-14)
    assert p[0] == pytest.approx(-G * m / np.sqrt(r2), rel=1e-14) and nb[0] == 0


@pytest.mark.skipif(not O.ref_available(), reason="reference kernel not built (needs /root/reference)")
@pytest.mark.parametrize("eps", [0.0, 0.5])
def test_monopole_oracle_vs_reference_kernel(eps):
    """The monopole restatement against the reference x86 kernel driven as
    CalcForceEpSpMonoSimd drives it (src/soft_force.hpp:427-480: run_epj on
    the superparticles' mass and centre, one Newton step on its rsqrt)."""
    eps2 = (eps * R_OUT) ** 2
    L = c1_lists(8192, seed=8)
    a = O.lists_args(L)
    acc, pot, _ = O.soft_force_walks(**a, eps2=eps2, rout2=R_OUT ** 2, part="spj", mono=True)
    racc, rpot, _ = O.ref_soft_force_walks(**a, eps2=eps2, rout2=R_OUT ** 2, part="spj", mono=True)
    ea, ep = rel_err(racc, acc), rel_err(rpot, pot)
    assert np.median(ea) < 2e-6 and ea.max() < 5e-5
    assert np.median(ep) < 2e-6 and ep.max() < 5e-5
    # a planted sign error (the repulsive form) is caught
    w = dict(a, spj_mass=-L.spj_mass)
    bad, _, _ = O.soft_force_walks(**w, eps2=eps2, rout2=R_OUT ** 2, part="spj", mono=True)
    assert np.median(rel_err(racc, bad)) > 1.0


# ---- closed forms ----------------------------------------------------------------

def _pair(xi, xj, mj, rsi=0.0, rsj=0.0, eps2=0.0, rout2=R_OUT ** 2, G=1.0, mode=0):
    w = dict(n_epi=[1], n_epj=[1], n_spj=[0], epi_pos=np.array([xi]), epi_rs=[rsi], id_epj=[0], id_spj=[],
             epj_pos=np.array([xj]), epj_mass=[mj], epj_rs=[rsj], spj_pos=[], spj_mass=[], spj_quad=[])
    a, p, nb = O.soft_force_walks(**w, eps2=eps2, rout2=rout2, G=G, nnb_mode=mode)
    return a[0], p[0], nb[0]


def test_two_body_outside_cutoff():
    xi, xj, m, G = np.array([0.1, -0.2, 0.3]), np.array([0.4, 0.2, -0.1]), 2.5e-3, 1.7
    d = xj - xi
    r = np.linalg.norm(d)
    a, p, _ = _pair(xi, xj, m, G=G)
    np.testing.assert_allclose(a, G * m * d / r ** 3, rtol=1e-14)
    assert p == pytest.approx(-G * m / r, rel=1e-14)


def test_two_body_inside_cutoff_is_linear():
    xi, xj, m = np.array([0.0, 0.0, 0.0]), np.array([0.003, -0.004, 0.0]), 1e-3
    a, p, _ = _pair(xi, xj, m)
    np.testing.assert_allclose(a, m * (xj - xi) / R_OUT ** 3, rtol=1e-14)
    assert p == pytest.approx(-m / R_OUT, rel=1e-14)


def test_self_pair():
    x = np.array([0.3, 0.1, -0.2])
    a, p, nb = _pair(x, x, 4e-3, rsi=0.01, rsj=0.01)
    assert np.all(a == 0) and p == pytest.approx(-4e-3 / R_OUT, rel=1e-15) and nb == 1


def test_neighbour_boundary_conventions():
    """Mode 0 (<), mode 1 (<=) and mode 2 (no eps) at r exactly equal to rs."""
    xi, xj = np.zeros(3), np.array([0.5, 0.0, 0.0])
    assert _pair(xi, xj, 1.0, rsi=0.5, mode=0)[2] == 0
    assert _pair(xi, xj, 1.0, rsi=0.5, mode=1)[2] == 1
    # eps enters modes 0/1 but not mode 2: r^2 + eps^2 > rs^2 > r^2
    assert _pair(xi, xj, 1.0, rsi=0.5000001, eps2=1e-6, mode=0)[2] == 0
    assert _pair(xi, xj, 1.0, rsi=0.5000001, eps2=1e-6, mode=2)[2] == 1
    # the larger of the two search radii counts
    assert _pair(xi, xj, 1.0, rsi=0.1, rsj=0.6, mode=0)[2] == 1


def test_newton_third_law_all_pairs():
    rng = np.random.default_rng(1)
    n = 400
    pos = rng.normal(size=(n, 3)) * 0.02
    m = rng.random(n) / n
    a, p, _ = O.direct_linear_cutoff(pos, m, 0.0, eps2=1e-8, rout2=R_OUT ** 2)
    tot = (m[:, None] * a).sum(0)
    assert np.linalg.norm(tot) < 1e-13 * (m[:, None] * np.abs(a)).sum()
    # potential energy counts every pair twice, the self term once per particle
    e_pair = 0.5 * (m * p).sum()
    assert e_pair < 0


def test_theta0_lists_equal_direct_sum():
    L = c1_lists(2048, theta=0.0, seed=7)
    assert L.n_spj.sum() == 0 and np.all(L.n_epj == 2048)
    a = O.lists_args(L)
    acc, pot, nnb = O.soft_force_walks(**a, eps2=0.0, rout2=R_OUT ** 2)
    idx = np.concatenate([np.arange(b, b + n) for b, n in zip(L.epi_begin, L.n_epi)])
    dacc, dpot, dnnb = O.direct_linear_cutoff(L.epj_pos, L.epj_mass, L.epj_rs, sample=idx, rout2=R_OUT ** 2)
    np.testing.assert_allclose(acc, dacc, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(pot, dpot, rtol=1e-12)
    assert np.array_equal(nnb, dnnb)


# ---- quadrupole superparticles ---------------------------------------------------

def _spj_vs_members(R, seed=3):
    rng = np.random.default_rng(seed)
    k = 200
    s = 0.05
    mem = rng.normal(size=(k, 3)) * s * np.array([1.0, 0.6, 0.3])
    mm = rng.random(k) / k
    M, com, q = O.moments(mem, mm)
    dirs = rng.normal(size=(64, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    xi = com + R * dirs
    n = xi.shape[0]
    w = dict(n_epi=[n], n_epj=[k], n_spj=[1], epi_pos=xi, epi_rs=np.zeros(n), id_epj=np.arange(k), id_spj=[0],
             epj_pos=mem, epj_mass=mm, epj_rs=np.zeros(k), spj_pos=com[None], spj_mass=[M], spj_quad=q[None])
    a_s, p_s, _ = O.soft_force_walks(**w, eps2=0.0, rout2=0.0, part="spj")
    a_d, p_d, _ = O.soft_force_walks(**w, eps2=0.0, rout2=0.0, part="epj")
    return np.median(rel_err(a_s, a_d)), np.median(rel_err(p_s, p_d))


def test_quadrupole_error_falls_as_octupole():
    """With the quadrupole included the truncation error is the octupole,
    ~(s/R)^3 relative; a wrong quadrupole term would leave (s/R)^2."""
    ea1, ep1 = _spj_vs_members(1.0)
    ea2, ep2 = _spj_vs_members(2.0)
    assert ea1 / ea2 > 6.0 and ep1 / ep2 > 6.0
    assert ea1 < 2e-3


def test_monopole_limit():
    rng = np.random.default_rng(2)
    xi = rng.normal(size=(10, 3))
    xs = np.array([[3.0, -2.0, 1.0]])
    w = dict(n_epi=[10], n_epj=[0], n_spj=[1], epi_pos=xi, epi_rs=np.zeros(10), id_epj=[], id_spj=[0],
             epj_pos=[], epj_mass=[], epj_rs=[], spj_pos=xs, spj_mass=[0.7], spj_quad=np.zeros((1, 6)))
    a, p, _ = O.soft_force_walks(**w, eps2=0.0, rout2=0.0, G=1.3)
    d = xs - xi
    r = np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(a, 1.3 * 0.7 * d / r[:, None] ** 3, rtol=1e-13)
    np.testing.assert_allclose(p, -1.3 * 0.7 / r, rtol=1e-13)


def test_moments_parallel_axis():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(50, 3))
    m = rng.random(50)
    M, c, q = O.moments(x, m)
    assert M == pytest.approx(m.sum())
    np.testing.assert_allclose(c, (m[:, None] * x).sum(0) / m.sum(), rtol=1e-13)
    # second moment about the origin = about the CoM + M c c^T
    q0 = np.array([(m * x[:, 0] ** 2).sum(), (m * x[:, 1] ** 2).sum(), (m * x[:, 2] ** 2).sum(),
                   (m * x[:, 0] * x[:, 1]).sum(), (m * x[:, 0] * x[:, 2]).sum(), (m * x[:, 1] * x[:, 2]).sum()])
    cc = np.array([c[0] ** 2, c[1] ** 2, c[2] ** 2, c[0] * c[1], c[0] * c[2], c[1] * c[2]])
    np.testing.assert_allclose(q + M * cc, q0, rtol=1e-12)
    _, _, q1 = O.moments(x[:1], m[:1])
    assert np.all(q1 == 0)


def test_nnb_ambiguity_band_brute_force():
    xi = np.zeros((1, 3))
    rs = 0.01
    xj = np.array([[rs * (1 + 1e-9), 0, 0], [rs * (1 - 1e-3), 0, 0], [rs * 2, 0, 0]])
    w = dict(n_epi=[1], n_epj=[3], epi_pos=xi, epi_rs=[rs], id_epj=[0, 1, 2], epj_pos=xj, epj_rs=np.zeros(3))
    amb = O.nnb_ambiguous(**w, centres=None, eps2=0.0, rel=1e-6)
    assert amb[0] == 1


# ---- changeover (src/changeover.hpp) ---------------------------------------------

def _w():
    L = O.orc()
    return L.orc_changeover_wpot, L.orc_changeover_wacc, L.orc_changeover_wacc1


@pytest.mark.parametrize("rin,rout", [(0.001, 0.01), (0.3, 1.0), (0.09, 0.1)])
def test_changeover_closed_forms_and_identities(rin, rout):
    wpot, wacc, wacc1 = _w()
    c = (rout - rin) / (rout + rin)
    assert wacc(0.5 * rin, rin, rout) == 1.0 and wacc(2 * rout, rin, rout) == 0.0
    assert wpot(0.5 * rin, rin, rout) == 1.0
    # outside r_out the potential weight is r (1+c)/r_out: the soft part is
    # the plain -Gm/r there, continuous at r_out
    assert wpot(3 * rout, rin, rout) == pytest.approx(3 * (1 + c), rel=1e-14)
    assert wpot(rout * (1 - 1e-12), rin, rout) == pytest.approx(1 + c, rel=1e-9)
    assert wacc(rin * (1 + 1e-12), rin, rout) == pytest.approx(1.0, abs=1e-9)
    rs = np.linspace(rin * 1.01, rout * 0.99, 57)
    h = 1e-4 * (rout - rin)

    def d5(f, r):  # five-point central difference, error O(h^4)
        return (-f(r + 2 * h, rin, rout) + 8 * f(r + h, rin, rout) - 8 * f(r - h, rin, rout)
                + f(r - 2 * h, rin, rout)) / (12 * h)

    for r in rs:
        # force weight = potential weight - r d(potential weight)/dr
        assert wacc(r, rin, rout) == pytest.approx(wpot(r, rin, rout) - r * d5(wpot, r), abs=1e-9)
        # calcAcc1W is dW_acc/dt for dr/dt = rdot
        assert wacc1(r, 0.37, rin, rout) == pytest.approx(d5(wacc, r) * 0.37, rel=1e-7,
                                                          abs=1e-9 / (rout - rin))


def test_changeover_two_picks_larger_rout():
    L = O.orc()
    r = 0.05
    assert L.orc_changeover_wacc_two(r, 0.01, 0.1, 0.02, 0.2) == L.orc_changeover_wacc(r, 0.02, 0.2)
    assert L.orc_changeover_wacc_two(r, 0.02, 0.2, 0.01, 0.1) == L.orc_changeover_wacc(r, 0.02, 0.2)
    assert L.orc_changeover_wpot_two(r, 0.01, 0.1, 0.01, 0.1) == L.orc_changeover_wpot(r, 0.01, 0.1)


def test_changeover_setr_mass_scaling():
    import ctypes as C
    L = O.orc()
    a, b = C.c_double(), C.c_double()
    L.orc_changeover_setr(8.0, 0.1, 1.0, C.byref(a), C.byref(b))
    assert a.value == pytest.approx(0.2) and b.value == pytest.approx(2.0)
    L.orc_changeover_setr(0.5, 0.1, 1.0, C.byref(a), C.byref(b))
    assert a.value == pytest.approx(0.1) and b.value == pytest.approx(1.0)


@pytest.mark.parametrize("eps", [0.0, 0.3, 2.0])
def test_changeover_correction_gives_weighted_newton(eps):
    """Linear-cutoff pair + the correction of src/hard.hpp:1404-1476 =
    (1 - W_acc) Newton for acc and -Gm (1 - W_pot)/r for the soft potential,
    with r the softened distance sqrt(|d|^2 + eps^2) (eps in units of r_in)."""
    rng = np.random.default_rng(8)
    G, r_out, r_in = 1.0, 0.01, 0.001
    eps2 = (eps * r_in) ** 2
    wpot, wacc, _ = _w()
    for _ in range(20):
        xi = rng.normal(size=3) * 0.01
        xj = xi + rng.normal(size=3) * 0.006
        mi, mj = 1e-4, 3e-4
        a_lin, p_lin, _ = _pair(xi, xj, mj, eps2=eps2, rout2=r_out ** 2, G=G)
        acc = a_lin.copy()
        ps = np.array([p_lin])
        pt = np.array([p_lin])
        O.orc().orc_changeover_correct_one(
            np.ascontiguousarray(xi), mi, r_in, r_out, 1, 1, np.ascontiguousarray(xj), np.array([mj]),
            np.array([r_in]), np.array([r_out]), np.array([2], np.int64), eps2, r_out, G, acc, ps, pt, None, None)
        d = xi - xj
        r = np.sqrt(d @ d + eps2)
        self_term = G * mi / r_out
        np.testing.assert_allclose(acc, -(1 - wacc(r, r_in, r_out)) * G * mj * d / r ** 3, rtol=1e-10, atol=1e-300)
        assert ps[0] - self_term == pytest.approx(-G * mj * (1 - wpot(r, r_in, r_out)) / r, rel=1e-10, abs=1e-15)
        assert pt[0] - self_term == pytest.approx(-G * mj / r, rel=1e-12)


# ---- changeover mode over walks (src/hard.hpp:1404-1476,1655-1691) --------------

def _co_two_body(r, rin=(0.001, 0.001), rout=(0.01, 0.01), m=(2e-4, 3e-4), rs=(0.015, 0.015), G=1.3,
                 r_out_glb=0.01, group=None):
    """Particles A (id 10) and B (id 20) at distance r along a random axis;
    walk 0 holds A as i, walk 1 holds B, both lists hold A and B; group
    [2][2] their group data (mass_backup, status) or None (singles)."""
    rng = np.random.default_rng(int(r * 1e7) % 1000)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    xa = np.array([0.1, -0.2, 0.3])
    pos = np.array([xa, xa + r * u])
    acc, ps, pt, nnb, amb = O.soft_force_walks_co(
        [1, 1], [2, 2], [0, 0], pos, np.array(rs), np.array([10, 20]), [0, 1, 0, 1], [], pos, np.array(m),
        np.array(rs), np.array([10, 20]), np.array(rin), np.array(rout), [], [], [], 0.0, r_out_glb ** 2, G,
        epj_group=group)
    return u, acc, ps, pt, nnb


LARGE = float(np.finfo(np.float32).max) / 16  # an unset c.m. address: status = -LARGE (src/ptcl.hpp:219-220)


@pytest.mark.parametrize("r", [0.0004, 0.004, 0.012])
def test_co_member_neighbour(r):
    """B is a group member whose c.m. holds its mass: its tree mass is zero
    and its backup mass m_b is used for the potential
    (src/hard.hpp:1459-1464): A gets no acc from B, pot_soft = -G m_b (1 -
    W_pot)/r, pot_tot = -G m_b/r; B (with a c.m. address) gets no self
    term, which its zero mass makes zero anyway."""
    wpot, wacc, _ = _w()
    G, mb, ma = 1.3, 3e-4, 2e-4
    u, acc, ps, pt, _ = _co_two_body(r, m=(ma, 0.0), G=G, group=[[0.0, 0.0], [mb, -7.0]])
    wa, wp = wacc(r, 0.001, 0.01), wpot(r, 0.001, 0.01)
    assert np.all(acc[0] == 0.0)
    assert ps[0] == pytest.approx(-G * mb * (1 - wp) / r, rel=1e-12, abs=1e-300)
    assert pt[0] == pytest.approx(-G * mb / r, rel=1e-12)
    np.testing.assert_allclose(acc[1], -G * ma * (1 - wa) * u / r ** 2, rtol=1e-12, atol=1e-300)
    assert ps[1] == pytest.approx(-G * ma * (1 - wp) / r, rel=1e-12, abs=1e-300)
    assert pt[1] == pytest.approx(-G * ma / r, rel=1e-12)


def test_co_member_without_cm_keeps_its_mass_and_loses_the_self_term():
    """A binary without artificial particles: members keep their tree mass,
    have no c.m. address (status -LARGE_FLOAT), and their self term is
    removed as for singles (src/hard.hpp:1665-1673); with backup mass equal
    to the tree mass the pair gives the single-pair closed forms."""
    G, r = 1.3, 0.004
    m = (2e-4, 3e-4)
    single = _co_two_body(r, m=m, G=G)
    grp = _co_two_body(r, m=m, G=G, group=[[m[0], -LARGE], [m[1], -LARGE]])
    for x, y in zip(single[1:], grp[1:]):
        np.testing.assert_allclose(x, y, rtol=1e-13, atol=1e-300)


@pytest.mark.parametrize("r", [0.0004, 0.004, 0.012])
def test_co_artificial_neighbour(r):
    """B is an artificial particle (status > 0): A's acc gets the changeover
    weight with B's mass, but B leaves no potential on A (the tree's
    linear-cutoff term is cancelled, src/hard.hpp:1466-1470); B keeps its
    own self term -G m_B / r_out (only singles and members without a c.m.
    lose it, src/hard.hpp:1669)."""
    wpot, wacc, _ = _w()
    G = 1.3
    m = (2e-4, 3e-4)
    u, acc, ps, pt, _ = _co_two_body(r, m=m, G=G, group=[[0.0, 0.0], [0.0, 3.0]])
    wa, wp = wacc(r, 0.001, 0.01), wpot(r, 0.001, 0.01)
    np.testing.assert_allclose(acc[0], G * m[1] * (1 - wa) * u / r ** 2, rtol=1e-12, atol=1e-300)
    assert ps[0] == 0.0 and pt[0] == 0.0
    self_b = -G * m[1] / 0.01
    assert ps[1] == pytest.approx(self_b - G * m[0] * (1 - wp) / r, rel=1e-12)
    assert pt[1] == pytest.approx(self_b - G * m[0] / r, rel=1e-12)


@pytest.mark.parametrize("r", [0.0004, 0.004, 0.009, 0.012])
def test_co_two_body_closed_forms(r):
    """Within the neighbour radius: acc = G m (1 - W_acc) dx / r^3, pot_soft =
    -G m (1 - W_pot) / r, pot_tot = -G m / r, no self term; W from the larger
    r_out (A's here, B's in the mirrored pair)."""
    wpot, wacc, _ = _w()
    G = 1.3
    rin, rout, m = (0.001, 0.0005), (0.01, 0.005), (2e-4, 3e-4)
    u, acc, ps, pt, nnb = _co_two_body(r, rin=rin, rout=rout, m=m, G=G)
    for i, j, sgn in ((0, 1, 1.0), (1, 0, -1.0)):
        wa, wp = wacc(r, rin[0], rout[0]), wpot(r, rin[0], rout[0])  # A has the larger r_out
        np.testing.assert_allclose(acc[i], sgn * G * m[j] * (1 - wa) * u / r ** 2, rtol=1e-12, atol=1e-300)
        assert ps[i] == pytest.approx(-G * m[j] * (1 - wp) / r, rel=1e-12, abs=1e-300)
        assert pt[i] == pytest.approx(-G * m[j] / r, rel=1e-12)
        assert nnb[i] == 2
    if r < rin[0]:
        assert np.all(acc == 0.0) and np.all(ps == 0.0)


def test_co_beyond_neighbour_radius_is_linear_cutoff_without_self():
    G, r = 1.3, 0.02  # > 0.99 * 0.015
    u, acc, ps, pt, nnb = _co_two_body(r, G=G)
    m = (2e-4, 3e-4)
    np.testing.assert_allclose(acc[0], G * m[1] * u / r ** 2, rtol=1e-12)
    assert ps[0] == pytest.approx(-G * m[1] / r, rel=1e-12) and pt[0] == ps[0]
    assert nnb[0] == 1  # self only (r > r_search)


def test_co_soft_potential_is_potential_of_acc():
    """-grad_i pot_soft = acc for a small cluster whose pairs are all within
    the neighbour radius, with unequal mass-scaled changeover radii (a wrong
    sign, a W_acc/W_pot mix-up or a wrong pair rule breaks this)."""
    rng = np.random.default_rng(12)
    n = 6
    base = rng.normal(size=(n, 3)) * 0.004
    m = rng.uniform(1e-5, 1e-3, n)
    fac = np.maximum(np.cbrt(m / m.mean()), 1.0)
    rin, rout = 0.002 * fac, 0.012 * fac
    rs = 2.5 * rout
    ids = np.arange(100, 100 + n)
    G = 0.7

    def evaluate(pos):
        lst = np.tile(np.arange(n), n)
        return O.soft_force_walks_co(np.ones(n, int), np.full(n, n), np.zeros(n, int), pos, rs, ids, lst, [], pos, m,
                                     rs, ids, rin, rout, [], [], [], 0.0, 0.01 ** 2, G)

    acc, ps, pt, _, _ = evaluate(base)
    h = 1e-7
    for i in range(n):
        for d in range(3):
            vals = []
            for s in (-2, -1, 1, 2):
                p = base.copy()
                p[i, d] += s * h
                vals.append(evaluate(p)[1][i])
            grad = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
            assert -grad == pytest.approx(acc[i, d], rel=2e-6, abs=1e-6 * np.abs(acc[i]).max())
    # and pot_tot is the plain Newtonian potential over the other particles
    for i in range(n):
        d = np.linalg.norm(base - base[i], axis=1)
        assert pt[i] == pytest.approx(-G * np.sum(np.delete(m / np.where(d > 0, d, 1), i)), rel=1e-12)


def test_co_missing_own_entry_is_an_error():
    pos = np.array([[0.0, 0.0, 0.0], [0.005, 0.0, 0.0]])
    with pytest.raises(RuntimeError):
        O.soft_force_walks_co([1], [1], [0], pos[:1], [0.015], [7], [1], [], pos, [1e-4, 1e-4], [0.015, 0.015],
                              [6, 8], [0.001, 0.001], [0.01, 0.01], [], [], [], 0.0, 1e-4)


def _brute_ngb_pairs(pos, rs):
    """All (i, j) with |x_i - x_j|^2 < max(rs_i, rs_j)^2 in FP64 (self
    included), by a k-d tree query widened to the largest radius and the
    exact FP64 filter; sorted by (i, j)."""
    from scipy.spatial import cKDTree
    pairs = cKDTree(pos).query_pairs(float(rs.max()) * (1 + 1e-9), output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    d = pos[i] - pos[j]
    r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    r = np.maximum(rs[i], rs[j])
    keep = r2 < r * r
    i, j = i[keep], j[keep]
    n = len(pos)
    selfp = np.arange(n)[rs > 0]
    a = np.concatenate([i, j, selfp])
    b = np.concatenate([j, i, selfp])
    o = np.lexsort((b, a))
    return a[o], b[o]


@pytest.mark.parametrize("f_bin,theta", [(0.5, 0.3), (0.0, 0.7)])
def test_ngb_oracle_equals_brute_force(f_bin, theta):
    """The neighbour-list restatement on the symmetric-search tree lists
    equals an all-pairs FP64 search: the restatement is right and the
    lists hold every neighbour (mass-scaled r_search, binaries)."""
    from petar_amd import host
    m, x, _ = host.cluster(12000, "plummer", kroupa=True, f_bin=f_bin, seed=17)
    rs = 1.5 * R_OUT * np.maximum(np.cbrt(m / m.mean()), 1.0)
    L = host.build_lists(x, m, rs, theta=theta)
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(L.epi_begin, L.n_epi)])
    assert np.array_equal(idx, np.arange(len(m)))
    cnt, off, ngb = O.ngb_walks(L.n_epi, L.n_epj, L.epj_pos[idx], L.epj_rs[idx], L.id_epj, L.epj_pos, L.epj_rs)
    a = np.repeat(np.arange(len(m)), cnt)
    o = np.lexsort((ngb, a))
    ba, bb = _brute_ngb_pairs(L.epj_pos, L.epj_rs)
    assert np.array_equal(a[o], ba) and np.array_equal(ngb[o], bb)
    assert cnt.min() >= 1 and (cnt > 1).sum() > 100  # self always, binaries give real neighbours


def test_ngb_oracle_catches_a_dropped_neighbour():
    """A planted mistake (the last neighbour of one particle dropped) is
    seen by the brute-force comparison."""
    from petar_amd import host
    m, x, _ = host.cluster(3000, "plummer", kroupa=True, f_bin=0.5, seed=18)
    rs = np.full(len(m), 1.5 * R_OUT)
    L = host.build_lists(x, m, rs, theta=0.3)
    cnt, off, ngb = O.ngb_walks(L.n_epi, L.n_epj, L.epj_pos, L.epj_rs, L.id_epj, L.epj_pos, L.epj_rs)
    k = int(np.argmax(cnt))
    keep = np.ones(ngb.size, bool)
    keep[off[k + 1] - 1] = False
    cnt2 = cnt.copy()
    cnt2[k] -= 1
    a = np.repeat(np.arange(len(m)), cnt2)
    o = np.lexsort((ngb[keep], a))
    ba, bb = _brute_ngb_pairs(L.epj_pos, L.epj_rs)
    assert not (a.size == ba.size and np.array_equal(a[o], ba) and np.array_equal(ngb[keep][o], bb))
