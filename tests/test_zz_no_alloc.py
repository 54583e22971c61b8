"""The allocation-free entry point pth_resident_eval_noalloc against the
allocating resident evaluation (pth_resident_create + run + results) on the
same inputs: acc, pot and nnb bit-identical, on two calls in a row (the
second on a smaller cluster, so it runs over what the first left in the
scratch array), with quadrupole and with monopole superparticles; and the
errors it returns without touching the GPU (an evaluation larger than the
scratch array, changeover mode, the plain form, an index out of range)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from petar_amd import soft_force as sf
from tests._common import R_OUT, c1_lists, rel_err

pytestmark = pytest.mark.gpu


def _resident(ctx, L):
    r = sf.Resident.from_lists(ctx, L)
    r.run()
    out = r.results()
    r.close()
    return out


def _same(a, b, label):
    for k, name in enumerate(("acc", "pot", "nnb")):
        d = int(np.count_nonzero(a[k] != b[k]))
        print(f"[{label}] {name}: {d} entries differ", flush=True)
        assert np.array_equal(a[k], b[k]), f"{label}: {name} differs in {d} entries"


@pytest.mark.parametrize("mono", [0, 1])
def test_noalloc_equals_resident(mono):
    big = c1_lists(16384, seed=1, f_bin=0.5)
    small = c1_lists(3000, seed=2)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, spj_monopole=mono)
    for label, L in (("16384", big), ("3000", small)):
        print(f"[noalloc mono={mono} N={label}] {L.n_walk} walks", flush=True)
        got = sf.eval_noalloc_lists(ctx, L)
        _same(got, _resident(ctx, L), f"mono={mono} N={label}")
        if not mono:
            racc, rpot, _ = O.soft_force_walks(**O.lists_args(L), eps2=0.0, rout2=R_OUT ** 2, G=1.0, nnb_mode=0)
            ea = rel_err(got[0], racc)
            print(f"[noalloc N={label}] acc err vs FP64 oracle median {np.median(ea):.2e}", flush=True)
            assert np.median(ea) < 1e-6
    # the first cluster again after the smaller one: same results as its first call
    again = sf.eval_noalloc_lists(ctx, big)
    _same(again, _resident(ctx, big), f"mono={mono} N=16384 again")
    ctx.close()


def test_noalloc_errors():
    L = c1_lists(3000, seed=2)
    huge = c1_lists(200000, seed=3)
    ctx = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0)
    with pytest.raises(sf.PthError, match="bytes of scratch"):
        sf.eval_noalloc_lists(ctx, huge)
    bad = L.id_epj.copy()
    bad[0] = len(L.epj_mass)
    idx = np.concatenate([np.arange(b, b + n) for b, n in zip(L.epi_begin, L.n_epi)])
    with pytest.raises(sf.PthError, match="out of range"):
        sf.eval_noalloc(ctx, L.n_epi, L.n_epj, L.n_spj, L.epj_pos[idx], L.epj_rs[idx], bad, L.id_spj, L.epj_pos,
                        L.epj_mass, L.epj_rs, L.spj_pos, L.spj_mass, L.spj_quad)
    # after the refusals the entry point still works
    _same(sf.eval_noalloc_lists(ctx, L), _resident(ctx, L), "after errors")
    ctx.close()
    for kw, msg in ((dict(changeover=1), "linear-cutoff mode only"), (dict(exec_mode=sf.EXEC_PLAIN), "fused form only")):
        c2 = sf.Context(0, eps2=0.0, rcut2=R_OUT ** 2, G=1.0, **kw)
        with pytest.raises(sf.PthError, match=msg):
            sf.eval_noalloc_lists(c2, L)
        c2.close()
