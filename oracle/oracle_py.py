"""ctypes bindings of the checkers: oracle/liboracle.so (FP64 restatement)
and oracle/_ref/libphantomquad_*.so (the reference x86 kernel built from
source).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_orc = None
_refs = {}


def orc():
    global _orc
    if _orc is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle` or __graft_entry__.build()")
        L = C.CDLL(path)
        wl = [C.c_int, _ip, _ip, _ip]
        # spj_quad is a raw pointer: NULL selects the monopole SPJ pair
        L.orc_soft_force_walks.argtypes = wl + [_dp, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _dp, C.c_void_p,
                                                C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp, _lp]
        L.orc_epj_walks.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _ip, _dp, _dp, _dp, C.c_double,
                                    C.c_double, C.c_double, C.c_int, _dp, _dp, _lp]
        L.orc_spj_walks.argtypes = [C.c_int, _ip, _ip, _dp, _ip, _dp, _dp, C.c_void_p, C.c_double,
                                    C.c_double, _dp, _dp]
        L.orc_nnb_ambiguous_walks.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _ip, _dp, _dp, C.c_void_p,
                                              C.c_double, C.c_double, _lp]
        L.orc_direct_linear_cutoff.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_void_p, C.c_double,
                                               C.c_double, C.c_double, C.c_int, _dp, _dp, _lp]
        L.orc_direct_newton.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_void_p, C.c_double, C.c_double,
                                        _dp, _dp]
        L.orc_moments.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        _llp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        L.orc_soft_force_walks_co.argtypes = wl + [_dp, _dp, _llp, _ip, _ip, _dp, _dp, _dp, _llp, _dp, _dp,
                                                   _dp, _dp, C.c_void_p] + [C.c_double] * 5 + [_dp, _dp, _dp, _lp,
                                                                                               _lp, C.c_void_p]
        L.orc_ngb_walks.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _ip, _dp, _dp, _llp, C.c_void_p, C.c_void_p]
        L.orc_ngb_walks.restype = C.c_longlong
        L.orc_changeover_setr.argtypes = [C.c_double] * 3 + [C.POINTER(C.c_double)] * 2
        for f in ("orc_changeover_wpot", "orc_changeover_wacc"):
            getattr(L, f).argtypes = [C.c_double] * 3
            getattr(L, f).restype = C.c_double
        L.orc_changeover_wacc1.argtypes = [C.c_double] * 4
        L.orc_changeover_wacc1.restype = C.c_double
        for f in ("orc_changeover_wpot_two", "orc_changeover_wacc_two"):
            getattr(L, f).argtypes = [C.c_double] * 5
            getattr(L, f).restype = C.c_double
        L.orc_changeover_correct_one.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_longlong,
                                                 C.c_int, _dp, _dp, _dp, _dp, _lp, C.c_double, C.c_double,
                                                 C.c_double, _dp, _dp, _dp, C.c_void_p, C.c_void_p]
        _orc = L
    return _orc


def cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return "avx512f" in fl and "avx512dq" in fl
    except OSError:
        pass
    return False


def ref_available(isa: str | None = None) -> bool:
    isa = isa or ("avx512" if cpu_has_avx512() else "avx2")
    return os.path.exists(os.path.join(HERE, "_ref", f"libphantomquad_{isa}.so"))


def ref(isa: str | None = None):
    """The reference phantomquad kernel; isa 'avx512' or 'avx2' (default:
    the best one this CPU runs)."""
    isa = isa or ("avx512" if cpu_has_avx512() else "avx2")
    if isa not in _refs:
        path = os.path.join(HERE, "_ref", f"libphantomquad_{isa}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: it is built from /root/reference by `make -C oracle ref`")
        L = C.CDLL(path)
        L.ref_isa_bits.restype = C.c_int
        L.ref_soft_force_walks.argtypes = [C.c_int, _ip, _ip, _ip, _dp, _dp, C.c_void_p, _ip, _ip, _dp,
                                           _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                           C.c_int, C.c_int, C.c_int, _dp, _dp, _lp]
        _refs[isa] = L
    return _refs[isa]


# ---------------------------------------------------------------------------
# convenience wrappers over a "walk set" (numpy arrays)

def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def soft_force_walks(n_epi, n_epj, n_spj, epi_pos, epi_rs, id_epj, id_spj, epj_pos, epj_mass, epj_rs,
                     spj_pos, spj_mass, spj_quad, eps2, rout2, G=1.0, nnb_mode=0, part="all", mono=False):
    """FP64 oracle over walks; part in {'all', 'epj', 'spj'}; mono selects the
    monopole SPJ pair (src/soft_force.hpp:125-158; spj_quad is ignored)."""
    n_epi, n_epj, n_spj = _i32(n_epi), _i32(n_epj), _i32(n_spj)
    ni = int(n_epi.sum())
    acc = np.zeros((ni, 3))
    pot = np.zeros(ni)
    nnb = np.zeros(ni, np.int64)
    epi_pos, epi_rs = _f64(epi_pos).reshape(-1, 3), _f64(epi_rs)
    id_epj = _i32(id_epj) if len(id_epj) else np.zeros(1, np.int32)
    id_spj = _i32(id_spj) if len(id_spj) else np.zeros(1, np.int32)
    epj_pos = _f64(epj_pos).reshape(-1, 3) if len(epj_pos) else np.zeros((1, 3))
    epj_mass = _f64(epj_mass) if len(epj_mass) else np.zeros(1)
    epj_rs = _f64(epj_rs) if len(epj_rs) else np.zeros(1)
    spj_pos = _f64(spj_pos).reshape(-1, 3) if len(spj_pos) else np.zeros((1, 3))
    spj_mass = _f64(spj_mass) if len(spj_mass) else np.zeros(1)
    spj_quad = _f64(spj_quad).reshape(-1, 6) if len(spj_quad) else np.zeros((1, 6))
    L = orc()
    nw = n_epi.size
    if ni == 0:
        return acc, pot, nnb
    qp = None if mono else spj_quad.ctypes.data
    if part == "all":
        rc = L.orc_soft_force_walks(nw, n_epi, n_epj, n_spj, epi_pos.ravel(), epi_rs, id_epj, id_spj,
                                    epj_pos.ravel(), epj_mass, epj_rs, spj_pos.ravel(), spj_mass,
                                    qp, eps2, rout2, G, nnb_mode, acc.ravel(), pot, nnb)
    elif part == "epj":
        rc = L.orc_epj_walks(nw, n_epi, n_epj, epi_pos.ravel(), epi_rs, id_epj, epj_pos.ravel(), epj_mass,
                             epj_rs, eps2, rout2, G, nnb_mode, acc.ravel(), pot, nnb)
    else:
        rc = L.orc_spj_walks(nw, n_epi, n_spj, epi_pos.ravel(), id_spj, spj_pos.ravel(), spj_mass,
                             qp, eps2, G, acc.ravel(), pot)
    if rc != 0:
        raise RuntimeError("oracle failed")
    return acc, pot, nnb


def ref_soft_force_walks(n_epi, n_epj, n_spj, epi_pos, epi_rs, id_epj, id_spj, epj_pos, epj_mass, epj_rs,
                         spj_pos, spj_mass, spj_quad, eps2, rout2, G=1.0, part="all", nthreads=0,
                         isa=None, epi_type=None, mono=False):
    """The reference phantomquad kernel over walks (soft_force.hpp wrapper
    semantics); mono: monopole superparticles (CalcForceEpSpMonoSimd)."""
    n_epi, n_epj, n_spj = _i32(n_epi), _i32(n_epj), _i32(n_spj)
    ni = int(n_epi.sum())
    acc = np.zeros((ni, 3))
    pot = np.zeros(ni)
    nnb = np.zeros(ni, np.int64)
    if ni == 0:
        return acc, pot, nnb
    id_epj = _i32(id_epj) if len(id_epj) else np.zeros(1, np.int32)
    id_spj = _i32(id_spj) if len(id_spj) else np.zeros(1, np.int32)
    epj_pos = _f64(epj_pos).reshape(-1, 3) if len(epj_pos) else np.zeros((1, 3))
    epj_mass = _f64(epj_mass) if len(epj_mass) else np.zeros(1)
    epj_rs = _f64(epj_rs) if len(epj_rs) else np.zeros(1)
    spj_pos = _f64(spj_pos).reshape(-1, 3) if len(spj_pos) else np.zeros((1, 3))
    spj_mass = _f64(spj_mass) if len(spj_mass) else np.zeros(1)
    spj_quad = _f64(spj_quad).reshape(-1, 6) if len(spj_quad) else np.zeros((1, 6))
    et = None if epi_type is None else _i32(epi_type)
    L = ref(isa)
    rc = L.ref_soft_force_walks(n_epi.size, n_epi, n_epj, n_spj, _f64(epi_pos).ravel(), _f64(epi_rs),
                                et.ctypes.data if et is not None else None, id_epj, id_spj,
                                epj_pos.ravel(), epj_mass, epj_rs, spj_pos.ravel(), spj_mass,
                                spj_quad.ravel(), eps2, rout2, G, int(part in ("all", "epj")),
                                (2 if mono else 1) if part in ("all", "spj") else 0, nthreads, acc.ravel(), pot,
                                nnb)
    if rc != 0:
        raise RuntimeError(f"reference kernel failed ({rc})")
    return acc, pot, nnb


def soft_force_walks_co(n_epi, n_epj, n_spj, epi_pos, epi_rs, epi_id, id_epj, id_spj, epj_pos, epj_mass, epj_rs,
                        epj_id, epj_rin, epj_rout, spj_pos, spj_mass, spj_quad, eps2, rout2, G=1.0, ngb_factor=0.99,
                        amb_rel=1e-5, mono=False, epj_group=None):
    """FP64 oracle of the changeover mode: acc, pot_soft, pot_tot, nnb and the
    per-i count of neighbour pairs at the search radius within amb_rel."""
    n_epi, n_epj, n_spj = _i32(n_epi), _i32(n_epj), _i32(n_spj)
    ni = int(n_epi.sum())
    acc, ps, pt = np.zeros((ni, 3)), np.zeros(ni), np.zeros(ni)
    nnb, amb = np.zeros(ni, np.int64), np.zeros(ni, np.int64)
    if ni == 0:
        return acc, ps, pt, nnb, amb

    def arr(a, dt=np.float64):
        # flat contiguous copy; a one-element dummy stands in for an empty array
        a = np.ascontiguousarray(a, dtype=dt).ravel()
        return a if a.size else np.zeros(1, dt)
    q = None if mono else arr(spj_quad)  # kept alive through the call
    grp = None if epj_group is None else arr(epj_group)
    rc = orc().orc_soft_force_walks_co(
        n_epi.size, n_epi, n_epj, n_spj, arr(epi_pos), arr(epi_rs), arr(epi_id, np.int64),
        arr(id_epj, np.int32), arr(id_spj, np.int32), arr(epj_pos), arr(epj_mass), arr(epj_rs),
        arr(epj_id, np.int64), arr(epj_rin), arr(epj_rout), arr(spj_pos), arr(spj_mass),
        None if q is None else q.ctypes.data, eps2, rout2, G, ngb_factor, amb_rel, acc.ravel(), ps, pt, nnb, amb,
        None if grp is None else grp.ctypes.data)
    if rc != 0:
        raise RuntimeError("orc_soft_force_walks_co: an i-particle has no own EPJ entry in its walk")
    return acc, ps, pt, nnb, amb


def lists_args(lists):
    """Walk-set arguments from a petar_amd.host.Lists object."""
    ni = lists.n_epi
    idx = np.concatenate([np.arange(b, b + n) for b, n in zip(lists.epi_begin, ni)]) if ni.size else np.zeros(0, np.int64)
    return dict(n_epi=ni, n_epj=lists.n_epj, n_spj=lists.n_spj, epi_pos=lists.epj_pos[idx],
                epi_rs=lists.epj_rs[idx], id_epj=lists.id_epj, id_spj=lists.id_spj,
                epj_pos=lists.epj_pos, epj_mass=lists.epj_mass, epj_rs=lists.epj_rs,
                spj_pos=lists.spj_pos, spj_mass=lists.spj_mass, spj_quad=lists.spj_quad)


def nnb_ambiguous(n_epi, n_epj, epi_pos, epi_rs, id_epj, epj_pos, epj_rs, centres=None, eps2=0.0,
                  rel=4.0 * 2.0 ** -24, **_):
    n_epi, n_epj = _i32(n_epi), _i32(n_epj)
    out = np.zeros(int(n_epi.sum()), np.int64)
    if out.size == 0:
        return out
    c = None if centres is None else _f64(centres).reshape(-1, 3)
    orc().orc_nnb_ambiguous_walks(n_epi.size, n_epi, n_epj, _f64(epi_pos).ravel(), _f64(epi_rs),
                                  _i32(id_epj) if len(id_epj) else np.zeros(1, np.int32),
                                  _f64(epj_pos).ravel(), _f64(epj_rs),
                                  c.ctypes.data if c is not None else None, eps2, rel, out)
    return out


def direct_linear_cutoff(pos, mass, rs, sample=None, eps2=0.0, rout2=1e-4, G=1.0, nnb_mode=0):
    pos = _f64(pos).reshape(-1, 3)
    n = pos.shape[0]
    idx = None if sample is None else _i32(sample)
    ns = n if idx is None else idx.size
    acc = np.zeros((ns, 3))
    pot = np.zeros(ns)
    nnb = np.zeros(ns, np.int64)
    orc().orc_direct_linear_cutoff(n, pos.ravel(), _f64(mass), _f64(np.broadcast_to(rs, (n,))), ns,
                                   idx.ctypes.data if idx is not None else None, eps2, rout2, G,
                                   nnb_mode, acc.ravel(), pot, nnb)
    return acc, pot, nnb


def direct_newton(pos, mass, sample=None, eps2=0.0, G=1.0):
    pos = _f64(pos).reshape(-1, 3)
    n = pos.shape[0]
    idx = None if sample is None else _i32(sample)
    ns = n if idx is None else idx.size
    acc = np.zeros((ns, 3))
    pot = np.zeros(ns)
    orc().orc_direct_newton(n, pos.ravel(), _f64(mass), ns, idx.ctypes.data if idx is not None else None,
                            eps2, G, acc.ravel(), pot)
    return acc, pot


def moments(pos, mass):
    pos = _f64(pos).reshape(-1, 3)
    m = np.zeros(1)
    c = np.zeros(3)
    q = np.zeros(6)
    orc().orc_moments(pos.shape[0], pos.ravel(), _f64(mass), m, c, q)
    return m[0], c, q


def ngb_walks(n_epi, n_epj, epi_pos, epi_rs, id_epj, epj_pos, epj_rs):
    """FP64 neighbour lists over walks (src/soft_force.hpp:11-34 on each
    walk's EPJ list): returns (count [ni], offset [ni + 1], ngb) with ngb the
    EPJ indices of every i's neighbours in list order, self included."""
    n_epi, n_epj = _i32(n_epi), _i32(n_epj)
    ni = int(n_epi.sum())
    epi_pos, epi_rs = _f64(epi_pos).reshape(-1, 3), _f64(epi_rs)
    epj_pos, epj_rs = _f64(epj_pos).reshape(-1, 3), _f64(epj_rs)
    ide = _i32(id_epj) if len(id_epj) else np.zeros(1, np.int32)
    cnt = np.zeros(max(ni, 1), np.int64)
    L = orc()
    args = (n_epi.size, n_epi, n_epj, epi_pos, epi_rs, ide, epj_pos, epj_rs, cnt)
    total = L.orc_ngb_walks(*args, None, None)
    off = np.zeros(ni + 1, np.int64)
    np.cumsum(cnt[:ni], out=off[1:])
    ngb = np.zeros(max(total, 1), np.int32)
    L.orc_ngb_walks(*args, off.ctypes.data, ngb.ctypes.data)
    return cnt[:ni], off, ngb[:total]

