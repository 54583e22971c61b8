"""Host side of the soft-force path: initial conditions and the FDPS-style
tree walk that produces interaction lists (C ABI in include/petar_host.h,
library petar_amd/lib/libpetar_host.so).

The lists come out in the form PeTar's FDPS multi-walk driver hands to the
GPU functor (src/force_gpu_cuda.cu:535-549): per walk (i-group) a
contiguous EPI range, and int32 index lists into one EPJ array (particles in
tree order) and one SPJ array (tree cells with quadrupole moments).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load libpetar_host.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(_LIB_DIR, "libpetar_host.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
    L = C.CDLL(path)
    L.pth_host_last_error.restype = C.c_char_p
    L.pth_mt_new.restype = C.c_void_p
    L.pth_mt_new.argtypes = [C.c_uint]
    L.pth_mt_free.argtypes = [C.c_void_p]
    L.pth_mt_int32.restype = C.c_uint
    L.pth_mt_int32.argtypes = [C.c_void_p]
    L.pth_mt_res53.restype = C.c_double
    L.pth_mt_res53.argtypes = [C.c_void_p]
    L.pth_mt_real2.restype = C.c_double
    L.pth_mt_real2.argtypes = [C.c_void_p]
    L.pth_ic_plummer_equal.argtypes = [C.c_long, C.c_double, C.c_double, C.c_uint, _dp, _dp, C.c_void_p]
    L.pth_ic_kroupa_masses.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_uint, _dp]
    L.pth_ic_cluster.argtypes = [C.c_long, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_uint, _dp, _dp, C.c_void_p]
    L.pth_tree_build.argtypes = [C.POINTER(C.c_void_p), C.c_long, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int]
    L.pth_tree_walk.argtypes = [C.c_void_p, C.c_double]
    L.pth_tree_sizes.argtypes = [C.c_void_p] + [C.POINTER(C.c_long)] * 4
    L.pth_tree_order.argtypes = [C.c_void_p, _lp]
    L.pth_tree_nodes.argtypes = [C.c_void_p, _dp, _dp, _dp, _lp, _lp, np.ctypeslib.ndpointer(np.int32)]
    L.pth_tree_groups.argtypes = [C.c_void_p, _lp, _ip, _ip, _ip]
    L.pth_tree_lists.argtypes = [C.c_void_p, _ip, _ip]
    L.pth_tree_free.argtypes = [C.c_void_p]
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib().pth_host_last_error().decode()}")


class MT19937:
    """MT19937 as FDPS's MTTS (init_genrand / genrand_res53 / genrand_real2)."""

    def __init__(self, seed: int):
        self._h = lib().pth_mt_new(seed)

    def int32(self) -> int:
        return lib().pth_mt_int32(self._h)

    def res53(self) -> float:
        return lib().pth_mt_res53(self._h)

    def real2(self) -> float:
        return lib().pth_mt_real2(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pth_mt_free(self._h)
            self._h = None


def plummer_equal(n: int, mass_glb: float = 1.0, eng: float = -0.25, seed: int = 0,
                  with_vel: bool = True):
    """Equal-mass Plummer model, the makePlummerModel recipe
    (src/particle_distribution_generator.hpp:173-250)."""
    mass = np.empty(n)
    pos = np.empty((n, 3))
    vel = np.empty((n, 3)) if with_vel else None
    _check(lib().pth_ic_plummer_equal(n, mass_glb, eng, seed, mass, pos,
                                      vel.ctypes.data if with_vel else None), "plummer")
    return mass, pos, vel


def kroupa_masses(n: int, m_lo=0.08, m_hi=150.0, m_tot=1.0, seed: int = 0):
    m = np.empty(n)
    _check(lib().pth_ic_kroupa_masses(n, m_lo, m_hi, m_tot, seed, m), "kroupa")
    return m


def cluster(n: int, model: str = "plummer", w0: float = 7.0, kroupa: bool = True,
            f_bin: float = 0.0, a_lo: float = 1e-5, a_hi: float = 1e-2,
            segregate: bool = False, seed: int = 0, with_vel: bool = False):
    """Star cluster in N-body units (G=1, M=1, E=-1/4), see pth_ic_cluster."""
    mass = np.empty(n)
    pos = np.empty((n, 3))
    vel = np.empty((n, 3)) if with_vel else None
    mid = {"plummer": 0, "king": 1}[model]
    _check(lib().pth_ic_cluster(n, mid, w0, int(kroupa), f_bin, a_lo, a_hi, int(segregate), seed,
                                mass, pos, vel.ctypes.data if with_vel else None), "cluster")
    return mass, pos, vel


@dataclass
class Lists:
    """Interaction lists of one soft-force evaluation in FDPS multi-walk form.

    EPJ records are the particles in tree order (epj_pos/epj_mass/epj_rs),
    SPJ records are the tree cells (spj_pos = centre of mass, spj_mass,
    spj_quad = xx,yy,zz,xy,xz,yz).  Walk w owns EPI particles
    [epi_begin[w], epi_begin[w]+n_epi[w]) of the EPJ array (i-groups are tree
    cells, so their particles are contiguous) and n_epj[w] / n_spj[w]
    consecutive entries of id_epj / id_spj.
    """
    order: np.ndarray
    epj_pos: np.ndarray
    epj_mass: np.ndarray
    epj_rs: np.ndarray
    spj_pos: np.ndarray
    spj_mass: np.ndarray
    spj_quad: np.ndarray
    node_begin: np.ndarray
    node_end: np.ndarray
    node_level: np.ndarray
    epi_begin: np.ndarray
    n_epi: np.ndarray
    n_epj: np.ndarray
    n_spj: np.ndarray
    id_epj: np.ndarray
    id_spj: np.ndarray
    theta: float = 0.3
    extra: dict = field(default_factory=dict)

    @property
    def n_walk(self) -> int:
        return int(self.n_epi.size)

    @property
    def epi_pos(self) -> np.ndarray:
        return self.epj_pos  # EPI of walk w = EPJ rows epi_begin[w]..

    def interactions(self):
        """(N_e, N_s): EPJ and SPJ pair interactions, sum over walks of n_i*n_j."""
        ne = int(np.dot(self.n_epi.astype(np.int64), self.n_epj.astype(np.int64)))
        ns = int(np.dot(self.n_epi.astype(np.int64), self.n_spj.astype(np.int64)))
        return ne, ns

    def walk_offsets(self):
        e = np.zeros(self.n_walk + 1, np.int64)
        s = np.zeros(self.n_walk + 1, np.int64)
        np.cumsum(self.n_epj, out=e[1:])
        np.cumsum(self.n_spj, out=s[1:])
        return e, s


def build_lists(pos, mass, r_search, theta=0.3, n_leaf_limit=20, n_group_limit=512,
                nthreads=0) -> Lists:
    """Tree + group walk (see petar_amd/csrc/host/tree.cpp)."""
    L = lib()
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    rs = np.ascontiguousarray(np.broadcast_to(r_search, (n,)), dtype=np.float64)
    h = C.c_void_p()
    _check(L.pth_tree_build(C.byref(h), n, pos.ravel(), mass, rs, n_leaf_limit, n_group_limit,
                            nthreads), "tree build")
    try:
        _check(L.pth_tree_walk(h, theta), "tree walk")
        nn, ng, ne, ns = C.c_long(), C.c_long(), C.c_long(), C.c_long()
        _check(L.pth_tree_sizes(h, C.byref(nn), C.byref(ng), C.byref(ne), C.byref(ns)), "sizes")
        nn, ng, ne, ns = nn.value, ng.value, ne.value, ns.value
        order = np.empty(n, np.int64)
        _check(L.pth_tree_order(h, order), "order")
        smass = np.empty(nn)
        scom = np.empty((nn, 3))
        squad = np.empty((nn, 6))
        nb = np.empty(nn, np.int64)
        ne_ = np.empty(nn, np.int64)
        lv = np.empty(nn, np.int32)
        _check(L.pth_tree_nodes(h, smass, scom.ravel(), squad.ravel(), nb, ne_, lv), "nodes")
        eb = np.empty(ng, np.int64)
        n_epi = np.empty(ng, np.int32)
        n_epj = np.empty(ng, np.int32)
        n_spj = np.empty(ng, np.int32)
        _check(L.pth_tree_groups(h, eb, n_epi, n_epj, n_spj), "groups")
        id_epj = np.empty(max(ne, 1), np.int32)
        id_spj = np.empty(max(ns, 1), np.int32)
        _check(L.pth_tree_lists(h, id_epj, id_spj), "lists")
    finally:
        L.pth_tree_free(h)
    return Lists(order=order, epj_pos=np.ascontiguousarray(pos[order]),
                 epj_mass=np.ascontiguousarray(mass[order]), epj_rs=np.ascontiguousarray(rs[order]),
                 spj_pos=scom, spj_mass=smass, spj_quad=squad, node_begin=nb, node_end=ne_,
                 node_level=lv, epi_begin=eb, n_epi=n_epi, n_epj=n_epj, n_spj=n_spj,
                 id_epj=id_epj[:ne], id_spj=id_spj[:ns], theta=theta)
