"""Python mirror of PeTar's GPU soft-force hooks over libpetar_hip.so.

The FDPS multi-walk interface (src/force_gpu_cuda.hpp:101-168) is kept:

    f = CalcForceWithLinearCutoffMultiWalk(my_rank, eps2, rcut2, G)
    f(tag, n_walk, epi, n_epi, id_epj, n_epj, id_spj, n_spj,
      epj, n_epj_tot, spj, n_spj_tot, send_flag)          # dispatch
    RetrieveForce(tag, n_walk, ni, force)                   # retrieve

with epi/force lists of per-walk numpy record arrays of dtype EPI_SOFT /
FORCE_SOFT and epj/spj record arrays of dtype EPJ_SOFT / SPJ_SOFT, whose
layouts are those of EPISoft, EPJSoft and ForceSoft (src/soft_ptcl.hpp:4-24,
271-376) and of the SPJQuadrupoleInAndOut members PeTar reads
(src/force_gpu_cuda.cu:596-608).  Results are assigned (not accumulated), as
RetrieveForceCUDA does (src/force_gpu_cuda.cu:866-875).

Resident evaluates a whole soft-force call that is kept in HBM (the form
bench.py times and the serving loop uses).

The HIP library is required: every entry point raises if it is missing or if
no GPU is visible; there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
import os

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libpetar_hip.so")
_lib = None

PTH_OK, PTH_ERR_ARG, PTH_ERR_CAPACITY, PTH_ERR_HIP, PTH_ERR_STATE, PTH_ERR_NOMEM = 0, -1, -2, -3, -4, -5
EXEC_FUSED, EXEC_PLAIN = 0, 1

# record layouts of the FDPS particle types (no KDKDK_4TH / SAVE_NEIGHBOR_ID)
EPI_SOFT = np.dtype([("id", "<i8"), ("pos", "<f8", 3), ("r_search", "<f8"), ("rank_org", "<i4"),
                     ("type", "<i4")])
EPJ_SOFT = np.dtype([("id", "<i8"), ("mass", "<f8"), ("pos", "<f8", 3), ("vel", "<f8", 3), ("r_in", "<f8"),
                     ("r_out", "<f8"), ("r_search", "<f8"), ("r_scale_next", "<f8"), ("group_data", "<f8", 2),
                     ("rank_org", "<i4"), ("adr_org", "<i4")])
SPJ_SOFT = np.dtype([("mass", "<f8"), ("pos", "<f8", 3),
                     ("quad", [("xx", "<f8"), ("yy", "<f8"), ("zz", "<f8"), ("xy", "<f8"), ("xz", "<f8"),
                               ("yz", "<f8")])])
# PS::SPJMonopoleInAndOut, the superparticle of a build without USE_QUAD
SPJ_SOFT_MONO = np.dtype([("mass", "<f8"), ("pos", "<f8", 3)])
FORCE_SOFT = np.dtype([("acc", "<f8", 3), ("pot", "<f8"), ("n_ngb", "<i8")])
# ForceSoft with the total potential the changeover mode also returns (pot
# is then the changeover-weighted soft potential)
FORCE_SOFT_CO = np.dtype([("acc", "<f8", 3), ("pot", "<f8"), ("n_ngb", "<i8"), ("pot_tot", "<f8")])
assert EPI_SOFT.itemsize == 48 and EPJ_SOFT.itemsize == 120 and FORCE_SOFT.itemsize == 40


class Params(C.Structure):
    _fields_ = [("eps2", C.c_double), ("rcut2", C.c_double), ("G", C.c_double), ("nnb_mode", C.c_int),
                ("exec_mode", C.c_int), ("chunk_j", C.c_int), ("max_walk", C.c_int), ("max_ipl", C.c_int),
                ("j_vgpr", C.c_int), ("changeover", C.c_int), ("ngb_factor", C.c_double), ("segments", C.c_int),
                ("streams", C.c_int), ("df_near", C.c_int), ("spj_monopole", C.c_int), ("graph", C.c_int),
                ("min_seg_pairs", C.c_double), ("persistent", C.c_int)]


class Profile(C.Structure):
    """pth_profile: times (s) and counts of the FDPS hooks."""
    _fields_ = [("copy_s", C.c_double), ("send_s", C.c_double), ("calc_s", C.c_double), ("recv_s", C.c_double),
                ("n_walk", C.c_longlong), ("n_epi", C.c_longlong), ("n_epj", C.c_longlong), ("n_spj", C.c_longlong),
                ("n_call", C.c_longlong)]


class EpjView(C.Structure):
    _fields_ = [(n, C.c_size_t) for n in ("stride", "off_pos", "off_mass", "off_rsearch", "off_id", "off_rin",
                                          "off_rout")] + [("off_group", C.c_longlong)]


class SpjView(C.Structure):
    _fields_ = [(n, C.c_size_t) for n in ("stride", "off_pos", "off_mass", "off_qxx", "off_qyy", "off_qzz",
                                          "off_qxy", "off_qxz", "off_qyz")]


class EpiView(C.Structure):
    _fields_ = [(n, C.c_size_t) for n in ("stride", "off_pos", "off_rsearch", "off_id")]


class ForceView(C.Structure):
    _fields_ = [("stride", C.c_size_t), ("off_acc", C.c_size_t), ("off_pot", C.c_size_t),
                ("off_nngb", C.c_size_t), ("nngb_bytes", C.c_int), ("off_pot_tot", C.c_longlong)]


EXPORTED = ("pth_params_default", "pth_create", "pth_set_params", "pth_get_params", "pth_create_error",
            "pth_destroy", "pth_last_error", "pth_device", "pth_send_j", "pth_dispatch", "pth_retrieve",
            "pth_resident_create", "pth_resident_changeover", "pth_resident_run", "pth_resident_sync", "pth_resident_results",
            "pth_resident_timing", "pth_resident_stats", "pth_resident_free", "pth_plan_debug",
            "pth_resident_replan", "pth_resident_autotune", "pth_gtree_build", "pth_gtree_sizes",
            "pth_gtree_download", "pth_gtree_timing", "pth_gtree_stats", "pth_gtree_free", "pth_resident_from_gtree",
            "pth_gtree_update", "pth_resident_refill", "pth_gtree_neighbors", "pth_gtree_neighbors_download",
            "pth_server_create", "pth_server_add_cluster", "pth_server_footprint", "pth_server_submit",
            "pth_server_wait", "pth_server_replay", "pth_server_results", "pth_server_stats", "pth_server_free",
            "pth_profile_enable", "pth_profile_take", "pth_server_memory", "pth_resident_eval_noalloc", "pth_dispatch_copies")


def lib():
    """Loads libpetar_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise RuntimeError(f"{_LIB} is missing: the HIP extension must be built (__graft_entry__.build())")
    L = C.CDLL(_LIB)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    L.pth_params_default.argtypes = [C.POINTER(Params)]
    L.pth_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(Params)]
    L.pth_set_params.argtypes = [vp, C.POINTER(Params)]
    L.pth_get_params.argtypes = [vp, C.POINTER(Params)]
    L.pth_profile_enable.argtypes = [vp, C.c_int]
    L.pth_profile_take.argtypes = [vp, C.POINTER(Profile)]
    L.pth_create_error.restype = C.c_char_p
    L.pth_destroy.argtypes = [vp]
    L.pth_last_error.argtypes = [vp]
    L.pth_last_error.restype = C.c_char_p
    L.pth_device.argtypes = [vp]
    L.pth_send_j.argtypes = [vp, vp, C.c_int, C.POINTER(EpjView), vp, C.c_int, C.POINTER(SpjView)]
    L.pth_dispatch.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp), ip, C.POINTER(EpiView), C.POINTER(vp), ip,
                               C.POINTER(vp), ip]
    L.pth_retrieve.argtypes = [vp, C.c_int, C.c_int, ip, C.POINTER(vp), C.POINTER(ForceView)]
    L.pth_dispatch_copies.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp), ip, C.POINTER(EpiView), C.POINTER(vp), ip,
                                      C.POINTER(EpjView), C.POINTER(vp), ip, C.POINTER(SpjView)]
    L.pth_resident_create.argtypes = [vp, C.POINTER(vp), C.c_int, ip, ip, ip, dp, dp, ip, ip, C.c_int, dp, dp, dp,
                                      C.c_int, dp, dp, dp]
    L.pth_resident_run.argtypes = [vp]
    L.pth_resident_sync.argtypes = [vp]
    L.pth_resident_changeover.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp, dp, vp]
    L.pth_resident_results.argtypes = [vp, dp, dp, C.POINTER(C.c_longlong), dp]
    L.pth_resident_timing.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.pth_resident_stats.argtypes = [vp] + [C.POINTER(C.c_longlong)] * 4
    L.pth_resident_free.argtypes = [vp]
    L.pth_resident_eval_noalloc.argtypes = [vp, C.c_int, ip, ip, ip, dp, dp, ip, ip, C.c_int, dp, dp, dp, C.c_int,
                                            dp, dp, dp, dp, dp, C.POINTER(C.c_longlong)]
    L.pth_server_create.argtypes = [vp, C.POINTER(vp), C.c_int, C.c_size_t, C.c_int]
    L.pth_server_add_cluster.argtypes = [vp, ip, C.c_int, ip, ip, ip, dp, dp, ip, ip, C.c_int, dp, dp, dp, C.c_int,
                                         dp, dp, dp]
    L.pth_server_footprint.argtypes = [vp, C.c_int]
    L.pth_server_footprint.restype = C.c_longlong
    L.pth_server_submit.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong)]
    L.pth_server_wait.argtypes = [vp, C.c_longlong, dp]
    L.pth_server_replay.argtypes = [vp, C.c_int, ip, dp, dp, dp, ip]
    L.pth_server_memory.argtypes = [vp] + [C.POINTER(C.c_longlong)] * 5
    L.pth_server_results.argtypes = [vp, C.c_int, dp, dp, C.POINTER(C.c_longlong)]
    L.pth_server_stats.argtypes = [vp] + [C.POINTER(C.c_longlong)] * 5
    L.pth_server_free.argtypes = [vp]
    L.pth_resident_replan.argtypes = [vp, C.c_int, C.c_int]
    L.pth_resident_autotune.argtypes = [vp, ip, C.c_int, ip, C.c_int, C.c_int, ip, ip, C.POINTER(C.c_float)]
    llp = C.POINTER(C.c_longlong)
    L.pth_gtree_build.argtypes = [vp, C.POINTER(vp), C.c_longlong, dp, dp, dp, C.c_int, C.c_int, C.c_double, C.c_int]
    L.pth_gtree_sizes.argtypes = [vp, llp, llp, llp, llp]
    L.pth_gtree_download.argtypes = [vp, llp, llp, llp, ip, dp, dp, dp, llp, ip, ip, ip, ip, ip]
    L.pth_gtree_timing.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.pth_gtree_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pth_resident_from_gtree.argtypes = [vp, C.POINTER(vp), vp]
    L.pth_gtree_update.argtypes = [vp, C.c_longlong, vp, vp, vp, C.c_int]
    L.pth_resident_refill.argtypes = [vp, vp]
    L.pth_gtree_neighbors.argtypes = [vp, llp]
    L.pth_gtree_neighbors_download.argtypes = [vp, ip, llp, ip, C.POINTER(C.c_float)]
    L.pth_gtree_free.argtypes = [vp]
    L.pth_plan_debug.restype = C.c_longlong
    L.pth_plan_debug.argtypes = [C.c_int, ip, ip, ip, C.c_int, C.c_int, C.c_int, C.c_int, ip, C.c_longlong,
                                 C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), ip, C.c_longlong,
                                 C.POINTER(C.c_longlong), ip]
    _lib = L
    return L


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class PthError(RuntimeError):
    pass


def default_params(eps2=0.0, rcut2=0.0, G=1.0, nnb_mode=0, exec_mode=EXEC_FUSED, chunk_j=0, max_walk=1000,
                   max_ipl=None, j_vgpr=None, changeover=0, ngb_factor=None, segments=None, streams=None,
                   df_near=None, spj_monopole=0, graph=None, min_seg_pairs=None, persistent=None):
    p = Params()
    lib().pth_params_default(C.byref(p))
    p.eps2, p.rcut2, p.G, p.nnb_mode, p.exec_mode, p.chunk_j, p.max_walk = (
        eps2, rcut2, G, nnb_mode, exec_mode, chunk_j, max_walk)
    if max_ipl is not None:
        p.max_ipl = max_ipl
    if j_vgpr is not None:
        p.j_vgpr = j_vgpr
    p.changeover = changeover
    if ngb_factor is not None:
        p.ngb_factor = ngb_factor
    if segments is not None:
        p.segments = segments
    if streams is not None:
        p.streams = streams
    if df_near is not None:
        p.df_near = df_near
    p.spj_monopole = spj_monopole
    if graph is not None:
        p.graph = graph
    if min_seg_pairs is not None:
        p.min_seg_pairs = min_seg_pairs
    if persistent is not None:
        p.persistent = persistent
    return p


# live contexts, closed (with their trees, evaluations and servers) when the
# interpreter exits, while the HIP runtime is still whole
_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_contexts():
    for ctx in list(_CONTEXTS):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """One device context (pth_create); device = rank % visible devices."""

    def __init__(self, rank=0, **params):
        L = lib()
        self._h = C.c_void_p()
        self.params = default_params(**params)
        # resident evaluations, trees and servers made on this context: they
        # use its device and stream, so they are closed before it is
        self._children = weakref.WeakSet()
        rc = L.pth_create(C.byref(self._h), rank, C.byref(self.params))
        if rc != PTH_OK:
            raise PthError(f"pth_create failed ({rc}): {L.pth_create_error().decode()}")
        _CONTEXTS.add(self)

    def _adopt(self, obj):
        self._children.add(obj)

    @property
    def handle(self):
        return self._h

    @property
    def device(self) -> int:
        return lib().pth_device(self._h)

    def check(self, rc, what):
        if rc != PTH_OK:
            raise PthError(f"{what} failed ({rc}): {lib().pth_last_error(self._h).decode()}")

    def get_params(self) -> Params:
        p = Params()
        self.check(lib().pth_get_params(self._h, C.byref(p)), "pth_get_params")
        return p

    def profile_enable(self, on=True):
        """Times the hooks' copies and kernels with HIP events (pth_profile_enable)."""
        self.check(lib().pth_profile_enable(self._h, int(bool(on))), "pth_profile_enable")

    def profile_take(self) -> dict:
        """What the hooks accumulated since the last take (pth_profile_take)."""
        p = Profile()
        self.check(lib().pth_profile_take(self._h, C.byref(p)), "pth_profile_take")
        return {k: getattr(p, k) for k, _ in Profile._fields_}

    def set_params(self, p: Params):
        """Replaces every parameter (pth_set_params); batches in flight keep
        the values they were dispatched with."""
        self.check(lib().pth_set_params(self._h, C.byref(p)), "pth_set_params")
        self.params = p

    def close(self):
        if self._h:
            for child in list(self._children):
                child.close()
            lib().pth_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _views(mono=False):
    f = EPJ_SOFT.fields
    # group_data holds ArtificialParticleInformation: mass_backup, status
    ev = EpjView(EPJ_SOFT.itemsize, f["pos"][1], f["mass"][1], f["r_search"][1], f["id"][1], f["r_in"][1],
                 f["r_out"][1], f["group_data"][1])
    sf = SPJ_SOFT.fields
    qf = SPJ_SOFT.fields["quad"][0].fields
    qo = sf["quad"][1]
    sv = SpjView(SPJ_SOFT.itemsize, sf["pos"][1], sf["mass"][1], qo + qf["xx"][1], qo + qf["yy"][1],
                 qo + qf["zz"][1], qo + qf["xy"][1], qo + qf["xz"][1], qo + qf["yz"][1])
    if mono:
        mf = SPJ_SOFT_MONO.fields
        sv = SpjView(SPJ_SOFT_MONO.itemsize, mf["pos"][1], mf["mass"][1], 0, 0, 0, 0, 0, 0)
    fi = EPI_SOFT.fields
    iv = EpiView(EPI_SOFT.itemsize, fi["pos"][1], fi["r_search"][1], fi["id"][1])
    return ev, sv, iv, force_view(FORCE_SOFT)


def force_view(dt):
    """View of a force record dtype with acc, pot, n_ngb (int64) and,
    optionally, pot_tot (written in changeover mode)."""
    ff = dt.fields
    pt = ff["pot_tot"][1] if "pot_tot" in ff else -1
    return ForceView(dt.itemsize, ff["acc"][1], ff["pot"][1], ff["n_ngb"][1], 8, pt)


class CalcForceWithLinearCutoffMultiWalk:
    """The dispatch functor of src/force_gpu_cuda.hpp:101-133 over the C ABI.

    One process-wide context per rank, created at the first call (as the
    reference's first-call initialisation, src/force_gpu_cuda.cu:547-571)
    and updated with pth_set_params when a functor with other parameters
    calls it (as the C++ adapter does), so j arrays sent through one functor
    serve the dispatches of another.  Contexts are closed at interpreter
    exit.
    """

    _ctx = {}

    def __init__(self, my_rank=0, eps2=0.0, rcut2=0.0, G=1.0, nnb_mode=0, exec_mode=EXEC_FUSED, chunk_j=0,
                 changeover=0, spj_monopole=0):
        self.my_rank, self.eps2, self.rcut2, self.G = my_rank, eps2, rcut2, G
        self.nnb_mode, self.exec_mode, self.chunk_j, self.changeover = nnb_mode, exec_mode, chunk_j, changeover
        self.spj_monopole = spj_monopole

    @classmethod
    def close_contexts(cls):
        """Closes the per-rank contexts (with the j arrays they hold); the next
        call creates a fresh one."""
        for c in cls._ctx.values():
            c.close()
        cls._ctx.clear()

    def _wanted(self):
        return dict(eps2=float(self.eps2), rcut2=float(self.rcut2), G=float(self.G), nnb_mode=int(self.nnb_mode),
                    exec_mode=int(self.exec_mode), chunk_j=int(self.chunk_j), changeover=int(self.changeover),
                    spj_monopole=int(self.spj_monopole))

    def context(self) -> Context:
        want = self._wanted()
        c = self._ctx.get(self.my_rank)
        if c is None or not c.handle:
            c = Context(self.my_rank, **want)
            self._ctx[self.my_rank] = c
            return c
        cur = c.get_params()
        if any(getattr(cur, k) != v for k, v in want.items()):
            for k, v in want.items():
                setattr(cur, k, v)
            c.set_params(cur)
        return c

    def __call__(self, tag, n_walk, epi, n_epi, id_epj, n_epj, id_spj, n_spj, epj, n_epj_tot, spj, n_spj_tot,
                 send_flag):
        c = self.context()
        ev, sv, iv, _ = _views(bool(self.spj_monopole))
        L = lib()
        if send_flag:
            epj = np.ascontiguousarray(epj, dtype=EPJ_SOFT)
            spj = np.ascontiguousarray(spj, dtype=SPJ_SOFT_MONO if self.spj_monopole else SPJ_SOFT)
            c.check(L.pth_send_j(c.handle, epj.ctypes.data, int(n_epj_tot), C.byref(ev), spj.ctypes.data,
                                 int(n_spj_tot), C.byref(sv)), "pth_send_j")
            return 0
        epi = [np.ascontiguousarray(e, dtype=EPI_SOFT) for e in epi[:n_walk]]
        ide = [_i32(a) for a in id_epj[:n_walk]]
        ids = [_i32(a) for a in id_spj[:n_walk]]
        pe = (C.c_void_p * max(n_walk, 1))(*[e.ctypes.data for e in epi])
        pj = (C.c_void_p * max(n_walk, 1))(*[a.ctypes.data for a in ide])
        ps = (C.c_void_p * max(n_walk, 1))(*[a.ctypes.data for a in ids])
        ne, nj, ns = _i32(n_epi), _i32(n_epj), _i32(n_spj)
        c.check(L.pth_dispatch(c.handle, tag, n_walk, pe, _ptr(ne, C.c_int), C.byref(iv), pj, _ptr(nj, C.c_int),
                               ps, _ptr(ns, C.c_int)), "pth_dispatch")
        return 0

    def retrieve(self, tag, n_walk, ni, force):
        c = self.context()
        dt = force[0].dtype if n_walk > 0 else FORCE_SOFT
        if dt not in (FORCE_SOFT, FORCE_SOFT_CO):
            raise TypeError("force records must be FORCE_SOFT or FORCE_SOFT_CO arrays")
        fv = force_view(dt)
        for f in force[:n_walk]:
            if f.dtype != dt or not f.flags.c_contiguous:
                raise TypeError("force records must be contiguous arrays of one dtype")
        pf = (C.c_void_p * max(n_walk, 1))(*[f.ctypes.data for f in force[:n_walk]])
        n = _i32(ni)
        c.check(lib().pth_retrieve(c.handle, tag, n_walk, _ptr(n, C.c_int), pf, C.byref(fv)), "pth_retrieve")
        return 0


class CalcForceWithLinearCutoff(CalcForceWithLinearCutoffMultiWalk):
    """The dispatch functor of a build without the multi-walk index
    (src/force_gpu_cuda.hpp:137-162): each walk comes with its own EPJ and
    SPJ record arrays, f(tag, n_walk, epi, n_epi, epj, n_epj, spj, n_spj),
    over pth_dispatch_copies; retrieve as the multi-walk functor's.  Shares
    the per-rank contexts of CalcForceWithLinearCutoffMultiWalk."""

    def __call__(self, tag, n_walk, epi, n_epi, epj, n_epj, spj, n_spj):
        c = self.context()
        ev, sv, iv, _ = _views(bool(self.spj_monopole))
        epi = [np.ascontiguousarray(e, dtype=EPI_SOFT) for e in epi[:n_walk]]
        epj = [np.ascontiguousarray(e, dtype=EPJ_SOFT) for e in epj[:n_walk]]
        spj = [np.ascontiguousarray(e, dtype=SPJ_SOFT_MONO if self.spj_monopole else SPJ_SOFT) for e in spj[:n_walk]]
        pe = (C.c_void_p * max(n_walk, 1))(*[e.ctypes.data for e in epi])
        pj = (C.c_void_p * max(n_walk, 1))(*[e.ctypes.data for e in epj])
        ps = (C.c_void_p * max(n_walk, 1))(*[e.ctypes.data for e in spj])
        ne, nj, ns = _i32(n_epi), _i32(n_epj), _i32(n_spj)
        c.check(lib().pth_dispatch_copies(c.handle, tag, n_walk, pe, _ptr(ne, C.c_int), C.byref(iv), pj,
                                          _ptr(nj, C.c_int), C.byref(ev), ps, _ptr(ns, C.c_int), C.byref(sv)),
                "pth_dispatch_copies")
        return 0


_last_functor = [None]


def RetrieveForce(tag, n_walk, ni, force, functor=None):
    """RetrieveForceCUDA (src/force_gpu_cuda.cu:831-880): waits and assigns
    acc/pot/n_ngb of the batch dispatched by `functor` (default: the last one
    constructed through make_functor)."""
    f = functor or _last_functor[0]
    if f is None:
        raise PthError("RetrieveForce: no functor has dispatched")
    return f.retrieve(tag, n_walk, ni, force)


def make_functor(*args, **kw):
    f = CalcForceWithLinearCutoffMultiWalk(*args, **kw)
    _last_functor[0] = f
    return f


class Resident:
    """A whole soft-force evaluation kept in HBM (pth_resident_*).  With a
    context made with spj_monopole=1 the superparticles are monopoles and
    spj_quad is not read."""

    def __init__(self, ctx: Context, n_epi, n_epj, n_spj, epi_pos, epi_rs, id_epj, id_spj, epj_pos, epj_mass, epj_rs,
                 spj_pos, spj_mass, spj_quad, _tree=None):
        self.ctx = ctx
        L = lib()
        self.co = bool(ctx.params.changeover)
        self._h = C.c_void_p()
        ctx._adopt(self)
        if _tree is not None:
            self.n_walk = int(_tree.sizes()["n_group"])
            self.ni = _tree.n
            ctx.check(L.pth_resident_from_gtree(ctx.handle, C.byref(self._h), _tree._h), "pth_resident_from_gtree")
            return
        self._keep = [_i32(n_epi), _i32(n_epj), _i32(n_spj), _f64(epi_pos).reshape(-1, 3), _f64(epi_rs),
                      _i32(id_epj) if len(id_epj) else np.zeros(1, np.int32),
                      _i32(id_spj) if len(id_spj) else np.zeros(1, np.int32),
                      _f64(epj_pos).reshape(-1, 3) if len(epj_pos) else np.zeros((1, 3)),
                      _f64(epj_mass) if len(epj_mass) else np.zeros(1),
                      _f64(epj_rs) if len(epj_rs) else np.zeros(1),
                      _f64(spj_pos).reshape(-1, 3) if len(spj_pos) else np.zeros((1, 3)),
                      _f64(spj_mass) if len(spj_mass) else np.zeros(1),
                      _f64(spj_quad).reshape(-1, 6) if len(spj_quad) else np.zeros((1, 6))]
        k = self._keep
        self.n_walk = k[0].size
        self.ni = int(k[0].sum())
        ctx.check(L.pth_resident_create(ctx.handle, C.byref(self._h), self.n_walk, _ptr(k[0], C.c_int),
                                        _ptr(k[1], C.c_int), _ptr(k[2], C.c_int), _ptr(k[3], C.c_double),
                                        _ptr(k[4], C.c_double), _ptr(k[5], C.c_int), _ptr(k[6], C.c_int),
                                        len(epj_mass), _ptr(k[7], C.c_double), _ptr(k[8], C.c_double),
                                        _ptr(k[9], C.c_double), len(spj_mass), _ptr(k[10], C.c_double),
                                        _ptr(k[11], C.c_double), _ptr(k[12], C.c_double)), "pth_resident_create")
        self._keep = None

    def changeover(self, epi_id, epj_id, epj_rin, epj_rout, epj_group=None):
        """Changeover mode: ids of the EPI (concatenated per walk), ids,
        changeover radii and group data ([n, 2]: mass_backup, status; None:
        every particle a single) of the EPJ."""
        ei = np.ascontiguousarray(epi_id, dtype=np.int64)
        ej = np.ascontiguousarray(epj_id, dtype=np.int64)
        ri, ro = _f64(epj_rin), _f64(epj_rout)
        g = None if epj_group is None else _f64(epj_group).reshape(-1, 2)
        self.ctx.check(lib().pth_resident_changeover(self._h, _ptr(ei, C.c_longlong), _ptr(ej, C.c_longlong),
                                                     _ptr(ri, C.c_double), _ptr(ro, C.c_double),
                                                     None if g is None else g.ctypes.data),
                       "pth_resident_changeover")

    @classmethod
    def from_lists(cls, ctx, lists, rin=None, rout=None, group=None):
        """A resident evaluation of host.build_lists output; in changeover
        mode, rin / rout are the per-particle changeover radii and group the
        group data ([n, 2] or None) in the lists' EPJ order (the EPJ ids are
        the particle indices)."""
        idx = np.concatenate([np.arange(b, b + n) for b, n in zip(lists.epi_begin, lists.n_epi)])
        r = cls(ctx, lists.n_epi, lists.n_epj, lists.n_spj, lists.epj_pos[idx], lists.epj_rs[idx], lists.id_epj,
                lists.id_spj, lists.epj_pos, lists.epj_mass, lists.epj_rs, lists.spj_pos, lists.spj_mass,
                lists.spj_quad)
        if r.co:
            r.changeover(idx, np.arange(len(lists.epj_mass)), rin, rout, group)
        return r

    def refill(self, tree):
        """Refills this evaluation from a (re)built GpuTree, reusing device
        memory (pth_resident_refill)."""
        self.ctx.check(lib().pth_resident_refill(self._h, tree._h), "pth_resident_refill")
        self.n_walk = int(tree.sizes()["n_group"])
        self.ni = tree.n

    @classmethod
    def from_gtree(cls, ctx, tree):
        """A resident evaluation fed from a GpuTree on the device (walk w =
        group w, particles in tree order; pth_resident_from_gtree)."""
        return cls(ctx, *([None] * 13), _tree=tree)

    def run(self):
        self.ctx.check(lib().pth_resident_run(self._h), "pth_resident_run")

    def sync(self):
        self.ctx.check(lib().pth_resident_sync(self._h), "pth_resident_sync")

    def results(self):
        """acc [ni, 3], pot [ni], nnb [ni] (+ pot_tot [ni] in changeover mode)."""
        acc = np.zeros((self.ni, 3))
        pot = np.zeros(self.ni)
        nnb = np.zeros(self.ni, np.int64)
        pt = np.zeros(self.ni)
        self.ctx.check(lib().pth_resident_results(self._h, _ptr(acc, C.c_double), _ptr(pot, C.c_double),
                                                  _ptr(nnb, C.c_longlong), _ptr(pt, C.c_double)),
                       "pth_resident_results")
        return (acc, pot, nnb, pt) if self.co else (acc, pot, nnb)

    def replan(self, max_ipl, chunk_j):
        self.ctx.check(lib().pth_resident_replan(self._h, max_ipl, chunk_j), "pth_resident_replan")

    def autotune(self, ipl=(1, 2, 3, 4), chunk=(1024, 2048, 4096), reps=3):
        """Times every (max_ipl, chunk_j) pair on this evaluation, keeps the
        fastest; returns (max_ipl, chunk_j, ms)."""
        a, b = _i32(ipl), _i32(chunk)
        bi, bc, bt = C.c_int(), C.c_int(), C.c_float()
        self.ctx.check(lib().pth_resident_autotune(self._h, _ptr(a, C.c_int), a.size, _ptr(b, C.c_int), b.size, reps,
                                                   C.byref(bi), C.byref(bc), C.byref(bt)), "pth_resident_autotune")
        return bi.value, bc.value, bt.value

    def timing(self):
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        self.ctx.check(lib().pth_resident_timing(self._h, C.byref(a), C.byref(b), C.byref(c)), "pth_resident_timing")
        return a.value, b.value, c.value

    def stats(self):
        v = [C.c_longlong() for _ in range(4)]
        self.ctx.check(lib().pth_resident_stats(self._h, *[C.byref(x) for x in v]), "pth_resident_stats")
        return dict(n_epep=v[0].value, n_epsp=v[1].value, n_items=v[2].value, staged_bytes=v[3].value)

    def close(self):
        if self._h:
            lib().pth_resident_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def eval_noalloc(ctx: Context, n_epi, n_epj, n_spj, epi_pos, epi_rs, id_epj, id_spj, epj_pos, epj_mass, epj_rs,
                 spj_pos, spj_mass, spj_quad):
    """One evaluation with the arguments of Resident, run through
    pth_resident_eval_noalloc (no device allocation: every device buffer
    comes from the library's per-GPU scratch array); returns acc [ni, 3],
    pot [ni], nnb [ni]."""
    k = [_i32(n_epi), _i32(n_epj), _i32(n_spj), _f64(epi_pos).reshape(-1, 3), _f64(epi_rs),
         _i32(id_epj) if len(id_epj) else np.zeros(1, np.int32),
         _i32(id_spj) if len(id_spj) else np.zeros(1, np.int32),
         _f64(epj_pos).reshape(-1, 3) if len(epj_pos) else np.zeros((1, 3)),
         _f64(epj_mass) if len(epj_mass) else np.zeros(1), _f64(epj_rs) if len(epj_rs) else np.zeros(1),
         _f64(spj_pos).reshape(-1, 3) if len(spj_pos) else np.zeros((1, 3)),
         _f64(spj_mass) if len(spj_mass) else np.zeros(1),
         _f64(spj_quad).reshape(-1, 6) if len(spj_quad) else np.zeros((1, 6))]
    ni = int(k[0].sum())
    acc = np.zeros((ni, 3))
    pot = np.zeros(ni)
    nnb = np.zeros(ni, np.int64)
    ctx.check(lib().pth_resident_eval_noalloc(ctx.handle, k[0].size, _ptr(k[0], C.c_int), _ptr(k[1], C.c_int),
                                              _ptr(k[2], C.c_int), _ptr(k[3], C.c_double), _ptr(k[4], C.c_double),
                                              _ptr(k[5], C.c_int), _ptr(k[6], C.c_int), len(epj_mass),
                                              _ptr(k[7], C.c_double), _ptr(k[8], C.c_double), _ptr(k[9], C.c_double),
                                              len(spj_mass), _ptr(k[10], C.c_double), _ptr(k[11], C.c_double),
                                              _ptr(k[12], C.c_double), _ptr(acc, C.c_double), _ptr(pot, C.c_double),
                                              _ptr(nnb, C.c_longlong)), "pth_resident_eval_noalloc")
    return acc, pot, nnb


def eval_noalloc_lists(ctx: Context, lists):
    """eval_noalloc on host.build_lists output (the arguments of
    Resident.from_lists)."""
    idx = np.concatenate([np.arange(b, b + n) for b, n in zip(lists.epi_begin, lists.n_epi)])
    return eval_noalloc(ctx, lists.n_epi, lists.n_epj, lists.n_spj, lists.epj_pos[idx], lists.epj_rs[idx],
                        lists.id_epj, lists.id_spj, lists.epj_pos, lists.epj_mass, lists.epj_rs, lists.spj_pos,
                        lists.spj_mass, lists.spj_quad)


class GpuTree:
    """Octree, moments, i-groups and interaction lists built on the device
    (pth_gtree_*), with the semantics of petar_amd.host.build_lists."""

    def __init__(self, ctx: Context, pos, mass, r_search=None, theta=0.3, n_leaf_limit=20, n_group_limit=512,
                 canonical=True):
        self.ctx = ctx
        self._h = C.c_void_p()
        ctx._adopt(self)
        pos = _f64(pos).reshape(-1, 3)
        self.n = pos.shape[0]
        mass = _f64(mass)
        rs = None if r_search is None else _f64(np.broadcast_to(r_search, (self.n,)))
        ctx.check(lib().pth_gtree_build(ctx.handle, C.byref(self._h), self.n, _ptr(pos, C.c_double),
                                        _ptr(mass, C.c_double), None if rs is None else _ptr(rs, C.c_double),
                                        n_leaf_limit, n_group_limit, theta, int(canonical)), "pth_gtree_build")

    def update(self, pos, mass, r_search=None):
        """Rebuilds for new particles, reusing device memory.  numpy arrays
        are host input; torch CUDA tensors (float64, contiguous) are read in
        place on the device."""
        if hasattr(pos, "data_ptr") and getattr(pos, "is_cuda", False):
            n = int(pos.shape[0])
            for a in (pos, mass) + (() if r_search is None else (r_search,)):
                if a.dtype != __import__("torch").float64 or not a.is_contiguous() or not a.is_cuda:
                    raise ValueError("device inputs must be contiguous float64 CUDA tensors")
            ptrs = [pos.data_ptr(), mass.data_ptr(), None if r_search is None else r_search.data_ptr()]
            keep, on_dev = None, 1
        else:
            p = _f64(pos).reshape(-1, 3)
            n = p.shape[0]
            keep = [p, _f64(mass)] + ([] if r_search is None else [_f64(np.broadcast_to(r_search, (n,)))])
            ptrs = [keep[0].ctypes.data, keep[1].ctypes.data, None if r_search is None else keep[2].ctypes.data]
            on_dev = 0
        self.n = n
        self.ctx.check(lib().pth_gtree_update(self._h, n, ptrs[0], ptrs[1], ptrs[2], on_dev), "pth_gtree_update")
        del keep

    def sizes(self):
        v = [C.c_longlong() for _ in range(4)]
        self.ctx.check(lib().pth_gtree_sizes(self._h, *[C.byref(x) for x in v]), "pth_gtree_sizes")
        return dict(zip(("n_node", "n_group", "n_epj_list", "n_spj_list"), (x.value for x in v)))

    def download(self):
        z = self.sizes()
        nn, ng = z["n_node"], z["n_group"]
        d = dict(order=np.zeros(self.n, np.int64), node_begin=np.zeros(nn, np.int64), node_end=np.zeros(nn, np.int64),
                 node_level=np.zeros(nn, np.int32), node_mass=np.zeros(nn), node_com=np.zeros((nn, 3)),
                 node_quad=np.zeros((nn, 6)), epi_begin=np.zeros(ng, np.int64), n_epi=np.zeros(ng, np.int32),
                 n_epj=np.zeros(ng, np.int32), n_spj=np.zeros(ng, np.int32),
                 id_epj=np.zeros(max(z["n_epj_list"], 1), np.int32), id_spj=np.zeros(max(z["n_spj_list"], 1), np.int32))
        ll, i32, f = C.c_longlong, C.c_int, C.c_double
        self.ctx.check(lib().pth_gtree_download(
            self._h, _ptr(d["order"], ll), _ptr(d["node_begin"], ll), _ptr(d["node_end"], ll),
            _ptr(d["node_level"], i32), _ptr(d["node_mass"], f), _ptr(d["node_com"], f), _ptr(d["node_quad"], f),
            _ptr(d["epi_begin"], ll), _ptr(d["n_epi"], i32), _ptr(d["n_epj"], i32), _ptr(d["n_spj"], i32),
            _ptr(d["id_epj"], i32), _ptr(d["id_spj"], i32)), "pth_gtree_download")
        d["id_epj"] = d["id_epj"][:z["n_epj_list"]]
        d["id_spj"] = d["id_spj"][:z["n_spj_list"]]
        return d

    def neighbors(self):
        """Neighbour lists on the device (pth_gtree_neighbors): returns
        (count [n], offset [n + 1], ngb, ms) indexed by input particle index,
        self included."""
        tot = C.c_longlong()
        self.ctx.check(lib().pth_gtree_neighbors(self._h, C.byref(tot)), "pth_gtree_neighbors")
        cnt = np.zeros(self.n, np.int32)
        off = np.zeros(self.n + 1, np.int64)
        ngb = np.zeros(max(tot.value, 1), np.int32)
        ms = C.c_float()
        self.ctx.check(lib().pth_gtree_neighbors_download(self._h, _ptr(cnt, C.c_int), _ptr(off, C.c_longlong),
                                                          _ptr(ngb, C.c_int), C.byref(ms)),
                       "pth_gtree_neighbors_download")
        return cnt, off, ngb[:tot.value], ms.value

    def timing(self):
        """Host wall ms of the last build: (tree, first walk pass and list
        sizes, writing pass or list move + canonical ordering with the host
        plan made meanwhile)."""
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        self.ctx.check(lib().pth_gtree_timing(self._h, C.byref(a), C.byref(b), C.byref(c)), "pth_gtree_timing")
        return a.value, b.value, c.value

    def _stats(self):
        v = [C.c_int(), C.c_int(), C.c_int()]
        self.ctx.check(lib().pth_gtree_stats(self._h, *(C.byref(x) for x in v)), "pth_gtree_stats")
        return [x.value for x in v]

    def walk_passes(self):
        """Walk passes of the last build: 1 when the lists were written in one
        pass into regions sized from the previous build, 2 otherwise."""
        return self._stats()[0]

    def tile_walk_passes(self):
        """Passes of the last neighbour search's candidate walk (1 or 2, as
        walk_passes)."""
        return self._stats()[1]

    def filter_passes(self):
        """Passes of the last neighbour search's exact filter (1 or 2)."""
        return self._stats()[2]

    def close(self):
        if self._h:
            lib().pth_gtree_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SERVE_FUSED, SERVE_PLAIN = 0, 1


class Server:
    """Many independent evaluations behind one queue (pth_server_*).

    mode SERVE_FUSED keeps admitted clusters resident in an HBM pool of
    hbm_budget bytes and runs each request as one graph launch on one of
    n_streams streams; SERVE_PLAIN replays the plain per-walk form per
    request.  Clusters are host.Lists objects (or the arguments of Resident)."""

    def __init__(self, ctx: Context, mode=SERVE_FUSED, hbm_budget=0, n_streams=4):
        self.ctx = ctx
        self.mode = mode
        self._h = C.c_void_p()
        ctx._adopt(self)
        self._ni = []
        self._nint = []
        ctx.check(lib().pth_server_create(ctx.handle, C.byref(self._h), mode, int(hbm_budget), n_streams),
                  "pth_server_create")

    def add(self, n_epi, n_epj, n_spj, epi_pos, epi_rs, id_epj, id_spj, epj_pos, epj_mass, epj_rs, spj_pos,
            spj_mass, spj_quad):
        k = [_i32(n_epi), _i32(n_epj), _i32(n_spj), _f64(epi_pos).reshape(-1, 3), _f64(epi_rs),
             _i32(id_epj) if len(id_epj) else np.zeros(1, np.int32),
             _i32(id_spj) if len(id_spj) else np.zeros(1, np.int32),
             _f64(epj_pos).reshape(-1, 3) if len(epj_pos) else np.zeros((1, 3)),
             _f64(epj_mass) if len(epj_mass) else np.zeros(1), _f64(epj_rs) if len(epj_rs) else np.zeros(1),
             _f64(spj_pos).reshape(-1, 3) if len(spj_pos) else np.zeros((1, 3)),
             _f64(spj_mass) if len(spj_mass) else np.zeros(1),
             _f64(spj_quad).reshape(-1, 6) if len(spj_quad) else np.zeros((1, 6))]
        cid = C.c_int()
        self.ctx.check(lib().pth_server_add_cluster(
            self._h, C.byref(cid), k[0].size, _ptr(k[0], C.c_int), _ptr(k[1], C.c_int), _ptr(k[2], C.c_int),
            _ptr(k[3], C.c_double), _ptr(k[4], C.c_double), _ptr(k[5], C.c_int), _ptr(k[6], C.c_int), len(epj_mass),
            _ptr(k[7], C.c_double), _ptr(k[8], C.c_double), _ptr(k[9], C.c_double), len(spj_mass),
            _ptr(k[10], C.c_double), _ptr(k[11], C.c_double), _ptr(k[12], C.c_double)), "pth_server_add_cluster")
        self._ni.append(int(k[0].sum()))
        self._nint.append(int(np.dot(k[0].astype(np.int64), (k[1] + k[2]).astype(np.int64))))
        return cid.value

    def add_lists(self, lists):
        idx = np.concatenate([np.arange(b, b + n) for b, n in zip(lists.epi_begin, lists.n_epi)])
        return self.add(lists.n_epi, lists.n_epj, lists.n_spj, lists.epj_pos[idx], lists.epj_rs[idx],
                        lists.id_epj, lists.id_spj, lists.epj_pos, lists.epj_mass, lists.epj_rs, lists.spj_pos,
                        lists.spj_mass, lists.spj_quad)

    def interactions(self, cluster):
        return self._nint[cluster]

    def footprint(self, cluster):
        return int(lib().pth_server_footprint(self._h, cluster))

    def submit(self, cluster):
        t = C.c_longlong()
        self.ctx.check(lib().pth_server_submit(self._h, cluster, C.byref(t)), "pth_server_submit")
        return t.value

    def wait(self, ticket):
        lat = np.zeros(1)
        self.ctx.check(lib().pth_server_wait(self._h, ticket, _ptr(lat, C.c_double)), "pth_server_wait")
        return float(lat[0])

    def replay(self, clusters, arrival_s, admitted=False):
        """Submits clusters[k] at arrival_s[k] seconds; returns the latencies
        (ms) and the wall time (s) from the first arrival to the last
        completion (and, with admitted=True, whether each request uploaded
        its cluster)."""
        n = len(clusters)
        cl = _i32(clusters) if n else np.zeros(1, np.int32)
        ar = _f64(arrival_s) if n else np.zeros(1)
        lat = np.zeros(max(n, 1))
        wall = np.zeros(1)
        adm = np.zeros(max(n, 1), np.int32)
        self.ctx.check(lib().pth_server_replay(self._h, n, _ptr(cl, C.c_int), _ptr(ar, C.c_double),
                                               _ptr(lat, C.c_double), _ptr(wall, C.c_double), _ptr(adm, C.c_int)),
                       "pth_server_replay")
        if admitted:
            return lat[:n], float(wall[0]), adm[:n].astype(bool)
        return lat[:n], float(wall[0])

    def memory(self):
        """pth_server_memory: resident bytes of all clusters, the largest
        per-stream scratch, pinned host bytes, graph captures and
        instantiations."""
        v = [C.c_longlong() for _ in range(5)]
        self.ctx.check(lib().pth_server_memory(self._h, *[C.byref(x) for x in v]), "pth_server_memory")
        return dict(zip(("resident_total", "scratch_max", "pinned_host", "captures", "instantiations"),
                        (x.value for x in v)))

    def results(self, cluster):
        n = self._ni[cluster]
        acc, pot, nnb = np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.int64)
        self.ctx.check(lib().pth_server_results(self._h, cluster, _ptr(acc, C.c_double), _ptr(pot, C.c_double),
                                                _ptr(nnb, C.c_longlong)), "pth_server_results")
        return acc, pot, nnb

    def stats(self):
        v = [C.c_longlong() for _ in range(5)]
        self.ctx.check(lib().pth_server_stats(self._h, *[C.byref(x) for x in v]), "pth_server_stats")
        return dict(zip(("done", "admitted", "evicted", "interactions", "pool_peak"), (x.value for x in v)))

    def close(self):
        if self._h:
            lib().pth_server_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_debug(n_epi, n_epj, n_spj, exec_mode=EXEC_FUSED, chunk_j=0, max_ipl=4, n_seg=1, segments=False):
    """The work decomposition (host only): items [k, 8] (column 2 packs
    n_i | ipl << 16), es_off, ss_off, ritems [k, 8] (and the segments
    [g, 12] when segments=True)."""
    L = lib()
    ne, nj, ns = _i32(n_epi), _i32(n_epj), _i32(n_spj)
    nw = ne.size
    nr = C.c_longlong()
    n = L.pth_plan_debug(nw, _ptr(ne, C.c_int), _ptr(nj, C.c_int), _ptr(ns, C.c_int), exec_mode, chunk_j, max_ipl,
                         n_seg, None, 0, None, None, None, 0, C.byref(nr), None)
    if n < 0:
        raise PthError(f"pth_plan_debug failed ({n})")
    items = np.zeros((max(n, 1), 8), np.int32)
    rit = np.zeros((max(nr.value, 1), 8), np.int32)
    es = np.zeros(nw + 1, np.int64)
    ss = np.zeros(nw + 1, np.int64)
    sg = np.zeros((32, 12), np.int32)
    n2 = L.pth_plan_debug(nw, _ptr(ne, C.c_int), _ptr(nj, C.c_int), _ptr(ns, C.c_int), exec_mode, chunk_j, max_ipl,
                          n_seg, _ptr(items, C.c_int), items.shape[0], _ptr(es, C.c_longlong), _ptr(ss, C.c_longlong),
                          _ptr(rit, C.c_int), rit.shape[0], C.byref(nr), _ptr(sg, C.c_int))
    if n2 < 0:
        raise PthError(f"pth_plan_debug failed ({n2})")
    out = (items[:n], es, ss, rit[:nr.value])
    return out + (sg[sg[:, 0] >= 0],) if segments else out
