"""PeTar snapshot files (pth_snapshot_* of include/petar_host.h).

A snapshot is a FileHeader (src/io.hpp:210-282) and one FPSoft record per
particle, as PeTar writes them with writeParticleAscii / writeParticleBinary
(src/petar.hpp:1905-1908).  Columns carry the names of PeTar's own reader
(tools/analysis/data.py:390-399).  `soft_inputs` turns a snapshot into the
inputs of one soft-force evaluation (positions, masses, r_search, changeover
radii, ids), so snapshots can drive the device tree, the resident evaluation
and the serving loop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import host

BINARY, ASCII = 0, 1
EXTERNAL, SE = 1, 2
INT_COLUMNS = ("binary_state", "id", "mass_bk", "status", "n_nb")

_bound = False


def _lib():
    global _bound
    L = host.lib()
    if not _bound:
        vp, cp = C.c_void_p, C.c_char_p
        L.pth_snapshot_read.argtypes = [cp, C.c_int, C.c_int, C.POINTER(vp)]
        L.pth_snapshot_new.argtypes = [C.c_longlong, C.c_int, C.POINTER(vp)]
        L.pth_snapshot_header.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                          C.POINTER(C.c_double), vp, vp]
        L.pth_snapshot_set_header.argtypes = [vp, C.c_longlong, C.c_double, vp, vp]
        L.pth_snapshot_column_width.argtypes = [vp, cp]
        L.pth_snapshot_column.argtypes = [vp, cp, vp]
        L.pth_snapshot_set_column.argtypes = [vp, cp, vp]
        L.pth_snapshot_write.argtypes = [vp, cp, C.c_int]
        L.pth_snapshot_free.argtypes = [vp]
        _bound = True
    return L


def _fmt(fmt):
    if fmt in (BINARY, "binary", "b", "B"):
        return BINARY
    if fmt in (ASCII, "ascii", "a", "A"):
        return ASCII
    raise ValueError(f"unknown snapshot format {fmt!r}")


def _flags(external, stellar_evolution):
    return (EXTERNAL if external else 0) | (SE if stellar_evolution else 0)


class Snapshot:
    """One snapshot held by the host library; columns by name."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def read(cls, path, fmt="ascii", external=False, stellar_evolution=False):
        h = C.c_void_p()
        host._check(_lib().pth_snapshot_read(str(path).encode(), _fmt(fmt), _flags(external, stellar_evolution),
                                             C.byref(h)), "pth_snapshot_read")
        return cls(h)

    @classmethod
    def new(cls, n, external=False, stellar_evolution=False, file_id=0, time=0.0, pos_offset=None, vel_offset=None):
        h = C.c_void_p()
        host._check(_lib().pth_snapshot_new(n, _flags(external, stellar_evolution), C.byref(h)), "pth_snapshot_new")
        s = cls(h)
        s.set_header(file_id, time, pos_offset, vel_offset)
        return s

    def header(self):
        fid, n, t = C.c_longlong(), C.c_longlong(), C.c_double()
        po, vo = np.zeros(3), np.zeros(3)
        host._check(_lib().pth_snapshot_header(self._h, C.byref(fid), C.byref(n), C.byref(t), po.ctypes.data,
                                               vo.ctypes.data), "pth_snapshot_header")
        return dict(file_id=fid.value, n=n.value, time=t.value, pos_offset=po, vel_offset=vo)

    @property
    def n(self):
        return self.header()["n"]

    def set_header(self, file_id=0, time=0.0, pos_offset=None, vel_offset=None):
        po = None if pos_offset is None else np.ascontiguousarray(pos_offset, dtype=np.float64)
        vo = None if vel_offset is None else np.ascontiguousarray(vel_offset, dtype=np.float64)
        host._check(_lib().pth_snapshot_set_header(self._h, file_id, time, None if po is None else po.ctypes.data,
                                                   None if vo is None else vo.ctypes.data), "pth_snapshot_set_header")

    def width(self, name):
        return _lib().pth_snapshot_column_width(self._h, name.encode())

    def __getitem__(self, name):
        w = self.width(name)
        if w == 0:
            raise KeyError(name)
        dt = np.int64 if name in INT_COLUMNS else np.float64
        out = np.zeros((self.n, w) if w > 1 else self.n, dt)
        host._check(_lib().pth_snapshot_column(self._h, name.encode(), out.ctypes.data), "pth_snapshot_column")
        return out

    def __setitem__(self, name, value):
        w = self.width(name)
        if w == 0:
            raise KeyError(name)
        dt = np.int64 if name in INT_COLUMNS else np.float64
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=dt), (self.n, w) if w > 1 else (self.n,)))
        host._check(_lib().pth_snapshot_set_column(self._h, name.encode(), a.ctypes.data), "pth_snapshot_set_column")

    def write(self, path, fmt="ascii"):
        host._check(_lib().pth_snapshot_write(self._h, str(path).encode(), _fmt(fmt)), "pth_snapshot_write")

    def close(self):
        if self._h:
            _lib().pth_snapshot_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def soft_inputs(snap: Snapshot):
    """The inputs of one soft-force evaluation from a snapshot: dict of pos
    [n, 3], mass, r_search, r_in, r_out (FP64) and id (int64), as PeTar's
    tree force reads them from the particle system (EPISoft / EPJSoft
    getters, src/soft_ptcl.hpp:271-376)."""
    return dict(pos=snap["pos"], mass=snap["mass"], r_search=snap["r_search"], r_in=snap["r_in"],
                r_out=snap["r_out"], id=snap["id"])
