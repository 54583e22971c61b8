"""Workloads and comparison helpers shared by the tests."""
from __future__ import annotations

import numpy as np

from petar_amd import host
from oracle import oracle_py as O

R_OUT = 0.01
R_SEARCH = 1.5 * R_OUT


def c1_lists(n=16384, theta=0.3, seed=1, f_bin=0.0, n_group_limit=512, r_out=R_OUT, model="plummer"):
    """The small plain configuration: Plummer, Kroupa masses, N-body units."""
    m, x, _ = host.cluster(n, model, kroupa=True, f_bin=f_bin, seed=seed)
    return host.build_lists(x, m, 1.5 * r_out, theta=theta, n_group_limit=n_group_limit)


def walk_centres(args):
    """Bounding-box centre of each walk's i-particles (what the GPU path uses)."""
    ne = np.asarray(args["n_epi"])
    pos = np.asarray(args["epi_pos"]).reshape(-1, 3)
    off = np.concatenate([[0], np.cumsum(ne)])
    c = np.zeros((ne.size, 3))
    for w in range(ne.size):
        if ne[w]:
            p = pos[off[w]:off[w + 1]]
            c[w] = 0.5 * (p.min(0) + p.max(0))
    return c


def rel_err(a, b):
    """Per-particle relative errors |a-b|/|b| for acc (rows) or pot."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 2:
        return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def nnb_check(nnb, nnb_ref, n_amb):
    """Neighbour counts must match wherever no pair sits in the FP32 band,
    and differ by at most the number of ambiguous pairs elsewhere."""
    d = np.abs(np.asarray(nnb) - np.asarray(nnb_ref))
    exact = n_amb == 0
    return int((d[exact] != 0).sum()), int((d > n_amb).sum()), int((~exact).sum())


def co_lists(n=16384, seed=1, f_bin=0.0, r_out=R_OUT, r_in=0.1 * R_OUT, theta=0.3, shift=(0.0, 0.0, 0.0)):
    """A changeover-mode workload: changeover radii scaled by
    max((m/<m>)^(1/3), 1) as ChangeOver::setR does (src/changeover.hpp:44-59),
    r_search = 1.5 r_out per particle, the cluster moved by shift.  Returns
    (lists, r_in, r_out) with the radii in the lists' EPJ (tree) order."""
    m, x, _ = host.cluster(n, "plummer", kroupa=True, f_bin=f_bin, seed=seed)
    x = x + np.asarray(shift, np.float64)
    mbar = m.mean()
    L = host.build_lists(x, m, 1.5 * r_out * np.maximum(np.cbrt(m / mbar), 1.0), theta=theta)
    fac = np.maximum(np.cbrt(L.epj_mass / mbar), 1.0)
    return L, r_in * fac, r_out * fac


def co_args(L, rin, rout):
    """Oracle arguments of a changeover-mode evaluation of host lists (the
    EPJ ids are the tree indices)."""
    a = O.lists_args(L)
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(L.epi_begin, L.n_epi)])
    a.update(epi_id=idx, epj_id=np.arange(len(L.epj_mass)), epj_rin=rin, epj_rout=rout)
    return a
