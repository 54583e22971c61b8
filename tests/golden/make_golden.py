"""Writes tests/golden/simd_test_scenario.npz.

The reference's only soft-force test, src/simd_test.cxx, compares the scalar,
SIMD and GPU kernels on one scenario but stores no vectors.  This script
rebuilds that scenario and records what the reference x86 kernel (compiled
from /root/reference by oracle/Makefile into oracle/_ref) and the FP64
oracle return for it:

  - particles: makePlummerModel(1.0, 2000, 2000, ..., -0.25) with the MT19937
    stream seeded by rank 0 (src/simd_test.cxx:71-85), changeover
    setR(1.0, 0.001, 0.01) and calcRSearch(1/2048) with search_factor 0, so
    r_search = r_out = 0.01 (src/ptcl.hpp:227-238, src/static_variables.hpp:3-4);
  - EPI = the first 1000 particles, EPJ = all 2000 (src/simd_test.cxx:62-64,
    111-125);
  - 1000 SPJ drawn with glibc rand() in the order of setSpj
    (src/simd_test.cxx:44-57; no srand call, so the default seed);
  - eps = 1e-4, r_out = 0.01, G = 1 (src/simd_test.cxx:66-68).

Outputs: the EP-EP and EP-SP results of the reference kernel (AVX-512 and
AVX2 builds, wrapper semantics of src/soft_force.hpp) and of the oracle in
the three neighbour conventions.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import ctypes
import os
import platform
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle_py as O  # noqa: E402
from petar_amd import host  # noqa: E402


def spj_from_glibc_rand(n_spj, N):
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(1)  # glibc's default state before any srand call
    RAND_MAX_F = float(np.float32(2147483647))  # (float)RAND_MAX = 2^31
    mass = np.zeros(n_spj)
    pos = np.zeros((n_spj, 3))
    quad = np.zeros((n_spj, 6))  # xx yy zz xy xz yz
    for i in range(n_spj):
        mass[i] = 1.0 / N + 0.001 / N * libc.rand() / RAND_MAX_F
        for d in range(3):
            pos[i, d] = 1.0 + 10.0 * libc.rand() / RAND_MAX_F
        # drawn in the order xx, yy, zz, xy, yz, xz
        xx = 10.0 * libc.rand() / RAND_MAX_F
        yy = 10.0 * libc.rand() / RAND_MAX_F
        zz = 10.0 * libc.rand() / RAND_MAX_F
        xy = 10.0 * libc.rand() / RAND_MAX_F
        yz = 10.0 * libc.rand() / RAND_MAX_F
        xz = 10.0 * libc.rand() / RAND_MAX_F
        quad[i] = (xx, yy, zz, xy, xz, yz)
    return mass, pos, quad


def main():
    n_epi, n_epj, n_spj = 1000, 2000, 1000
    N = max(n_epi, n_epj)
    mass, pos, _ = host.plummer_equal(N, 1.0, -0.25, seed=0, with_vel=False)
    rs = np.full(N, 0.01)
    eps2, rout2, G = 1e-8, 1e-4, 1.0
    smass, spos, squad = spj_from_glibc_rand(n_spj, float(N))
    walk = dict(n_epi=np.array([n_epi], np.int32), n_epj=np.array([n_epj], np.int32),
                n_spj=np.array([n_spj], np.int32), epi_pos=pos[:n_epi], epi_rs=rs[:n_epi],
                id_epj=np.arange(n_epj, dtype=np.int32), id_spj=np.arange(n_spj, dtype=np.int32),
                epj_pos=pos, epj_mass=mass, epj_rs=rs, spj_pos=spos, spj_mass=smass, spj_quad=squad)
    out = dict(epj_pos=pos, epj_mass=mass, epj_rs=rs, spj_pos=spos, spj_mass=smass, spj_quad=squad,
               n_epi=n_epi, eps2=eps2, rout2=rout2, G=G)
    for mode in (0, 1, 2):
        a, p, nb = O.soft_force_walks(**walk, eps2=eps2, rout2=rout2, G=G, nnb_mode=mode, part="epj")
        out[f"oracle_epj_mode{mode}_acc"], out[f"oracle_epj_mode{mode}_pot"], out[f"oracle_epj_mode{mode}_nnb"] = a, p, nb
    a, p, _ = O.soft_force_walks(**walk, eps2=eps2, rout2=rout2, G=G, part="spj")
    out["oracle_spj_acc"], out["oracle_spj_pot"] = a, p
    for isa in ("avx512", "avx2"):
        if not O.ref_available(isa):
            raise SystemExit(f"reference kernel {isa} not built: run `make -C oracle ref` first")
        if isa == "avx512" and not O.cpu_has_avx512():
            raise SystemExit("this CPU has no AVX-512; the fixture needs both ISAs")
        a, p, nb = O.ref_soft_force_walks(**walk, eps2=eps2, rout2=rout2, G=G, part="epj", isa=isa, nthreads=1)
        out[f"ref_{isa}_epj_acc"], out[f"ref_{isa}_epj_pot"], out[f"ref_{isa}_epj_nnb"] = a, p, nb
        a, p, _ = O.ref_soft_force_walks(**walk, eps2=eps2, rout2=rout2, G=G, part="spj", isa=isa, nthreads=1)
        out[f"ref_{isa}_spj_acc"], out[f"ref_{isa}_spj_pot"] = a, p
    out["cpu_model"] = np.array(platform.processor() or "unknown")
    path = os.path.join(ROOT, "tests", "golden", "simd_test_scenario.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
