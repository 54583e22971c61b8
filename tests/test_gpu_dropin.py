"""The C++ drop-in (include/petar_hip_fdps.hpp) driven in FDPS's multi-walk
call order by tests/cpp/fdps_replay, checked against the FP64 oracle on the
lists the program used (tolerances as in test_gpu_parity)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py as O
from tests._common import nnb_check, rel_err, walk_centres

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "fdps_replay")


def _read(path):
    with open(path, "rb") as f:
        n, nn, ng, nle, nls = np.fromfile(f, np.int64, 5)
        d = dict(epj_pos=np.fromfile(f, np.float64, 3 * n).reshape(-1, 3), epj_mass=np.fromfile(f, np.float64, n),
                 epj_rs=np.fromfile(f, np.float64, n), spj_pos=np.fromfile(f, np.float64, 3 * nn).reshape(-1, 3),
                 spj_mass=np.fromfile(f, np.float64, nn), spj_quad=np.fromfile(f, np.float64, 6 * nn).reshape(-1, 6),
                 epi_begin=np.fromfile(f, np.int64, ng), n_epi=np.fromfile(f, np.int32, ng),
                 n_epj=np.fromfile(f, np.int32, ng), n_spj=np.fromfile(f, np.int32, ng),
                 id_epj=np.fromfile(f, np.int32, nle), id_spj=np.fromfile(f, np.int32, nls))
        ni = int(d["n_epi"].sum())
        res = (np.fromfile(f, np.float64, 3 * ni).reshape(-1, 3), np.fromfile(f, np.float64, ni),
               np.fromfile(f, np.int64, ni))
        d.update(epj_id=np.fromfile(f, np.int64, n), epj_rin=np.fromfile(f, np.float64, n),
                 epj_rout=np.fromfile(f, np.float64, n))
        pot_tot = np.fromfile(f, np.float64, ni)
        d["epj_group"] = np.fromfile(f, np.float64, 2 * n).reshape(-1, 2)
    return d, res + (pot_tot,)


@pytest.mark.parametrize("exe,n,walk_limit", [("fdps_replay", 20000, 200), ("fdps_replay", 3000, 7),
                                              ("fdps_replay_mono", 20000, 200), ("fdps_replay_chunk", 20000, 200),
                                              ("fdps_replay_copies", 20000, 200), ("fdps_replay_copies", 3000, 7)])
def test_cpp_dropin_fdps_replay(tmp_path, exe, n, walk_limit):
    """The adapter as a default PeTar GPU build compiles it (USE_QUAD), as a
    build configured with --disable-quad compiles it (monopole
    superparticles: the oracle's monopole pair, src/soft_force.hpp:125-158),
    with -DPETAR_HIP_CHUNK_J=256 (work items of at most 256 list entries:
    long lists split over several items and summed by k_reduce), and as a
    build without PARTICLE_SIMULATOR_GPU_MULIT_WALK_INDEX, whose functor
    (CalcForceWithLinearCutoffCUDA, src/force_gpu_cuda.hpp:137-162) receives
    per-walk copies of the j records (fdps_replay_copies)."""
    path = os.path.join(ROOT, "tests", "cpp", exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (build() compiles it)")
    mono = exe.endswith("_mono")
    out = tmp_path / "replay.bin"
    r_out = 0.01
    print(f"[dropin] {exe} N={n} walk_limit={walk_limit}", flush=True)
    p = subprocess.run([path, str(n), "11", "0.3", str(walk_limit), str(r_out), str(out)], capture_output=True,
                       text=True, timeout=240)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    d, (acc, pot, nnb, _) = _read(out)
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(d["epi_begin"], d["n_epi"])])
    a = dict(n_epi=d["n_epi"], n_epj=d["n_epj"], n_spj=d["n_spj"], epi_pos=d["epj_pos"][idx],
             epi_rs=d["epj_rs"][idx], id_epj=d["id_epj"], id_spj=d["id_spj"], epj_pos=d["epj_pos"],
             epj_mass=d["epj_mass"], epj_rs=d["epj_rs"], spj_pos=d["spj_pos"], spj_mass=d["spj_mass"],
             spj_quad=d["spj_quad"])
    racc, rpot, rnnb = O.soft_force_walks(**a, eps2=0.0, rout2=r_out ** 2, mono=mono)
    ea, ep = rel_err(acc, racc), rel_err(pot, rpot)
    print(f"[dropin] acc med {np.median(ea):.2e} p99 {np.percentile(ea, 99):.2e} max {ea.max():.2e} "
          f"pot med {np.median(ep):.2e}", flush=True)
    assert np.median(ea) <= 1e-6 and np.median(ep) <= 1e-6
    assert np.all(pot != 99.0)
    # The workload has 30% binaries: in spatially extended (halo) walks the
    # FP32 walk-relative coordinates carry ~1e-7 x extent absolute error,
    # which close pairs in the linear-cutoff regime turn into relative errors
    # up to ~1e-2 of small net accelerations.  The tail must be no worse than
    # the reference kernel's (absolute FP32 coordinates) on the same lists.
    if O.ref_available():
        pacc, _, _ = O.ref_soft_force_walks(**a, eps2=0.0, rout2=r_out ** 2, mono=mono)
        er = rel_err(pacc, racc)
        print(f"[dropin] reference kernel: acc med {np.median(er):.2e} p99 {np.percentile(er, 99):.2e} "
              f"max {er.max():.2e}", flush=True)
        assert np.percentile(ea, 99) <= np.percentile(er, 99) and ea.max() <= er.max()
    else:
        assert np.percentile(ea, 99) <= 1e-4 and ea.max() <= 2e-2
    amb = O.nnb_ambiguous(**a, centres=walk_centres(a), eps2=0.0)
    bad_exact, bad_band, _ = nnb_check(nnb, rnnb, amb)
    assert bad_exact == 0 and bad_band == 0


@pytest.mark.parametrize("exe", ["fdps_replay_co", "fdps_replay_copies_co"])
def test_cpp_dropin_changeover_mode(tmp_path, exe):
    """The adapter compiled with PETAR_HIP_CHANGEOVER (ForceSoft with
    pot_tot, mass-scaled changeover radii, EPJSoft group data with group
    members, members of groups without artificial particles and artificial
    particles) against the FP64 changeover oracle on the lists and group
    data the program used; with the multi-walk-index functor and with the
    functor that receives per-walk copies of the j records."""
    exe = os.path.join(ROOT, "tests", "cpp", exe)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (build() compiles it)")
    out = tmp_path / "replay_co.bin"
    r_out = 0.01
    p = subprocess.run([exe, "20000", "13", "0.3", "200", str(r_out), str(out)], capture_output=True, text=True,
                       timeout=240)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    d, (acc, pot, nnb, pot_tot) = _read(out)
    idx = np.concatenate([np.arange(b, b + k) for b, k in zip(d["epi_begin"], d["n_epi"])])
    a = dict(n_epi=d["n_epi"], n_epj=d["n_epj"], n_spj=d["n_spj"], epi_pos=d["epj_pos"][idx],
             epi_rs=d["epj_rs"][idx], epi_id=d["epj_id"][idx], id_epj=d["id_epj"], id_spj=d["id_spj"],
             epj_pos=d["epj_pos"], epj_mass=d["epj_mass"], epj_rs=d["epj_rs"], epj_id=d["epj_id"],
             epj_rin=d["epj_rin"], epj_rout=d["epj_rout"], spj_pos=d["spj_pos"], spj_mass=d["spj_mass"],
             spj_quad=d["spj_quad"])
    grp = d["epj_group"]
    assert (grp[:, 1] < 0).sum() > 0 and (grp[:, 1] > 0).sum() > 0
    racc, rps, rpt, rnnb, amb = O.soft_force_walks_co(**a, eps2=0.0, rout2=r_out ** 2, amb_rel=1e-4, epj_group=grp)
    ea, es, et = rel_err(acc, racc), rel_err(pot[amb == 0], rps[amb == 0]), rel_err(pot_tot, rpt)
    print(f"[dropin co] acc med {np.median(ea):.2e} p99 {np.percentile(ea, 99):.2e}; pot_soft med "
          f"{np.median(es):.2e}; pot_tot med {np.median(et):.2e} p99 {np.percentile(et, 99):.2e}; "
          f"i with neighbours {(rps != rpt).sum()}", flush=True)
    assert np.all(pot != 99.0) and np.all(pot_tot != 99.0)
    assert np.median(ea) <= 1e-6 and np.median(es) <= 1e-6 and np.median(et) <= 1e-6
    assert np.percentile(et, 99) <= 1e-5 and np.percentile(es, 99) <= 1e-5


def test_cpp_dropin_gpu_profile(tmp_path):
    """The adapter built with GPU_PROFILE (as every reference GPU build is,
    and under C++11): gpu_counter holds the replayed sums (walks, i, EPJ and
    SPJ list entries, dispatch calls) over all passes, gpu_profile's four
    times are positive, the dump lines carry the nine columns PeTar's
    profile reader expects, and clear() zeroes both."""
    path = os.path.join(ROOT, "tests", "cpp", "fdps_replay_prof")
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (build() compiles it)")
    out = tmp_path / "replay_prof.bin"
    n, walk_limit, passes = 20000, 200, 3
    p = subprocess.run([path, str(n), "11", "0.3", str(walk_limit), "0.01", str(out), "0.3", str(passes - 1)],
                       capture_output=True, text=True, timeout=240)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    d, _ = _read(out)
    lines = p.stdout.splitlines()
    cnt = [l for l in lines if "gpu_counter" in l][0].split()
    got = {cnt[k]: int(cnt[k + 1]) for k in range(2, len(cnt) - 1, 2)}
    n_group = d["n_epi"].size
    want = dict(n_walk=passes * n_group, n_epi=passes * int(d["n_epi"].sum()),
                n_epj=passes * int(d["n_epj"].sum()), n_spj=passes * int(d["n_spj"].sum()),
                n_call=passes * ((n_group + walk_limit - 1) // walk_limit))
    assert got == want
    prof = [l for l in lines if "gpu_profile" in l][0].split()
    times = {prof[k]: float(prof[k + 1]) for k in range(2, len(prof) - 1, 2)}
    assert set(times) == {"copy", "send", "receive", "calc_force"} and all(v > 0 for v in times.values())
    names = lines.index([l for l in lines if l.strip().startswith("copy")][0])
    vals = lines[names + 1].split()
    assert len(vals) == 9
    np.testing.assert_allclose([float(v) for v in vals[4:]], [want[k] / passes for k in
                                                              ("n_walk", "n_epi", "n_epj", "n_spj", "n_call")])


@pytest.mark.parametrize("exe", ["fdps_replay", "fdps_replay_co"])
def test_cpp_dropin_list_runs_bit_identical(tmp_path, exe):
    """pth_dispatch sends each batch's index lists either as they are or as
    runs of consecutive indices expanded on the device (the default when
    that at least halves their bytes; PTH_LIST_RUNS=0 / 1 forces one form),
    runs the 1-i items on a second stream beside the longer classes
    (PTH_CLASS_STREAMS=0: one stream), and has the kernels write the
    results straight into pinned host memory (PTH_ZERO_COPY=0: device
    buffers and a copy): every combination gives the same forces bit for
    bit, in the linear-cutoff and the changeover adapter, with batches of
    200 walks and of 7."""
    path = os.path.join(ROOT, "tests", "cpp", exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (build() compiles it)")
    for walk_limit in (200, 7):
        outs = []
        for mode, streams, zc in (("0", "1", "1"), ("1", "1", "1"), ("1", "0", "0")):
            out = tmp_path / f"replay_{mode}_{streams}_{zc}_{walk_limit}.bin"
            env = dict(os.environ, PTH_LIST_RUNS=mode, PTH_CLASS_STREAMS=streams, PTH_ZERO_COPY=zc)
            print(f"[dropin runs] {exe} PTH_LIST_RUNS={mode} PTH_CLASS_STREAMS={streams} PTH_ZERO_COPY={zc} "
                  f"walk_limit={walk_limit}", flush=True)
            p = subprocess.run([path, "12000", "17", "0.3", str(walk_limit), "0.01", str(out)], capture_output=True,
                               text=True, timeout=240, env=env)
            print(p.stdout, p.stderr, flush=True)
            assert p.returncode == 0
            outs.append(_read(out)[1])
        for k, name in enumerate(("acc", "pot", "nnb", "pot_tot")):
            for o in outs[1:]:
                assert np.array_equal(outs[0][k], o[k]), f"{exe} walk_limit={walk_limit}: {name} differs"


def test_cpp_dropin_long_run_lists_bit_identical(tmp_path):
    """At 400,000 particles some walks' SPJ lists hold more than 1,024 runs
    of consecutive indices, so their expansion on the device takes several
    chunks of runs through LDS: lists sent as runs (PTH_LIST_RUNS=1) give
    the same forces bit for bit as lists sent as they are."""
    path = os.path.join(ROOT, "tests", "cpp", "fdps_replay")
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (build() compiles it)")
    outs = []
    for mode in ("0", "1"):
        out = tmp_path / f"replay_long_{mode}.bin"
        env = dict(os.environ, PTH_LIST_RUNS=mode)
        print(f"[dropin long runs] PTH_LIST_RUNS={mode}", flush=True)
        p = subprocess.run([path, "400000", "11", "0.3", "200", "0.01", str(out)], capture_output=True, text=True,
                           timeout=240, env=env)
        print(p.stdout, p.stderr, flush=True)
        assert p.returncode == 0
        d, res = _read(out)
        outs.append(res)
    eoff = np.concatenate([[0], np.cumsum(d["n_spj"])])
    runs = [1 + int(np.count_nonzero(np.diff(d["id_spj"][eoff[w]:eoff[w + 1]]) != 1))
            for w in range(d["n_spj"].size) if d["n_spj"][w] > 0]
    assert max(runs) > 1024
    for k, name in enumerate(("acc", "pot", "nnb")):
        assert np.array_equal(outs[0][k], outs[1][k]), f"{name} differs"
