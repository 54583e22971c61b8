"""PeTar snapshot files: the layout is pinned by PeTar's own reader
(tools/analysis/data.py: header dtype at :103,110, record keys at
:390-399) restated as numpy structured dtypes, and by the fprintf formats of
the writeAscii members (src/io.hpp:246-262, src/particle_base.hpp,
src/ptcl.hpp:147-152, src/changeover.hpp:160-163,
src/soft_ptcl.hpp:103-112), reproduced here as Python format strings; files
written either way are read back exactly."""
import numpy as np
import pytest

from petar_amd import host
from petar_amd import snapshot as S


def _record_dtype(external=False, se=False):
    """The record of tools/analysis/data.py:390-399 (default particle type,
    interrupt mode none / base)."""
    keys = [("mass", "f8"), ("pos", "f8", 3), ("vel", "f8", 3), ("binary_state", "i8")]
    if se:
        keys += [("radius", "f8"), ("dm", "f8"), ("time_record", "f8"), ("time_interrupt", "f8")]
    keys += [("r_search", "f8"), ("id", "i8"), ("mass_bk", "i8"), ("status", "i8"), ("r_in", "f8"), ("r_out", "f8"),
             ("acc_soft", "f8", 3), ("pot", "f8"), ("pot_soft", "f8")]
    if external:
        keys += [("pot_ext", "f8")]
    keys += [("n_nb", "i8")]
    return np.dtype(keys)


def _header_dtype(external=False):
    k = [("file_id", "i8"), ("n_glb", "i8"), ("time", "f8")]
    if external:
        k += [("x", "f8"), ("y", "f8"), ("z", "f8"), ("vx", "f8"), ("vy", "f8"), ("vz", "f8")]
    return np.dtype(k)


def _records(n, seed, external=False, se=False):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, _record_dtype(external, se))
    for name in r.dtype.names:
        shape = r[name].shape
        if r.dtype[name].base.kind == "i":
            r[name] = rng.integers(-2**40, 2**40, size=shape)
        else:
            r[name] = rng.standard_normal(shape) * 10.0 ** rng.integers(-12, 12, size=shape)
    return r


def _c(fmt):
    """A C printf format as Python's % operator reads it (%lld is %d)."""
    return fmt.replace("%lld", "%d")


def _ascii_text(hdr, rec, external=False, se=False):
    """A file as PeTar's writeAscii members print it."""
    e = "%26.17e"
    if external:
        lines = [_c("%lld %lld %26.17e %26.17e %26.17e %26.17e %26.17e %26.17e %26.17e\n") % (
            hdr["file_id"], hdr["n"], hdr["time"], *hdr["pos_offset"], *hdr["vel_offset"])]
    else:
        lines = [_c("%lld %lld %26.17e\n") % (hdr["file_id"], hdr["n"], hdr["time"])]
    for r in rec:
        s = _c(" ".join([e] * 7) + " %lld") % (r["mass"], *r["pos"], *r["vel"], r["binary_state"])
        if se:
            s += " ".join([e] * 4) % (r["radius"], r["dm"], r["time_record"], r["time_interrupt"])
        s += _c("%26.17e %lld %lld %lld") % (r["r_search"], r["id"], r["mass_bk"], r["status"])
        s += "%26.17e %26.17e " % (r["r_in"], r["r_out"])
        s += "%26.17e %26.17e %26.17e %26.17e %26.17e " % (*r["acc_soft"], r["pot"], r["pot_soft"])
        if external:
            s += "%26.17e " % r["pot_ext"]
        s += _c("%lld\n") % r["n_nb"]
        lines.append(s)
    return "".join(lines)


def _fill(snap, rec):
    for name in rec.dtype.names:
        snap[name] = rec[name]


def _same(snap, rec):
    for name in rec.dtype.names:
        assert np.array_equal(snap[name], rec[name]), name


@pytest.mark.parametrize("external,se", [(False, False), (True, False), (False, True), (True, True)])
def test_binary_layout_matches_petar_reader(tmp_path, external, se):
    rec = _records(257, 1, external, se)
    hdr = dict(file_id=7, n=rec.size, time=1.25, pos_offset=np.array([0.1, -0.2, 0.3]),
               vel_offset=np.array([1.0, 2.0, -3.0]))
    # ours -> PeTar's reader dtypes
    s = S.Snapshot.new(rec.size, external, se, hdr["file_id"], hdr["time"], hdr["pos_offset"], hdr["vel_offset"])
    _fill(s, rec)
    p = tmp_path / "a.bin"
    s.write(p, "binary")
    raw = p.read_bytes()
    h = np.frombuffer(raw, _header_dtype(external), count=1)
    assert h["file_id"][0] == 7 and h["n_glb"][0] == rec.size and h["time"][0] == 1.25
    if external:
        assert (h["x"][0], h["vy"][0]) == (0.1, 2.0)
    back = np.frombuffer(raw, _record_dtype(external, se), offset=_header_dtype(external).itemsize)
    assert back.size == rec.size and back.tobytes() == rec.tobytes()
    # PeTar's layout -> ours
    q = tmp_path / "b.bin"
    q.write_bytes(np.array([(7, rec.size, 1.25) + ((0.1, -0.2, 0.3, 1.0, 2.0, -3.0) if external else ())],
                           _header_dtype(external)).tobytes() + rec.tobytes())
    r = S.Snapshot.read(q, "binary", external, se)
    hh = r.header()
    assert hh["n"] == rec.size and hh["file_id"] == 7 and hh["time"] == 1.25
    if external:
        assert np.array_equal(hh["vel_offset"], [1.0, 2.0, -3.0])
    _same(r, rec)


@pytest.mark.parametrize("external,se", [(False, False), (True, True)])
def test_ascii_as_petar_prints_it(tmp_path, external, se):
    rec = _records(300, 2, external, se)
    hdr = dict(file_id=3, n=rec.size, time=0.5, pos_offset=np.array([1e-3, 2e-3, 3e-3]),
               vel_offset=np.array([-1e-3, 0.0, 5.0]))
    p = tmp_path / "petar.dat"
    p.write_text(_ascii_text(hdr, rec, external, se))
    s = S.Snapshot.read(p, "ascii", external, se)
    h = s.header()
    assert (h["file_id"], h["n"], h["time"]) == (3, rec.size, 0.5)
    if external:
        assert np.array_equal(h["pos_offset"], hdr["pos_offset"])
    _same(s, rec)  # %26.17e round-trips FP64 exactly
    # our ASCII writer reads back exactly, and so does binary <-> ASCII
    q = tmp_path / "ours.dat"
    s.write(q, "ascii")
    _same(S.Snapshot.read(q, "ascii", external, se), rec)
    b = tmp_path / "ours.bin"
    s.write(b, "binary")
    _same(S.Snapshot.read(b, "binary", external, se), rec)


def test_layout_mismatch_and_damage_are_errors(tmp_path):
    rec = _records(10, 3)
    s = S.Snapshot.new(rec.size)
    _fill(s, rec)
    b, a = tmp_path / "x.bin", tmp_path / "x.dat"
    s.write(b, "binary")
    s.write(a, "ascii")
    with pytest.raises(RuntimeError, match="check the build flags"):
        S.Snapshot.read(b, "binary", external=True)
    with pytest.raises(RuntimeError, match="check the build flags"):
        S.Snapshot.read(a, "ascii", stellar_evolution=True)
    h = tmp_path / "h.dat"
    h.write_text("1 0 0.0 4.0\n")
    with pytest.raises(RuntimeError, match="extra items on the header"):
        S.Snapshot.read(h, "ascii")
    t = tmp_path / "t.bin"
    t.write_bytes(b.read_bytes()[:-8])
    with pytest.raises(RuntimeError, match="need"):
        S.Snapshot.read(t, "binary")
    lines = a.read_text().splitlines(keepends=True)
    (tmp_path / "short.dat").write_text("".join(lines[:-1]))
    with pytest.raises(RuntimeError, match="records for a header count"):
        S.Snapshot.read(tmp_path / "short.dat", "ascii")
    lines[4] = lines[4].replace("e", "x", 1)
    (tmp_path / "bad.dat").write_text("".join(lines))
    with pytest.raises(RuntimeError, match="record 3"):
        S.Snapshot.read(tmp_path / "bad.dat", "ascii")
    with pytest.raises(RuntimeError, match="cannot open"):
        S.Snapshot.read(tmp_path / "missing.dat")
    with pytest.raises(KeyError):
        s["pot_ext"]


def test_snapshot_drives_the_tree(tmp_path):
    """A cluster written as a snapshot and read back gives the same tree
    lists as the arrays it came from."""
    m, x, v = host.cluster(5000, "plummer", kroupa=True, f_bin=0.3, seed=9, with_vel=True)
    rs = 0.015 * np.maximum(np.cbrt(m / m.mean()), 1.0)
    s = S.Snapshot.new(m.size, time=2.0)
    s["mass"], s["pos"], s["vel"], s["r_search"] = m, x, v, rs
    s["id"] = np.arange(1, m.size + 1)
    s["r_in"], s["r_out"] = rs / 15.0, rs / 1.5
    p = tmp_path / "data.0"
    s.write(p, "ascii")
    inp = S.soft_inputs(S.Snapshot.read(p, "ascii"))
    assert np.array_equal(inp["pos"], x) and np.array_equal(inp["id"], np.arange(1, m.size + 1))
    a = host.build_lists(inp["pos"], inp["mass"], inp["r_search"])
    b = host.build_lists(x, m, rs)
    assert np.array_equal(a.id_epj, b.id_epj) and np.array_equal(a.spj_quad, b.spj_quad)
