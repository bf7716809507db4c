"""On-disk frames in the reference's binary formats (src/viewer/grid_dump.cpp)
against files the reference itself wrote (tests/golden/frame32.npz, made by
make_golden_frame.py from the reference build): host-side parts, no GPU.

* The reference's layout file describes exactly the layers the library makes
  from the same split map (node indices, coordinates, flags, leaf and ghost
  orders, half-tilt list), and a layout from a different split map is refused.
* Reading the reference's tilt file gives the reference's tilt state, and
  writing it back reproduces every record of the reference's file (the
  reference leaves the records' padding bytes uninitialised; they are not
  compared).
* The layout the library writes carries the same records and lists as the
  reference's (whose map order is its hash map's, plus default records left
  by lookups).
* Truncated and mismatched files are refused without changing the state.
* Where the reference build is present, the reference reads the library's
  layout and tilt files in its own restart and ends in the same tilt state.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden

FIXTURE = os.path.join(ROOT, "tests", "golden", "frame32.npz")
FILES = ("layout", "tilt", "density", "heat", "velocity")

TILT = np.dtype([("e", "<f8"), ("flags", "u1", 6), ("p0", "u1", 2), ("hd", "<i4"), ("p1", "<u4")])
REC = np.dtype([("key", "<i8"), ("index", "<u8"), ("x", "<u2"), ("y", "<u2"), ("z", "<u2"), ("p0", "<u2"),
                ("flag", "<u4"), ("p1", "<u4")])


@pytest.fixture(scope="module")
def frame():
    return load_golden(FIXTURE)


def write_files(g, d):
    paths = {k: os.path.join(str(d), k + ".bin") for k in FILES}
    for k in FILES:
        g["file_" + k].tofile(paths[k])
    return paths


def host_context(g, split=None):
    import astg
    c = astg.Context(-1)
    c.set_topology((g["split"] if split is None else split).astype(np.int32), int(g["R"]), int(g["levels"]),
                   int(g["closed"]))
    return c


def parse_layout(path):
    b = open(path, "rb").read()
    off = 0

    def take(dt, n=1):
        nonlocal off
        a = np.frombuffer(b, dt, n, off)
        off += a.nbytes
        return a
    levels = []
    L = int(take("<i4")[0])
    for _ in range(L):
        recs = take(REC, int(take("<i4")[0]))
        ghosts = take("<i8", int(take("<i4")[0]))
        leaves = take("<i8", int(take("<i4")[0]))
        recs = recs[recs["flag"] != 0]  # default records from lookups of absent coordinates
        recs = np.sort(recs, order="index")
        levels.append((recs[["key", "index", "x", "y", "z", "flag"]], ghosts, leaves))
    half = take("<i8", int(take("<i4")[0]))
    assert off == len(b)
    return levels, half


def parse_tilt(b):
    """Records of a tilt file without their padding bytes."""
    b = b.tobytes()
    off, recs = 0, []
    L = int(np.frombuffer(b, "<i4", 1, off)[0])
    off += 4
    for _ in range(L):
        n = int(np.frombuffer(b, "<i4", 1, off)[0])
        off += 4
        r = np.frombuffer(b, TILT, n, off)
        off += r.nbytes
        recs.append(r[["e", "flags", "hd"]])
    assert off == len(b)
    return np.concatenate(recs)


def tilt_of(c, L):
    return [c.tilt(l) for l in range(L)]


def test_reference_layout_and_tilt_read(frame, tmp_path):
    g = frame
    L = int(g["levels"])
    p = write_files(g, tmp_path)
    c = host_context(g)
    c.read_frame(layout=p["layout"], tilt=p["tilt"])
    for l in range(L):
        e, bits, hd = c.tilt(l)
        n = len(g["restart_tilt_e_%d" % l])
        assert len(e) == n
        # tiltE part (bits 0-5), e and half direction: the restarted reference's
        assert np.array_equal(e, g["restart_tilt_e_%d" % l])
        assert np.array_equal(bits & 0x3F, g["restart_tilt_bits_%d" % l] & 0x3F)
        assert np.array_equal(hd, g["restart_tilt_hd_%d" % l])
        # the numbering's open / fixed bits stay those of the layout, as the
        # reference's restart keeps tilt_layout
        assert np.array_equal(bits & 0xC0, g["restart_tilt_bits_%d" % l] & 0xC0)
    out = tmp_path / "tilt_out.bin"
    c.write_frame(tilt=str(out))
    assert np.array_equal(parse_tilt(np.fromfile(str(out), np.uint8)), parse_tilt(g["file_tilt"]))


def test_layout_written_matches_reference(frame, tmp_path):
    g = frame
    p = write_files(g, tmp_path)
    c = host_context(g)
    out = str(tmp_path / "layout_out.bin")
    c.write_frame(layout=out)
    ours, half = parse_layout(out)
    theirs, rhalf = parse_layout(p["layout"])
    assert np.array_equal(half, rhalf) and len(ours) == len(theirs)
    for (a, ga, la), (b, gb, lb) in zip(ours, theirs):
        assert np.array_equal(a, b) and np.array_equal(ga, gb) and np.array_equal(la, lb)
    c.read_frame(layout=out)  # and it passes the layout check


def test_mismatched_and_truncated_files_are_refused(frame, tmp_path):
    import astg
    g = frame
    L = int(g["levels"])
    p = write_files(g, tmp_path)
    # a different split map (the finest cells coarsened by one level): the
    # layout check fails
    split2 = g["split"].copy()
    split2[split2 == 0] = 1
    c2 = host_context(g, split=split2)
    with pytest.raises(astg.AstgError) as ei:
        c2.read_frame(layout=p["layout"])
    assert ei.value.code == astg.ASTG_ERR_IO
    # truncated tilt file: refused, state unchanged
    c = host_context(g)
    before = tilt_of(c, L)
    short = tmp_path / "tilt_short.bin"
    g["file_tilt"][:-10].tofile(str(short))
    with pytest.raises(astg.AstgError):
        c.read_frame(tilt=str(short))
    for (e0, b0, h0), (e1, b1, h1) in zip(before, tilt_of(c, L)):
        assert np.array_equal(e0, e1) and np.array_equal(b0, b1) and np.array_equal(h0, h1)
    # fields need a device context; a missing file is an error
    with pytest.raises(astg.AstgError):
        c.read_frame(density=p["density"])
    with pytest.raises(astg.AstgError):
        c.read_frame(layout=str(tmp_path / "missing.bin"))


def bad_half_direction(g, tmp_path):
    """The reference's tilt file with its first record turned into a half-tilt
    whose half direction is -1 (no face: not a state the reference writes)."""
    b = bytearray(g["file_tilt"].tobytes())
    off = 8  # level count, then the first level's record count
    rec = np.frombuffer(bytes(b[off:off + TILT.itemsize]), TILT).copy()
    rec["flags"][0][5] = 1
    rec["hd"][0] = -1
    b[off:off + TILT.itemsize] = rec.tobytes()
    path = tmp_path / "tilt_bad_half.bin"
    path.write_bytes(bytes(b))
    return str(path)


def test_half_tilt_without_direction_is_refused(frame, tmp_path):
    import astg
    g = frame
    L = int(g["levels"])
    c = host_context(g)
    before = tilt_of(c, L)
    with pytest.raises(astg.AstgError) as ei:
        c.read_frame(tilt=bad_half_direction(g, tmp_path))
    assert ei.value.code == astg.ASTG_ERR_IO
    for (e0, b0, h0), (e1, b1, h1) in zip(before, tilt_of(c, L)):
        assert np.array_equal(e0, e1) and np.array_equal(b0, b1) and np.array_equal(h0, h1)


def test_reference_restarts_from_our_files(frame, tmp_path, monkeypatch):
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build (oracle/_ref) not present")
    g = frame
    L = int(g["levels"])
    p = write_files(g, tmp_path)
    c = host_context(g)
    c.read_frame(layout=p["layout"], tilt=p["tilt"])
    ours = {"layout": str(tmp_path / "our_layout.bin"), "tilt": str(tmp_path / "our_tilt.bin")}
    c.write_frame(**ours)
    r = ref.RefSolver(g["split"], int(g["R"]), L, int(g["closed"]))
    monkeypatch.chdir(tmp_path)  # restart writes a preview image under ./astg_ref_harness/
    r.restart(ours["layout"], ours["tilt"], p["density"], p["heat"], p["velocity"])
    for l in range(L):
        e, bits, hd = r.tilt(l)
        n = len(g["restart_tilt_e_%d" % l])
        assert np.array_equal(e[:n], g["restart_tilt_e_%d" % l])
        assert np.array_equal(bits[:n], g["restart_tilt_bits_%d" % l])
        assert np.array_equal(hd[:n], g["restart_tilt_hd_%d" % l])
    r.close()
