"""Restart from the reference's frame files on the device, against the
reference's own restart from the same files (tests/golden/frame32.npz, made by
make_golden_frame.py): the smoke configuration at 32^3, 3 levels, after 3
reference steps with tilt adaptivity.

* After astg_read_frame: tilt state identical; density and heat identical
  (restart loads them as written); velocity and dual grid, which restart
  recomputes (ghosts, updateDualGrid, dual ghosts), within 1e-12 of the
  largest magnitude.
* One more step of the smoke configuration on both sides (projection to 1e-10,
  so two correct solvers agree far below the solve tolerance): velocity and
  scalars within 1e-8 of the largest magnitude per level, tilt bits identical.
* Writing the frame back: scalar files byte-identical to the reference's,
  velocity equal to the restarted state, tilt records identical.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden
from test_frames import FILES, parse_tilt, write_files

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(ROOT, "tests", "golden", "frame32.npz")


def close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def frame():
    return load_golden(FIXTURE)


def restarted(g, paths):
    import astg
    c = astg.Context(0)
    c.set_topology(g["split"].astype(np.int32), int(g["R"]), int(g["levels"]), int(g["closed"]))
    c.read_frame(**paths)
    return c


def test_restart_matches_reference(frame, tmp_path):
    g = frame
    L = int(g["levels"])
    c = restarted(g, write_files(g, tmp_path))
    for l in range(L):
        e, bits, hd = c.tilt(l)
        assert np.array_equal(e, g["restart_tilt_e_%d" % l])
        assert np.array_equal(bits, g["restart_tilt_bits_%d" % l])
        assert np.array_equal(hd, g["restart_tilt_hd_%d" % l])
        for f in (0, 1):
            o, t = c.scalar(f, l)
            assert np.array_equal(o, g["restart_s%d_oct_%d" % (f, l)])
            assert np.array_equal(t, g["restart_s%d_tilt_%d" % (f, l)])
        m, t8 = c.velocity(l)
        assert close(m, g["restart_mac_%d" % l], 1e-12), l
        assert close(t8, g["restart_tilt8_%d" % l], 1e-12), l
        do, dt = c.dual(l)
        assert close(do, g["restart_dual_oct_%d" % l], 1e-12), l
        assert close(dt, g["restart_dual_tilt_%d" % l], 1e-12), l


def test_step_after_restart_matches_reference(frame, tmp_path):
    import astg
    from astg import scenes
    g = frame
    L = int(g["levels"])
    c = restarted(g, write_files(g, tmp_path))
    info = {"dt": float(g["dt"]), "rho": float(g["rho"]), "source": scenes.SMOKE_SOURCE, "buoyancy": (1.0, 0.0)}
    res = scenes.smoke_step(c, info, True, astg.default_params(tol=float(g["tol"]), max_iterations=3000))
    print("step after restart: n=%d iterations %d" % (res.n, res.iterations), flush=True)
    for l in range(L):
        _, bits, _ = c.tilt(l)
        assert np.array_equal(bits, g["next_tilt_bits_%d" % l])
        m, t8 = c.velocity(l)
        assert close(m, g["next_mac_%d" % l], 1e-8), l
        assert close(t8, g["next_tilt8_%d" % l], 1e-8), l
        for f in (0, 1):
            o, t = c.scalar(f, l)
            assert close(o, g["next_s%d_oct_%d" % (f, l)], 1e-8), (f, l)
            assert close(t, g["next_s%d_tilt_%d" % (f, l)], 1e-8), (f, l)


def test_frame_round_trip(frame, tmp_path):
    g = frame
    L = int(g["levels"])
    c = restarted(g, write_files(g, tmp_path))
    out = {k: str(tmp_path / ("out_" + k + ".bin")) for k in FILES}
    c.write_frame(**out)
    for k in ("density", "heat"):
        assert np.array_equal(np.fromfile(out[k], np.uint8), g["file_" + k])
    assert np.array_equal(parse_tilt(np.fromfile(out["tilt"], np.uint8)), parse_tilt(g["file_tilt"]))
    # the written velocity is the restarted state, and reading it back changes nothing
    c2 = restarted(g, {"layout": out["layout"], "tilt": out["tilt"], "density": out["density"],
                       "heat": out["heat"], "velocity": out["velocity"]})
    for l in range(L):
        m, t8 = c2.velocity(l)
        assert close(m, g["restart_mac_%d" % l], 1e-12), l
        assert close(t8, g["restart_tilt8_%d" % l], 1e-12), l


def test_bad_tilt_file_leaves_a_working_context(frame, tmp_path):
    """A tilt file that names a half-tilt without a face is refused before any
    state changes; the context keeps its tilt state and the next projection
    runs on consistent tables."""
    import astg
    from test_frames import bad_half_direction
    g = frame
    L = int(g["levels"])
    c = restarted(g, write_files(g, tmp_path))
    before = [c.tilt(l) for l in range(L)]
    prm = astg.default_params(tol=1e-8)
    r0 = c.project(0.5 / int(g["R"]), 2.0, prm)
    with pytest.raises(astg.AstgError) as ei:
        c.read_frame(tilt=bad_half_direction(g, tmp_path))
    assert ei.value.code == astg.ASTG_ERR_IO
    for (e0, b0, h0), l in zip(before, range(L)):
        e1, b1, h1 = c.tilt(l)
        assert np.array_equal(e0, e1) and np.array_equal(b0, b1) and np.array_equal(h0, h1)
    r1 = c.project(0.5 / int(g["R"]), 2.0, prm)
    assert r1.converged and r1.n == r0.n and r1.nnz == r0.nnz
