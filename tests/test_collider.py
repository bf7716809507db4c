"""The collider path on the host: boundary SDF sampling, the -3h numbering cut
and the cutting-cell weights against the reference run with its own sphere
collider (tests/golden/col_*.npz, written by make_golden_collider.py).

* The library's sphere SDF equals the reference collider's SDF grid to
  rounding (1e-15 absolute).
* Numbering (updateIndexParallel with the cut, octree_omnigrid_solver3.cpp:
  292-347) and the CSR pattern: bit-exact.
* Weights (initCuttingCellWeights, 1169-1238, fractionInside util.h:340-349):
  within 1e-14 (the reference divides under -ffast-math), including the
  half-tilt quirks; every weight below 1 is where the reference has one.
* Removing the collider restores the no-collider numbering and weights.
"""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden

COL = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "col_*.npz")))


@pytest.fixture(params=COL, ids=[os.path.basename(p)[:-4] for p in COL])
def col(request):
    return load_golden(request.param)


def collider_context(g, device=-1, sdf=None):
    import astg
    c = astg.Context(device)
    R, L = int(g["R"]), int(g["levels"])
    c.set_topology(g["split"].astype(np.int32), R, L, int(g["closed"]))
    if int(g["open_tilts"]):
        c.set_tilt_open(0, g["open_mask_l0"])
    c.set_boundary_sdf(g["sdf"] if sdf is None else sdf)
    return c


def test_sphere_sdf_matches_reference_collider(col):
    import astg
    R = int(col["R"])
    cx, cy, cz, r = col["collider"]
    s = astg.sphere_sdf(R, (cx, cy, cz), r)
    assert np.max(np.abs(s - col["sdf"])) < 1e-15


def test_numbering_with_collider(col):
    c = collider_context(col)
    cut = 0
    for l in range(int(col["levels"])):
        o, t, n = c.numbering(l)
        assert n == len(col["b"])
        assert np.array_equal(o, col["pid_oct_%d" % l]), "octagon numbering, level %d" % l
        assert np.array_equal(t, col["pid_tilt_%d" % l]), "tilt numbering, level %d" % l
    rs, cl = c.pattern()
    assert np.array_equal(rs, col["A_rowstart"]) and np.array_equal(cl, col["A_col"])
    # where cells lie more than 3h inside the collider, the cut removes them
    c.set_boundary_sdf(None)
    _, _, n0 = c.numbering(0)
    R = int(col["R"])
    cut = (col["sdf"] <= -3.0 / R).sum()
    assert (n0 > len(col["b"])) == (cut > 0)


def test_weights_with_collider(col):
    c = collider_context(col)
    below = 0
    for l in range(int(col["levels"])):
        m, t = c.weights(l)
        assert np.max(np.abs(m - col["w_mac_%d" % l])) <= 1e-14, l
        assert np.max(np.abs(t - col["w_tilt_%d" % l])) <= 1e-14, l
        assert np.array_equal(m < 1, col["w_mac_%d" % l] < 1)
        below += int((col["w_mac_%d" % l] < 1).sum())
    assert below > 0  # the collider cuts faces


def test_point_inside_sdf_changes_weights_only_near_it(col):
    """A positive SDF everywhere is the no-collider case: weights 1 except
    the half-tilt records, as without an SDF."""
    import astg
    g = col
    R = int(g["R"])
    c1 = collider_context(g, sdf=np.full(R ** 3, 5.0))
    c0 = astg.Context(-1)
    c0.set_topology(g["split"].astype(np.int32), R, int(g["levels"]), int(g["closed"]))
    if int(g["open_tilts"]):
        c0.set_tilt_open(0, g["open_mask_l0"])
    for l in range(int(g["levels"])):
        for a, b in zip(c1.weights(l), c0.weights(l)):
            assert np.array_equal(a, b)
        for a, b in zip(c1.numbering(l), c0.numbering(l)):
            assert np.array_equal(np.asarray(a), np.asarray(b))


def test_bad_sdf_is_refused(col):
    import astg
    g = col
    R = int(g["R"])
    c = astg.Context(-1)
    c.set_topology(g["split"].astype(np.int32), R, int(g["levels"]), int(g["closed"]))
    bad = np.array(g["sdf"])
    bad[5] = np.nan
    with pytest.raises(astg.AstgError) as e:
        c.set_boundary_sdf(bad)
    assert e.value.code == astg.ASTG_ERR_ARG
