"""Tilt-adaptivity fixtures (tests/golden/adj_*.npz, written from the
reference's adjustE by tests/golden/make_golden_adjust.py): consistent with the
rules of adjustE (octree_omnigrid_solver3.cpp:1125-1166) against the host
topology, and reproducible from the reference build when it is present."""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden, make_context

ADJ = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "adj_*.npz")))
OPEN, FIXED = 64, 128


def scene(p):
    j = load_golden(p)
    g = load_golden(os.path.join(ROOT, "tests", "golden", str(j["scene"]) + ".npz"))
    a = load_golden(os.path.join(ROOT, "tests", "golden", "adv_" + str(j["scene"]) + ".npz"))
    return g, a, j


def test_fixtures_follow_the_adjust_rules():
    assert len(ADJ) >= 3
    flips = 0
    for p in ADJ:
        g, a, j = scene(p)
        c = make_context(g, -1)  # host topology: the tilt state before adjustE
        L = int(g["levels"])
        h0 = 1.0 / int(g["R"])
        for l in range(L):
            nleaf = c.level_info(l)[1]
            e0, b0, _ = c.tilt(l)
            e1, b1 = j["adj_e_%d" % l], j["adj_bits_%d" % l]
            assert b1.shape == b0.shape
            leaf = np.arange(len(b0)) < nleaf
            fixed = leaf & ((b0 & FIXED) != 0)
            free = leaf & ~fixed
            # fixed tilts keep their state
            assert np.array_equal(b1[fixed], b0[fixed]) and np.array_equal(e1[fixed], e0[fixed])
            open1 = (b1 & OPEN) != 0
            # committed sizes: h/2 when open, 0 when closed
            h = h0 * (1 << l)
            assert np.all(e1[free & open1] == h * 0.5) and np.all(e1[free & ~open1] == 0.0)
            # open tilts whose density left (0.02, 1000) are the ones that closed
            d = a["s0_in_tilt_%d" % l]
            was_open = free & ((b0 & OPEN) != 0)
            out_of_range = (d < float(j["d_threshold"])) | (d > float(j["max_threshold"]))
            assert np.array_equal(open1[was_open], ~out_of_range[was_open])
            flips += int((open1 != ((b0 & OPEN) != 0))[free].sum())
            # the adaptive step commits the same decision
            assert np.array_equal(j["step_bits_%d" % l], b1) and np.array_equal(j["step_e_%d" % l], e1)
    assert flips > 100  # the fixtures exercise both directions of the update


def test_fixture_reproducible_from_reference():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build not present (oracle/_ref)")
    import importlib.util
    import shutil
    import tempfile
    spec = importlib.util.spec_from_file_location("mgj", os.path.join(ROOT, "tests", "golden", "make_golden_adjust.py"))
    mgj = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mgj)
    name = "s32_l4_open_tilts"
    old = load_golden(os.path.join(ROOT, "tests", "golden", "adj_%s.npz" % name))
    tmp = tempfile.mkdtemp()
    try:
        for f in (name + ".npz", "adv_" + name + ".npz"):
            shutil.copy(os.path.join(ROOT, "tests", "golden", f), tmp)
        mgj.HERE = tmp
        mgj.make(name)
        new = load_golden(os.path.join(tmp, "adj_%s.npz" % name))
    finally:
        shutil.rmtree(tmp)
    assert sorted(old) == sorted(new)
    for k in old:
        if old[k].dtype.kind in "fi":
            assert np.array_equal(old[k], new[k]), k
