"""Advection fixtures: well-formed, and reproducible from the reference build
(when oracle/_ref is present in this container)."""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden

ADV = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "adv_*.npz")))


def test_fixtures_cover_the_chain():
    assert len(ADV) >= 3
    for p in ADV:
        a = load_golden(p)
        g = load_golden(os.path.join(ROOT, "tests", "golden", str(a["scene"]) + ".npz"))
        L = int(a["levels"])
        for l in range(L):
            n = g["u_mac_out_%d" % l].size // 3
            for key in ("s0_in_oct", "s0_ghost_oct", "s0_adv_oct", "s1_adv_tilt"):
                assert a["%s_%d" % (key, l)].shape == (n,), (p, key, l)
            assert a["dual_oct_%d" % l].shape == (n, 3)
            assert a["vel_adv_tilt_%d" % l].shape == (n * 8,) or a["vel_adv_tilt_%d" % l].size == n * 8
        assert a["q_vel"].shape == a["pts"].shape and a["q_den"].shape == (len(a["pts"]),)
        # a midpoint back-trace moves a point by at most dt * max |u| over the field
        assert np.all(np.isfinite(a["q_bt"]))


def test_fixture_reproducible_from_reference():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build not present (oracle/_ref)")
    import importlib.util
    spec = importlib.util.spec_from_file_location("mga", os.path.join(ROOT, "tests", "golden", "make_golden_advect.py"))
    mga = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mga)
    name = "s32_two_spheres"
    old = load_golden(os.path.join(ROOT, "tests", "golden", "adv_%s.npz" % name))
    import shutil
    import tempfile
    tmp = tempfile.mkdtemp()
    try:
        src = os.path.join(ROOT, "tests", "golden", name + ".npz")
        shutil.copy(src, tmp)
        mga.HERE = tmp
        mga.make(name)
        new = load_golden(os.path.join(tmp, "adv_%s.npz" % name))
    finally:
        shutil.rmtree(tmp)
    for k in old:
        if old[k].dtype.kind == "f":
            assert np.allclose(new[k], old[k], rtol=0, atol=1e-12 * max(1.0, np.abs(old[k]).max())), k
