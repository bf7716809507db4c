"""The smoke-configuration fixture (tests/golden/smoke64.npz, written by
tests/golden/make_golden_smoke.py from the reference's own onUpdate): built on
the same split map as astg.scenes.smoke_64, and reproducible from the
reference build when it is present."""
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden

SMOKE = os.path.join(ROOT, "tests", "golden", "smoke64.npz")


def test_fixture_uses_the_scene_split_map():
    from astg import scenes
    z = load_golden(SMOKE)
    assert np.array_equal(z["split"], scenes.smoke_64_split(int(z["R"]), int(z["levels"])))
    # the run moved smoke: density above the source level and open tilts
    assert z["final_s0_oct_1"].sum() > z["step1_s0_oct_1"].sum() > 0
    assert ((z["final_bits_0"] & 64) != 0).sum() > ((z["step1_bits_0"] & 64) != 0).sum()


def test_fixture_reproducible_from_reference():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build not present (oracle/_ref)")
    import importlib.util
    import shutil
    import tempfile
    spec = importlib.util.spec_from_file_location("mgs", os.path.join(ROOT, "tests", "golden", "make_golden_smoke.py"))
    mgs = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mgs)
    old = load_golden(SMOKE)
    tmp = tempfile.mkdtemp()
    try:
        mgs.HERE = tmp
        mgs.main()
        new = load_golden(os.path.join(tmp, "smoke64.npz"))
    finally:
        shutil.rmtree(tmp)
    assert sorted(old) == sorted(new)
    for k in old:
        assert np.array_equal(old[k], new[k]), k
