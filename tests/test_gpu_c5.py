"""The full-scale adaptive configuration, reduced to 64^3 and 3 levels.

astg.scenes.full_scale_compare: two r=0.05 sources with +-0.5 lateral inflow,
addForce and the step (projection to 1e-6, tilt adaptivity) for 30 steps, the
layout rebuilt from the density field at steps 10 and 20 (fields carried over
by sampling), first in FP64, then in mixed precision on the FP64 run's
layouts. The rebuild has no reference counterpart (the reference builds its
layout once), so the mixed run is measured against the FP64 run:
  * every projection converges and every sampled field is finite (a NaN
    would raise in the solver or in the sampling);
  * the rebuilds change the layout (the leaf count follows the smoke);
  * per field, at steps 10, 20 and 30 on a 16^3 lattice: density and heat
    within 1e-5 relative L2 of FP64, velocity within 1e-3 (both runs stop at
    the 1e-6 infinity-norm residual, a few iterations apart, so the velocity
    differs by about that tolerance times the iteration-to-iteration change:
    measured 1e-7 and 6e-5).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_c5_reduced_mixed_against_fp64():
    from astg import scenes
    cmp = scenes.full_scale_compare(R=64, levels=3, steps=30, rebuild=10, lattice=16,
                                    log=lambda m: print(m, flush=True))
    f64, mx = cmp["fp64"], cmp["mixed"]
    assert len(f64["splits"]) == 2 and len(mx["splits"]) == 2
    leaves = [scenes.leaf_count(scenes.full_scale_initial_split(64, 3))] + \
             [scenes.leaf_count(s.astype(np.int32)) for s in f64["splits"]]
    print("leaves at steps 0, 10, 20: %s" % leaves)
    assert len(set(leaves)) > 1
    for o in (f64, mx):
        assert min(o["iterations"]) > 0
    for e in cmp["errors"]:
        print("step %d: " % e["step"] + "  ".join("%s %.2e" % (f, e[f]["rel_l2"]) for f in ("density", "heat",
                                                                                         "velocity")), flush=True)
        assert e["density"]["rel_l2"] < 1e-5 and e["heat"]["rel_l2"] < 1e-5
        assert e["velocity"]["rel_l2"] < 1e-3
    assert [e["step"] for e in cmp["errors"]] == [10, 20, 30]
