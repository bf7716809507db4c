"""The plain smoke configuration on the device against the reference's own
onUpdate (tests/golden/smoke64.npz): 64^3, 3 levels, a rising source, buoyancy
and tilt adaptivity, 50 steps from rest (astg.scenes.smoke_step: source,
addForce, constraint and ghosts, projection, post-projection chain, adjustE,
advection). Both sides project to a 1e-10 relative residual, far from the
stopping boundary (at 1e-6 two correct solvers legitimately differ by up to
~1e-6 in the small velocities of the coarse levels). After steps 1, 10, 25 and
50: tilt states bit-exact, density, heat and velocity within 1e-8 relative L2
per level."""
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden

pytestmark = pytest.mark.gpu

SMOKE = os.path.join(ROOT, "tests", "golden", "smoke64.npz")


def rel2(x, y):
    n = np.linalg.norm(y)
    return np.linalg.norm(np.asarray(x) - np.asarray(y)) / (n if n > 0 else 1.0)


def check(ctx, z, tag, L):
    for l in range(L):
        _, bits, _ = ctx.tilt(l)
        assert np.array_equal(bits, z["%s_bits_%d" % (tag, l)]), (tag, l)
        for f in (0, 1):
            o, t = ctx.scalar(f, l)
            assert rel2(o, z["%s_s%d_oct_%d" % (tag, f, l)]) < 1e-8, (tag, f, l)
            assert rel2(t, z["%s_s%d_tilt_%d" % (tag, f, l)]) < 1e-8, (tag, f, l)
        m, t = ctx.velocity(l)
        assert rel2(m, z["%s_vel_mac_%d" % (tag, l)]) < 1e-8, (tag, l)
        assert rel2(t, z["%s_vel_tilt_%d" % (tag, l)]) < 1e-8, (tag, l)


def test_smoke_configuration_matches_reference():
    import astg
    from astg import scenes
    z = load_golden(SMOKE)
    ctx, info = scenes.smoke_64(device=0, R=int(z["R"]), levels=int(z["levels"]))
    L = info["levels"]
    prm = astg.default_params(tol=float(z["tol"]), max_iterations=3000)
    checkpoints = set(int(c) for c in z["checkpoints"])
    for k in range(1, int(z["steps"]) + 1):
        res = scenes.smoke_step(ctx, info, k == 1, prm)
        assert res.converged == 1
        if k in checkpoints:
            print("smoke step %d: checking against the reference" % k, flush=True)
            check(ctx, z, "step%d" % k, L)
    check(ctx, z, "final", L)
    ctx.close()
