"""Writes tests/golden/smoke64.npz: the plain smoke configuration (64^3,
3 levels, one rising source; astg.scenes.smoke_64) run through the reference's
own onUpdate (octree_omnigrid_solver3.cpp:222-249: sceneSetup, addForce,
constraint, project, advect with tilt adaptivity) for 50 steps from rest.

Runs only where oracle/_ref/libastgref.so exists (the reference solver built
from its own sources by oracle/ref.mk); the split map comes from the host-only
astg library. The fixture is data only: the split map and scene constants,
and after steps 1, 10, 25 and 50 the density, heat and velocity per node of
every level and the tilt state.

    python tests/golden/make_golden_smoke.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ast-grid_amd"))

from oracle import ref  # noqa: E402
from astg import scenes  # noqa: E402

STEPS = 50
CHECKPOINTS = (1, 10, 25)  # and the last step, stored as "final"
TOL = 1e-10  # far from the stopping boundary, so two correct solvers agree to ~1e-10


def snapshot(r, L, sizes, tag, out):
    for l in range(L):
        for f in (0, 1):
            o, t = r.scalar(f, l)
            out["%s_s%d_oct_%d" % (tag, f, l)], out["%s_s%d_tilt_%d" % (tag, f, l)] = o[:sizes[l]], t[:sizes[l]]
        m, t = r.velocity(l)
        out["%s_vel_mac_%d" % (tag, l)], out["%s_vel_tilt_%d" % (tag, l)] = m[:sizes[l]], t[:sizes[l]]
        _, bits, _ = r.tilt(l)
        out["%s_bits_%d" % (tag, l)] = bits[:sizes[l]]


def main():
    R, L = 64, 3
    split = scenes.smoke_64_split(R, L)
    cx, cy, cz, rad = scenes.SMOKE_SOURCE
    r = ref.RefSolver(split, R, L, scenes.CLOSED_XY)
    r.set_source((cx, cy, cz), rad, 1.0, 1.0, 1.0)
    r.set_buoyancy(1.0, 0.0)
    r.set_tilt_adaptivity(True)
    sizes = [r.level_size(l) for l in range(L)]
    out = {"R": np.array(R), "levels": np.array(L), "split": split.astype(np.int32), "steps": np.array(STEPS),
           "dt": np.array(0.5 / R), "rho": np.array(2.0), "tol": np.array(TOL),
           "checkpoints": np.array(CHECKPOINTS)}
    for k in range(1, STEPS + 1):
        r.step(0.5 / R, tol=TOL)
        if k in CHECKPOINTS:
            snapshot(r, L, sizes, "step%d" % k, out)
    snapshot(r, L, sizes, "final", out)
    r.close()
    np.savez_compressed(os.path.join(HERE, "smoke64.npz"), **out)
    print("wrote smoke64.npz (%d steps; sizes %s)" % (STEPS, sizes))


if __name__ == "__main__":
    main()
