"""Writes tests/golden/frame32.npz: a frame in the reference's binary formats
and the reference's own restart from it.

The plain smoke configuration at 32^3, 3 levels (astg.scenes.smoke_64_split
with R=32), run through the reference's onUpdate for 3 steps with tilt
adaptivity (oracle/_ref, the reference built from its own sources). Then:

  * the five frame files the reference writes (dumpOctreeLayout3dBinary,
    dumpOctreeOmniTiltE3dBinary, dumpOctreeOmniScalar3dBinary for density and
    heat, dumpOctreeOmniFaceCentered3dBinary; src/viewer/grid_dump.cpp), kept
    byte for byte as uint8 arrays;
  * a fresh reference solver on the same split map restarted from those files
    (octree_omnigrid_solver3.cpp:1015-1057): its tilt state, velocity, dual
    grid and scalars per level;
  * that restarted solver after one more step (tolerance 1e-10): velocity,
    scalars and tilt bits per level.

    python tests/golden/make_golden_frame.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ast-grid_amd"))

from oracle import ref  # noqa: E402
from astg import scenes  # noqa: E402

R, L, STEPS = 32, 3, 3
TOL = 1e-10
FILES = ("layout", "tilt", "density", "heat", "velocity")


def state(r, sizes, tag, out, dual=True):
    for l in range(L):
        n = sizes[l]
        e, bits, hd = r.tilt(l)
        out["%s_tilt_e_%d" % (tag, l)], out["%s_tilt_bits_%d" % (tag, l)] = e[:n], bits[:n]
        out["%s_tilt_hd_%d" % (tag, l)] = hd[:n]
        m, t = r.velocity(l)
        out["%s_mac_%d" % (tag, l)], out["%s_tilt8_%d" % (tag, l)] = m[:n], t[:n]
        for f in (0, 1):
            o, t = r.scalar(f, l)
            out["%s_s%d_oct_%d" % (tag, f, l)], out["%s_s%d_tilt_%d" % (tag, f, l)] = o[:n], t[:n]
        if dual:
            do, dt = r.dual(l)
            out["%s_dual_oct_%d" % (tag, l)], out["%s_dual_tilt_%d" % (tag, l)] = do[:n], dt[:n]


def main():
    split = scenes.smoke_64_split(R, L)
    cx, cy, cz, rad = scenes.SMOKE_SOURCE
    dt = 0.5 / R
    out = {"R": np.array(R), "levels": np.array(L), "split": split.astype(np.int32), "dt": np.array(dt),
           "rho": np.array(2.0), "tol": np.array(TOL), "closed": np.array(scenes.CLOSED_XY)}

    def solver():
        r = ref.RefSolver(split, R, L, scenes.CLOSED_XY)
        r.set_source((cx, cy, cz), rad, 1.0, 1.0, 1.0)
        r.set_buoyancy(1.0, 0.0)
        r.set_tilt_adaptivity(True)
        return r

    with tempfile.TemporaryDirectory() as tmp:
        paths = {k: os.path.join(tmp, k + ".bin") for k in FILES}
        r = solver()
        # node counts before any step: lookups of absent coordinates later add
        # default records to the reference's maps, which its level sizes count
        sizes = [r.level_size(l) for l in range(L)]
        for _ in range(STEPS):
            r.step(dt, tol=TOL)
        r.write_frame(**paths)
        r.close()
        for k in FILES:
            out["file_" + k] = np.fromfile(paths[k], np.uint8)
        cwd = os.getcwd()
        os.chdir(tmp)  # restart writes a preview image under ./astg_ref_harness/
        try:
            r2 = solver()
            r2.restart(**paths)
        finally:
            os.chdir(cwd)
        state(r2, sizes, "restart", out)
        r2.step(dt, tol=TOL)
        state(r2, sizes, "next", out, dual=False)
        r2.close()
    np.savez_compressed(os.path.join(HERE, "frame32.npz"), **out)
    print("wrote frame32.npz (sizes %s, files %s bytes)" % (sizes, [len(out["file_" + k]) for k in FILES]))


if __name__ == "__main__":
    main()
