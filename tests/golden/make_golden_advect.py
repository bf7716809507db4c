"""Writes tests/golden/adv_*.npz: the reference's post-projection chain,
point sampling and one advection step on the projected velocity of a
projection fixture (tests/golden/<scene>.npz).

Runs only where oracle/_ref/libastgref.so exists (the reference solver built
from its own sources by oracle/ref.mk). Each fixture is data only:
  * inputs: seeded uniform density and heat (octagon and tilt value per node,
    every level), the fixture's corrected velocity, query points, dt = 0.5/R;
  * after onUpdate's chain between project and advect
    (octree_omnigrid_solver3.cpp:234-241): velocity per node (MAC, tilt) and
    the dual grid (cell and vertex vectors);
  * the scalars after OctreeOmniGrid3::updateGhost;
  * point queries: omni velocity sample, MAC sample, BackTrace, density sample;
  * after advect(dt) (850-914): density, heat and velocity per node.

    python tests/golden/make_golden_advect.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402

SCENES = ["s32_two_spheres", "s32_closed_box_tilts", "s32_l4_open_tilts"]
NPTS = 1000


def make(name):
    z = np.load(os.path.join(HERE, name + ".npz"))
    g = {k: z[k] for k in z.files}
    R, L = int(g["R"]), int(g["levels"])
    r = ref.RefSolver(g["split"], R, L, int(g["closed"]))
    if int(g["open_tilts"]):
        r.set_tilt_open(0, g["open_mask_l0"])
    for l in range(L):
        r.set_velocity(l, g["u_mac_out_%d" % l], g["u_tilt_out_%d" % l])
    rng = np.random.default_rng(7)
    out = {"scene": np.array(name), "levels": np.array(L)}
    sizes = [r.level_size(l) for l in range(L)]
    for f in (0, 1):
        for l in range(L):
            o, t = rng.random(sizes[l]), rng.random(sizes[l])
            r.set_scalar(f, l, o, t)
            out["s%d_in_oct_%d" % (f, l)], out["s%d_in_tilt_%d" % (f, l)] = o, t
    r.post_project()
    for l in range(L):
        out["vel_post_mac_%d" % l], out["vel_post_tilt_%d" % l] = r.velocity(l)
        out["dual_oct_%d" % l], out["dual_tilt_%d" % l] = r.dual(l)
    for f in (0, 1):
        r.scalar_update_ghost(f)
        for l in range(L):
            out["s%d_ghost_oct_%d" % (f, l)], out["s%d_ghost_tilt_%d" % (f, l)] = r.scalar(f, l)
    pts = rng.random((NPTS, 3)) * 0.98 + 0.01
    dt = 0.5 / R
    out["pts"], out["dt"] = pts, np.array(dt)
    out["q_vel"] = r.sample_points(0, pts)
    out["q_mac"] = r.sample_points(1, pts)
    out["q_bt"] = r.sample_points(2, pts, dt)
    out["q_den"] = r.sample_scalar(0, pts)
    r.advect(dt)
    for l in range(L):
        # the reference's temporaries may be longer than the layer: keep node values
        for f in (0, 1):
            o, t = r.scalar(f, l)
            out["s%d_adv_oct_%d" % (f, l)], out["s%d_adv_tilt_%d" % (f, l)] = o[:sizes[l]], t[:sizes[l]]
        m, t = r.velocity(l)
        out["vel_adv_mac_%d" % l], out["vel_adv_tilt_%d" % l] = m[:sizes[l]], t[:sizes[l]]
    np.savez_compressed(os.path.join(HERE, "adv_" + name + ".npz"), **out)
    r.close()


if __name__ == "__main__":
    for s in SCENES:
        make(s)
        print("wrote adv_%s.npz" % s)
