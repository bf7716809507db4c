"""Writes tests/golden/adj_*.npz: the reference's adjustE (tilt adaptivity,
octree_omnigrid_solver3.cpp:1125-1166) and one adaptive advect(dt) (850-914
with enableTiltAdaptivity: adjustE(0.02, dt, EPSILON, 1000) first, 852-853),
starting from the inputs of tests/golden/adv_<scene>.npz.

Runs only where oracle/_ref/libastgref.so exists (the reference solver built
from its own sources by oracle/ref.mk). Each fixture is data only:
  * the thresholds and dt;
  * after post_project and adjustE alone: the tilt state adjustE writes
    (TiltENode e and flags, tilt_layout open/fixed) per layer node;
  * after post_project and one adaptive advect(dt): density, heat and
    velocity per node and the tilt state the solver's fields then use;
  * then the next step's projection on the new tilt state (constraint and
    ghosts, project(dt) to 1e-6 with rho = 2): velocity per node, and the
    PCG iteration count and residual of that solve.

    python tests/golden/make_golden_adjust.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402

SCENES = ["s32_two_spheres", "s32_closed_box_tilts", "s32_l4_open_tilts"]
D_THRESHOLD, V_THRESHOLD, MAX_THRESHOLD = 0.02, float(np.finfo(np.float64).eps), 1000.0


def solver(g, a):
    R, L = int(g["R"]), int(g["levels"])
    r = ref.RefSolver(g["split"], R, L, int(g["closed"]))
    if int(g["open_tilts"]):
        r.set_tilt_open(0, g["open_mask_l0"])
    for l in range(L):
        r.set_velocity(l, g["u_mac_out_%d" % l], g["u_tilt_out_%d" % l])
    for f in (0, 1):
        for l in range(L):
            r.set_scalar(f, l, a["s%d_in_oct_%d" % (f, l)], a["s%d_in_tilt_%d" % (f, l)])
    r.post_project()
    return r, L


def make(name):
    g = dict(np.load(os.path.join(HERE, name + ".npz")))
    a = dict(np.load(os.path.join(HERE, "adv_" + name + ".npz")))
    dt = float(a["dt"])
    out = {"scene": np.array(name), "dt": np.array(dt), "d_threshold": np.array(D_THRESHOLD),
           "v_threshold": np.array(V_THRESHOLD), "max_threshold": np.array(MAX_THRESHOLD)}
    r, L = solver(g, a)
    sizes = [r.level_size(l) for l in range(L)]
    r.adjust_e(D_THRESHOLD, dt, V_THRESHOLD, MAX_THRESHOLD)
    for l in range(L):
        e, bits, _ = r.tilt(l, which=1)
        out["adj_e_%d" % l], out["adj_bits_%d" % l] = e[:sizes[l]], bits[:sizes[l]]
    r.close()
    r, L = solver(g, a)
    r.set_tilt_adaptivity(True)
    r.advect(dt)
    for l in range(L):
        for f in (0, 1):
            o, t = r.scalar(f, l)
            out["s%d_step_oct_%d" % (f, l)], out["s%d_step_tilt_%d" % (f, l)] = o[:sizes[l]], t[:sizes[l]]
        m, t = r.velocity(l)
        out["vel_step_mac_%d" % l], out["vel_step_tilt_%d" % l] = m[:sizes[l]], t[:sizes[l]]
        e, bits, _ = r.tilt(l, which=0)
        out["step_e_%d" % l], out["step_bits_%d" % l] = e[:sizes[l]], bits[:sizes[l]]
    r.post_project()
    r.project(float(g["dt"]), tol=float(g["tol"]))
    for l in range(L):
        m, t = r.velocity(l)
        out["vel_proj_mac_%d" % l], out["vel_proj_tilt_%d" % l] = m[:sizes[l]], t[:sizes[l]]
    # the iteration count and residual that project() only logs, from the same
    # PCGSolver on the same system
    _, res, it = r.resolve(r.L.ref_matrix_n(r.h))
    out["proj_iters"], out["proj_residual"] = np.array(it), np.array(res)
    r.close()
    np.savez_compressed(os.path.join(HERE, "adj_" + name + ".npz"), **out)
    return out


if __name__ == "__main__":
    for s in SCENES:
        make(s)
        print("wrote adj_%s.npz" % s)
