"""Writes tests/golden/col_*.npz: scenes with the reference's sphere collider.

Runs only where oracle/_ref/libastgref.so exists (the reference solver built
from its own sources by oracle/ref.mk). The harness installs the collider the
reference's own initObstacle installs (a Sphere3 of radius 0.1 at
(0.5, 0.5, 0.45) through RigidBodyCollider3, octree_omnigrid_solver3.cpp:
152-157). Each fixture is data only:
  * the split map, the collider and its SDF grid (R^3 cell-centred values);
  * per level: tilt state, cutting-cell weights (initCuttingCellWeights with
    the collider, 1169-1238), unknown numbering (with the -3h cut, 292-347);
  * one projection of seeded N(0,1) face velocities (tolerance 1e-6): A, b,
    pressure, iterations, corrected velocity;
  * onUpdate's chain between project and advect (234-241) with the slip
    constraint (omni_boundary_solver3.h:26-100): velocity and dual grid. The
    MAC part runs the harness's serial form of the reference's loop (the
    reference's own loop updates faces in place while sampling neighbours on
    its thread pool, so its result depends on the pool's chunking); the
    parallel reference result is kept too, with the count of faces where the
    two differ;
  * point queries: boundary SDF, omni velocity, MAC, BackTrace (with the SDF
    clip, advection.h:49-59), density;
  * one advect(dt) (850-914) and wipeBoundary (1240-1292).

    python tests/golden/make_golden_collider.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ast-grid_amd"))

import astg  # noqa: E402  (split-map generator only)
from oracle import ref  # noqa: E402

CLOSED_XY = astg.WALL_XNEG | astg.WALL_XPOS | astg.WALL_YNEG | astg.WALL_YPOS
COLLIDER = (0.5, 0.5, 0.45, 0.1)  # the reference's initObstacle sphere

# name: (R, levels, refinement spheres, band, open level-0 tilts in the refinement spheres)
# The reference's tilt slip constraint (omni_boundary_solver3.h:26-50) writes
# a 3-vector at index k = 0..7; its build (no NDEBUG) stops on Eigen's index
# assertion as soon as an open tilt within h of the collider has a slot
# k >= 3 inside it. The scenes keep open tilts (T-joints included) away from
# the collider: in col_s32 it lies inside the finest level, in col_s32_l1
# inside level 1, with the refinement and the opened tilts elsewhere.
# col_s32_big widens the sphere to radius 0.25 around the same centre, so the
# -3h numbering cut removes cells (at radius 0.1 no cell centre is 3h deep).
SCENES = {
    "col_s32": (32, 3, [0.5, 0.5, 0.45, 0.16], [0.05, 0.12], False, 0.1),
    "col_s32_l1": (32, 3, [0.2, 0.2, 0.2, 0.05], [0.02, 0.75], True, 0.1),
    "col_s32_big": (32, 3, [0.5, 0.5, 0.45, 0.3], [0.05, 0.12], False, 0.25),
}
NPTS = 1000


def open_mask(xyz, R, spheres):
    pos = xyz.astype(np.float64) / R
    inside = np.zeros(len(xyz), bool)
    for s in np.asarray(spheres, np.float64).reshape(-1, 4):
        inside |= np.linalg.norm(pos - s[:3], axis=1) < s[3]
    return inside.astype(np.uint8)


def build(split, R, L, open_m, collider):
    r = ref.RefSolver(split, R, L, CLOSED_XY, collider=collider)
    if open_m is not None:
        r.set_tilt_open(0, open_m)
    return r


def make(name, R, L, spheres, band, open_tilts, radius, seed):
    split = astg.split_spheres(R, L, spheres, band)
    collider = COLLIDER[:3] + (radius,)
    out = {"R": R, "levels": L, "closed": CLOSED_XY, "split": split.astype(np.int8),
           "spheres": np.asarray(spheres, np.float64), "open_tilts": int(open_tilts),
           "collider": np.asarray(collider, np.float64)}
    r0 = ref.RefSolver(split, R, L, CLOSED_XY, collider=collider)
    open_m = None
    if open_tilts:
        xyz, _ = r0.nodes(0)
        open_m = open_mask(xyz[:r0.level_size(0, ghost=False)], R, spheres)
        out["open_mask_l0"] = open_m
    r0.close()
    r = build(split, R, L, open_m, collider)
    out["sdf"] = r.sdf()
    rng = np.random.default_rng(seed)
    sizes = [r.level_size(l) for l in range(L)]
    for l in range(L):
        e, bits, hd = r.tilt(l)
        wm, wt = r.weights(l)
        out["tilt_bits_%d" % l] = bits.astype(np.uint8)
        out["w_mac_%d" % l], out["w_tilt_%d" % l] = wm, wt
        mac = rng.standard_normal((sizes[l], 3)).astype(np.float32).astype(np.float64)
        til = rng.standard_normal((sizes[l], 8)).astype(np.float32).astype(np.float64)
        out["u_mac_in_%d" % l], out["u_tilt_in_%d" % l] = mac, til
        r.set_velocity(l, mac, til)
    dt, rho, tol = 0.5 / R, 2.0, 1e-6
    out["dt"], out["rho"], out["tol"] = dt, rho, tol
    r.project(dt, tol=tol, maxit=3000)
    for l in range(L):
        out["pid_oct_%d" % l], out["pid_tilt_%d" % l] = r.pindex(l)
        out["u_mac_out_%d" % l], out["u_tilt_out_%d" % l] = r.velocity(l)
    rowstart, col, val, b, p = r.system()
    out["A_rowstart"], out["A_col"], out["A_val"], out["b"], out["p"] = rowstart, col, val, b, p
    _, res, it = r.resolve(len(b))
    out["pcg_iters"], out["pcg_residual"] = it, res
    # scalars, then the chain after project (serial MAC constraint)
    for f in (0, 1):
        for l in range(L):
            o, t = rng.random(sizes[l]), rng.random(sizes[l])
            r.set_scalar(f, l, o, t)
            out["s%d_in_oct_%d" % (f, l)], out["s%d_in_tilt_%d" % (f, l)] = o, t
    r.post_project_serial()
    for l in range(L):
        out["vel_post_mac_%d" % l], out["vel_post_tilt_%d" % l] = r.velocity(l)
        out["dual_oct_%d" % l], out["dual_tilt_%d" % l] = r.dual(l)
    for f in (0, 1):
        r.scalar_update_ghost(f)
    pts = rng.random((NPTS, 3)) * 0.98 + 0.01
    # half of the points near the collider, where the clip and the slip act
    c = np.asarray(collider[:3])
    d = rng.standard_normal((NPTS // 2, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts[:NPTS // 2] = c + d * rng.uniform(0.5 * radius, 1.5 * radius, (NPTS // 2, 1))
    out["pts"] = pts
    out["q_sdf"] = r.sample_sdf(pts)
    out["q_vel"] = r.sample_points(0, pts)
    out["q_mac"] = r.sample_points(1, pts)
    out["q_bt"] = r.sample_points(2, pts, dt)
    out["q_den"] = r.sample_scalar(0, pts)
    r.advect(dt)
    for l in range(L):
        for f in (0, 1):
            o, t = r.scalar(f, l)
            out["s%d_adv_oct_%d" % (f, l)], out["s%d_adv_tilt_%d" % (f, l)] = o[:sizes[l]], t[:sizes[l]]
        m, t = r.velocity(l)
        out["vel_adv_mac_%d" % l], out["vel_adv_tilt_%d" % l] = m[:sizes[l]], t[:sizes[l]]
    r.wipe_boundary()
    for l in range(L):
        for f in (0, 1):
            o, t = r.scalar(f, l)
            out["s%d_wipe_oct_%d" % (f, l)], out["s%d_wipe_tilt_%d" % (f, l)] = o[:sizes[l]], t[:sizes[l]]
    r.close()
    # the reference's own (thread-pool) chain on the same projected state
    rp = build(split, R, L, open_m, collider)
    for l in range(L):
        rp.set_velocity(l, out["u_mac_out_%d" % l], out["u_tilt_out_%d" % l])
    rp.post_project()
    ndiff = 0
    for l in range(L):
        m, _ = rp.velocity(l)
        out["vel_post_mac_pool_%d" % l] = m
        ndiff += int((np.abs(m - out["vel_post_mac_%d" % l]) > 1e-12 * max(1.0, np.abs(m).max())).sum())
    out["pool_vs_serial_faces"] = ndiff
    out["pool_workers"] = int(ref.lib().ref_num_workers())
    rp.close()
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("%-12s n=%5d nnz=%6d iters=%3d pool/serial faces differing: %d (%d workers) -> %s (%d KB)" % (
        name, len(b), len(col), it, ndiff, out["pool_workers"], os.path.basename(path), os.path.getsize(path) // 1024))


if __name__ == "__main__":
    ref.lib().ref_quiet()
    for i, (name, args) in enumerate(SCENES.items()):
        make(name, *args, seed=2000 + i)
