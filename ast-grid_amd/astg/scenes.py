"""Seeded synthetic scenes of the benchmark configurations.

Every random quantity comes from numpy's PCG64 generator seeded explicitly, so
a scene is a pure function of (name, seed).

* ``projection_256``: 256^3-effective octree, 4 levels, finest cells inside 6
  random spheres (centres U[0.15,0.85]^3, radii U[0.05,0.15]) with graded bands
  around them; every non-fixed level-0 tilt inside a sphere opened with
  e = h0/2; face velocities N(0,1) sampled in FP32 on every MAC face and tilt
  slot; closed x and y walls, open z walls (the reference's default flags,
  octree_omnigrid_solver3.cpp:155-156); dt = 0.5/256, rho = 2.
* ``smoke_64``: the plain smoke configuration: 64^3, 3 levels, one source
  sphere r=0.08 rising along +z, the reference's buoyancy axis
  (addForce, octree_omnigrid_solver3.cpp:917-972): centre (0.5,0.5,0.12);
  level 0 within 4 finest cells of the source, one level coarser every 8 cells
  of that level further out; closed x and y walls, open z walls; fields start
  at rest; every step applies the source (density 1, inflow 1) and a buoyancy
  of 1.0 x density (no heat term); dt = 0.5/64, rho = 2, tilt adaptivity on.
"""
import numpy as np

from . import Context, split_spheres, WALL_XNEG, WALL_XPOS, WALL_YNEG, WALL_YPOS, WALL_ZNEG, WALL_ZPOS

CLOSED_XY = WALL_XNEG | WALL_XPOS | WALL_YNEG | WALL_YPOS

# Band distances (absolute, unit cube) for refining level l+1 -> l, tuned once so
# that seed 0 of projection_256 gives 3.0M +- 10% leaves, then frozen.
BAND_256 = (0.092, 0.17, 0.32)


def random_spheres(seed, count=6, lo=0.15, hi=0.85, rmin=0.05, rmax=0.15):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, size=(count, 3))
    r = rng.uniform(rmin, rmax, size=(count, 1))
    return np.hstack([c, r])


def inside_mask(xyz, res, spheres):
    pos = xyz.astype(np.float64) / res
    m = np.zeros(len(xyz), bool)
    for s in spheres:
        m |= np.linalg.norm(pos - s[:3], axis=1) < s[3]
    return m.astype(np.uint8)


def fill_velocity(ctx, levels, seed):
    """N(0,1) face velocities drawn in FP32 (widened to FP64), every MAC face and tilt slot."""
    rng = np.random.default_rng(seed + 7919)
    for l in range(levels):
        nn, _ = ctx.level_info(l)
        mac = rng.standard_normal((nn, 3), dtype=np.float32).astype(np.float64)
        til = rng.standard_normal((nn, 8), dtype=np.float32).astype(np.float64)
        ctx.set_velocity(l, mac, til)


def projection_256(device=0, seed=0, R=256, levels=4, band=BAND_256, velocities=True):
    spheres = random_spheres(seed)
    split = split_spheres(R, levels, spheres.ravel(), band)
    ctx = Context(device)
    ctx.set_topology(split, R, levels, CLOSED_XY)
    xyz, _ = ctx.nodes(0)
    nleaf = ctx.level_info(0)[1]
    ctx.set_tilt_open(0, inside_mask(xyz[:nleaf], R, spheres))
    if velocities:
        fill_velocity(ctx, levels, seed)
    leaves = sum(ctx.level_info(l)[1] for l in range(levels))
    info = {"R": R, "levels": levels, "leaves": leaves, "spheres": spheres, "dt": 0.5 / R, "rho": 2.0,
            "seed": seed}
    return ctx, info


SMOKE_SOURCE = (0.5, 0.5, 0.12, 0.08)


def smoke_64_split(R=64, levels=3):
    h0 = 1.0 / R
    band = [4 * h0]
    for l in range(1, levels - 1):
        band.append(band[-1] + 8 * h0 * (1 << l))
    return split_spheres(R, levels, np.array([SMOKE_SOURCE]).ravel(), band)


def smoke_64(device=0, R=64, levels=3):
    split = smoke_64_split(R, levels)
    ctx = Context(device)
    ctx.set_topology(split, R, levels, CLOSED_XY)
    info = {"R": R, "levels": levels, "dt": 0.5 / R, "rho": 2.0, "split": split, "closed": CLOSED_XY,
            "source": SMOKE_SOURCE, "buoyancy": (1.0, 0.0)}
    return ctx, info


def smoke_step(ctx, info, first, params):
    """One step of the smoke configuration: sceneSetup (the source), addForce,
    then onUpdate's constrain / project / advect chain (with tilt adaptivity)."""
    cx, cy, cz, r = info["source"]
    ctx.apply_sphere_source((cx, cy, cz), r, 1.0, 1.0, 1.0, update_ghost=first)
    ctx.add_force(info["dt"], *info["buoyancy"])
    ctx.set_tilt_adaptivity(True)
    return ctx.step(info["dt"], info["rho"], params)
