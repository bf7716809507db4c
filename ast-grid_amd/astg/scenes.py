"""Seeded synthetic scenes of the benchmark configurations.

Every random quantity comes from numpy's PCG64 generator seeded explicitly, so
a scene is a pure function of (name, seed).

* ``projection_256``: 256^3-effective octree, 4 levels, finest cells inside 6
  random spheres (centres U[0.15,0.85]^3, radii U[0.05,0.15]) with graded bands
  around them; every non-fixed level-0 tilt inside a sphere opened with
  e = h0/2; face velocities N(0,1) sampled in FP32 on every MAC face and tilt
  slot; closed x and y walls, open z walls (the reference's default flags,
  octree_omnigrid_solver3.cpp:155-156); dt = 0.5/256, rho = 2.
* ``batch_scene`` / ``batch_systems``: the batched-solve workload: 24
  independent octree Poisson systems with leaf counts log-uniform in
  [5e4, 2e6]; scene k has 1-6 random spheres, 3-5 levels and R = 64, 128 or
  256 by its target count, its sphere radii and refinement bands scaled by one
  factor found by a secant search on log(leaves) until within 10% of the
  target; level-0 tilts inside the spheres opened; N(0,1) FP32 face velocities.
* ``full_512``: the full-scale size: 512^3-effective octree, 5 levels, the
  projection_256 spheres with radii x 1.2 and bands 1.2 * 24 * 2^l / 512
  (seed 0: 18.0 M leaves), level-0 tilts inside the spheres opened, N(0,1)
  FP32 face velocities; closed x and y walls, open z walls; dt = 0.5/512.
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


def velocity_frames(sizes, seed):
    """(mac (n,3), tilt (n,8)) per level for level node counts `sizes`: N(0,1)
    drawn in FP32 and widened to FP64, every MAC face and tilt slot."""
    rng = np.random.default_rng(seed + 7919)
    for nn in sizes:
        mac = rng.standard_normal((nn, 3), dtype=np.float32).astype(np.float64)
        til = rng.standard_normal((nn, 8), dtype=np.float32).astype(np.float64)
        yield mac, til


def fill_velocity(ctx, levels, seed):
    """N(0,1) face velocities drawn in FP32 (widened to FP64), every MAC face and tilt slot."""
    sizes = [ctx.level_info(l)[0] for l in range(levels)]
    for l, (mac, til) in enumerate(velocity_frames(sizes, seed)):
        ctx.set_velocity(l, mac, til)


def projection_256_split(seed=0, R=256, levels=4, band=BAND_256):
    """(split map, spheres) of projection_256."""
    spheres = random_spheres(seed)
    return split_spheres(R, levels, spheres.ravel(), band), spheres


def projection_256(device=0, seed=0, R=256, levels=4, band=BAND_256, velocities=True):
    split, spheres = projection_256_split(seed, R, levels, band)
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


def full_512(device=0, seed=0, scale=1.2):
    R, levels = 512, 5
    spheres = random_spheres(seed)
    spheres[:, 3] *= scale
    band = [scale * 24.0 * (1 << l) / R for l in range(levels - 1)]
    split = split_spheres(R, levels, spheres.ravel(), band)
    ctx = Context(device)
    ctx.set_topology(split, R, levels, CLOSED_XY)
    del split
    xyz, _ = ctx.nodes(0)
    nleaf = ctx.level_info(0)[1]
    ctx.set_tilt_open(0, inside_mask(xyz[:nleaf], R, spheres))
    del xyz
    fill_velocity(ctx, levels, seed)
    leaves = sum(ctx.level_info(l)[1] for l in range(levels))
    return ctx, {"R": R, "levels": levels, "leaves": leaves, "spheres": spheres, "dt": 0.5 / R, "rho": 2.0,
                 "seed": seed}


def leaf_count(split):
    """Leaves of a split map: a level-l leaf covers 8^l finest cells."""
    c = np.bincount(split)
    return int(round(sum(c[l] / 8.0 ** l for l in range(len(c)))))


def batch_targets(seed=0, count=24, lo=5e4, hi=2e6):
    rng = np.random.default_rng(seed + 104729)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count))


def batch_scene_params(seed, k, target):
    """(R, levels, spheres, band) of batch scene k, leaves within 10% of target."""
    rng = np.random.default_rng([seed, k, 7])
    R = 64 if target < 1.2e5 else (128 if target < 6e5 else 256)
    levels = int(rng.integers(3, 6))
    if R == 64:
        levels = min(levels, 4)
    m = int(rng.integers(1, 7))
    base = random_spheres(int(rng.integers(0, 2 ** 31)), count=m)

    def build(sc):
        sph = base.copy()
        sph[:, 3] *= sc
        band = [sc * 24.0 * (1 << l) / R for l in range(levels - 1)]
        split = split_spheres(R, levels, sph.ravel(), band)
        return sph, band, split

    # leaves grow about like scale^2 (refined shells): secant on the logs
    s0, s1 = 1.0, None
    sph, band, split = build(s0)
    l0 = leaf_count(split)
    s1 = s0 * (target / l0) ** 0.5
    for _ in range(8):
        if abs(np.log(l0 / target)) < np.log(1.1):
            break
        sph, band, split = build(s1)
        l1 = leaf_count(split)
        if abs(np.log(l1 / target)) < np.log(1.1) or l1 == l0:
            l0 = l1
            break
        slope = (np.log(l1) - np.log(l0)) / (np.log(s1) - np.log(s0))
        slope = min(max(slope, 0.5), 4.0)
        s0, l0 = s1, l1
        s1 = s0 * (target / l0) ** (1.0 / slope)
    return R, levels, sph, band, split


def batch_scene(seed, k, target, device=0):
    R, levels, sph, band, split = batch_scene_params(seed, k, target)
    ctx = Context(device)
    ctx.set_topology(split, R, levels, CLOSED_XY)
    xyz, _ = ctx.nodes(0)
    nleaf = ctx.level_info(0)[1]
    ctx.set_tilt_open(0, inside_mask(xyz[:nleaf], R, sph))
    fill_velocity(ctx, levels, seed * 1000 + k)
    leaves = sum(ctx.level_info(l)[1] for l in range(levels))
    return ctx, {"R": R, "levels": levels, "leaves": leaves, "spheres": len(sph), "dt": 0.5 / R, "rho": 2.0}


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


def diff_smoke(device=0, R=128, levels=3, seed=0):
    """The differentiable smoke configuration: the smoke topology at R^3
    effective (level 0 near the source, coarser outward), frozen; density 1
    in the source sphere (cells and vertices); the parameter is the MAC
    velocity u0 = 0.1 N(0,1) (seeded); the target is 3 Gaussian blobs
    (sigma 0.06, centres U[0.25, 0.75]^3) at the leaf cell centres. Closed x
    and y walls, open z walls, no collider; dt = 0.5/R, rho = 2."""
    split = smoke_64_split(R, levels)
    ctx = Context(device)
    ctx.set_topology(split, R, levels, CLOSED_XY)
    ctx.set_tilt_adaptivity(False)
    cx, cy, cz, r = SMOKE_SOURCE
    ctx.apply_sphere_source((cx, cy, cz), r, density=1.0, heat=0.0, speed=0.0, update_ghost=True)
    G, off = ctx.field_layout()
    rng = np.random.default_rng(seed + 31337)
    centres = rng.uniform(0.25, 0.75, size=(3, 3))
    target = np.zeros(G)
    leaf = np.zeros(G)
    for l in range(levels):
        xyz, _ = ctx.nodes(l)
        nleaf = ctx.level_info(l)[1]
        res = R >> l
        h = 1.0 / res
        pos = (xyz[:nleaf].astype(np.float64) + 0.5) * h
        valid = np.all(xyz[:nleaf] < res - 1, axis=1)
        t = np.zeros(nleaf)
        for c in centres:
            t += np.exp(-np.sum((pos - c) ** 2, axis=1) / (2 * 0.06 ** 2))
        target[off[l]:off[l] + nleaf] = t * valid
        leaf[off[l]:off[l] + nleaf] = valid
    u0 = 0.1 * np.random.default_rng(seed + 4242).standard_normal(3 * G)
    info = {"R": R, "levels": levels, "dt": 0.5 / R, "rho": 2.0, "G": G, "target": target, "leaf_cells": leaf,
            "u0": u0, "centres": centres, "leaves": int(sum(ctx.level_info(l)[1] for l in range(levels)))}
    return ctx, info
