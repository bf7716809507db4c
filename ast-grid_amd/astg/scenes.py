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
* ``full_scale_run``: the adaptive full-scale configuration at a chosen size:
  two sources r=0.05 at (0.3,0.5,0.1) and (0.7,0.5,0.1) with lateral inflow
  +-0.5 in x, buoyancy along +z, tilt adaptivity, the layout rebuilt from the
  density field every few steps (fields carried over by the device samplers).
* ``smoke_64``: the plain smoke configuration: 64^3, 3 levels, one source
  sphere r=0.08 rising along +z, the reference's buoyancy axis
  (addForce, octree_omnigrid_solver3.cpp:917-972): centre (0.5,0.5,0.12);
  level 0 within 4 finest cells of the source, one level coarser every 8 cells
  of that level further out; closed x and y walls, open z walls; fields start
  at rest; every step applies the source (density 1, inflow 1) and a buoyancy
  of 1.0 x density (no heat term); dt = 0.5/64, rho = 2, tilt adaptivity on.
"""
import numpy as np

from . import Context, split_distance, split_spheres, WALL_XNEG, WALL_XPOS, WALL_YNEG, WALL_YPOS, WALL_ZNEG, WALL_ZPOS

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


# ---- the full-scale adaptive configuration ---------------------------------
# Two sources of radius 0.05 near the floor with lateral inflow toward the
# middle (+-0.5 in x), buoyancy along +z (the reference's axis; addForce
# 917-972), closed x and y walls, open z walls, tilt adaptivity on. The layout
# is rebuilt from the density field every `rebuild` steps: a build-only
# extension (the reference builds its layout once, initData 68).
FULL_SOURCES = ((0.3, 0.5, 0.1, 0.05, (0.5, 0.0, 0.0)), (0.7, 0.5, 0.1, 0.05, (-0.5, 0.0, 0.0)))

# HALF_TILT_SLOT_INDEX rows (constants.h): the 4 tilt slots and the face slot
# of a half-tilt for each half direction
_HALF_SLOT = ((0, 3, 4, 7, 1), (0, 1, 4, 5, 2), (0, 1, 2, 3, 5), (1, 2, 5, 6, 0), (2, 3, 6, 7, 4), (4, 5, 6, 7, 3))
_TB_HALF = 32


# Refinement bands of the full-scale layouts, in domain units: finest within
# 0.2 of the smoke, one level coarser at 0.3, 0.45 and 0.7. The same layout at
# every R: a reduced run is the full-scale configuration at a coarser
# resolution. At 512^3 and 5 levels the initial layout (the two sources) has
# 15.9 M leaves.
FULL_BANDS = (0.2, 0.3, 0.45, 0.7)


def _bands(R, levels):
    return list(FULL_BANDS[:levels - 1])


def full_scale_initial_split(R, levels):
    sph = np.array([[cx, cy, cz, r] for cx, cy, cz, r, _ in FULL_SOURCES])
    return split_spheres(R, levels, sph.ravel(), _bands(R, levels))


def _tilt_dirs():
    s = 1.0 / np.sqrt(3.0)
    k = np.arange(8)
    k4 = k & 3
    up = np.where(k < 4, 1.0, -1.0)
    return np.stack([np.where((k4 == 0) | (k4 == 3), up, -up) * s, np.where(k4 < 2, up, -up) * s,
                     np.full(8, s)], axis=1)


def _tilt_offsets():
    """positionTilt's offset signs (octree_omni_tilt_grid3.h:33-40), per slot."""
    k = np.arange(8)
    k4 = k & 3
    return np.stack([np.where((k4 == 0) | (k4 == 3), 1.0, -1.0), np.where(k4 < 2, 1.0, -1.0),
                     np.where(k < 4, 1.0, -1.0)], axis=1)


def _sample(ctx, what, pts):
    v = ctx.sample_points(what, pts)
    if not np.isfinite(v).all():
        raise RuntimeError("sampled field holds %d non-finite values" % int((~np.isfinite(v)).sum()))
    return v


def transfer_fields(old, new, R, levels):
    """Carries velocity, density and heat from `old` to `new` (another layout of
    the same domain) by sampling the old context's samplers at the new
    context's degrees of freedom: MAC faces (the omni velocity's axis
    component), tilt slots (its projection on the slot direction, at the slot
    position), half-tilt face slots, cell centres and vertices."""
    import astg
    dirs, offs = _tilt_dirs(), _tilt_offsets()
    for l in range(levels):
        xyz, _ = new.nodes(l)
        nn = len(xyz)
        h = 1.0 / (R >> l)
        base = xyz.astype(np.float64) * h
        mac = np.empty((nn, 3))
        for a in range(3):
            p = base + 0.5 * h
            p[:, a] = base[:, a]
            mac[:, a] = _sample(old, astg.SAMPLE_VELOCITY, p)[:, a]
        e, bits, hd = new.tilt(l)
        til = np.empty((nn, 8))
        for k in range(8):
            p = base + offs[k][None, :] * (e[:, None] / 3.0)
            til[:, k] = _sample(old, astg.SAMPLE_VELOCITY, p) @ dirs[k]
        half = np.flatnonzero((bits & _TB_HALF) != 0)
        if len(half):
            # a half-tilt keeps its 4 tilt slots and its face slot (the velocity's
            # component along the face's axis); its other slots are zero
            vv = _sample(old, astg.SAMPLE_VELOCITY, base[half])
            d = hd[half].astype(np.int64)
            rows = np.flatnonzero(d >= 0)
            slots = np.asarray(_HALF_SLOT)[d[rows]]
            kf = slots[:, 4]
            ax = np.where(kf <= 1, 0, np.where((kf == 2) | (kf == 4), 1, 2))
            sub = til[half]
            keep = np.zeros((len(half), 8), bool)
            keep[rows[:, None], slots[:, :4]] = True
            keep[rows, kf] = True
            sub[rows, kf] = vv[rows, ax]
            sub[~keep] = 0.0
            til[half] = sub
        new.set_velocity(l, mac, til)
        for f, what in ((astg.FIELD_DENSITY, astg.SAMPLE_DENSITY), (astg.FIELD_HEAT, astg.SAMPLE_HEAT)):
            new.set_scalar(f, l, _sample(old, what, base + 0.5 * h), _sample(old, what, base))
    new.scalar_update_ghost(astg.FIELD_DENSITY)
    new.scalar_update_ghost(astg.FIELD_HEAT)


def density_split(ctx, R, levels, threshold=0.01, lattice=256):
    """A new split map from the density field: finest where the density at a
    cell centre of a lattice at most `lattice` cells wide (the finest cells,
    or blocks of 2^s of them above that) exceeds `threshold` or a source sits,
    graded outward with the smoke layouts' bands (astg_split_distance)."""
    import astg
    from scipy import ndimage
    shift = 0
    while (R >> shift) > lattice and shift + 1 < levels:
        shift += 1
    n = R >> shift
    c = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    del X, Y, Z
    d = _sample(ctx, astg.SAMPLE_DENSITY, pts)
    region = d > threshold
    for cx, cy, cz, r, _ in FULL_SOURCES:
        region |= ((pts[:, 0] - cx) ** 2 + (pts[:, 1] - cy) ** 2 + (pts[:, 2] - cz) ** 2) < r * r
    del pts
    region = region.reshape(n, n, n)
    dist = ndimage.distance_transform_edt(~region) / n  # z, y, x order: x fastest in ravel
    return split_distance(R, levels, dist.astype(np.float32).ravel(), _bands(R, levels), shift)


def full_scale_run(R=128, levels=4, steps=100, rebuild=10, precision=0, tol=1e-6, lattice=32, device=0,
                   splits=None, log=None):
    """The adaptive full-scale configuration: `steps` steps from rest, each the
    two sources, addForce and onUpdate's chain (projection in `precision` to
    `tol`, tilt adaptivity), the layout rebuilt from the density every
    `rebuild` steps (density_split), or from `splits[i]` at the i-th rebuild
    when given, so that a second run can replay the first one's layouts.
    Returns per-step wall and projection times, rebuild times, PCG
    iterations, unknowns, the split maps used and, every `rebuild` steps
    (before the rebuild), the density, heat and velocity sampled at the
    centres of a lattice^3 grid."""
    import time
    import astg
    import torch
    dt, rho = 0.5 / R, 2.0
    prm = astg.default_params(tol=tol, max_iterations=3000, precision=precision)
    split = full_scale_initial_split(R, levels)
    ctx = Context(device)
    ctx.set_topology(split, R, levels, CLOSED_XY)
    c = (np.arange(lattice) + 0.5) / lattice
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    probe = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    out = {"step_ms": [], "project_ms": [], "rebuild_ms": [], "iterations": [], "snapshots": [], "unknowns": [],
           "splits": []}
    first = True
    for k in range(1, steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for cx, cy, cz, r, v in FULL_SOURCES:
            ctx.apply_sphere_source((cx, cy, cz), r, 1.0, 1.0, update_ghost=first, velocity=v)
        first = False
        ctx.add_force(dt, 1.0, 0.0)
        ctx.set_tilt_adaptivity(True)
        res = ctx.step(dt, rho, prm)
        torch.cuda.synchronize()
        out["step_ms"].append((time.perf_counter() - t0) * 1e3)
        out["project_ms"].append(res.ms_total)
        out["iterations"].append(res.iterations)
        out["unknowns"].append(res.n)
        if k % rebuild == 0:
            out["snapshots"].append({"step": k, "density": _sample(ctx, astg.SAMPLE_DENSITY, probe),
                                     "heat": _sample(ctx, astg.SAMPLE_HEAT, probe),
                                     "velocity": _sample(ctx, astg.SAMPLE_VELOCITY, probe)})
            if k < steps:
                t0 = time.perf_counter()
                i = len(out["splits"])
                nsplit = splits[i] if splits is not None else density_split(ctx, R, levels)
                out["splits"].append(np.asarray(nsplit, np.uint8))  # levels < 16: a byte per cell
                new = Context(device)
                new.set_topology(nsplit, R, levels, CLOSED_XY)
                transfer_fields(ctx, new, R, levels)
                ctx.close()
                ctx = new
                torch.cuda.synchronize()
                out["rebuild_ms"].append((time.perf_counter() - t0) * 1e3)
                if log:
                    log("step %d: layout rebuilt (%.0f ms), %d unknowns before" % (k, out["rebuild_ms"][-1], res.n))
    ctx.close()
    return out


def full_scale_compare(R=128, levels=4, steps=100, rebuild=10, tol=1e-6, lattice=32, device=0, log=None):
    """The full-scale run in FP64, then in mixed precision on the FP64 run's
    layouts. Per snapshot and field: the mixed run's relative L2 error and
    largest absolute difference against the FP64 run at the lattice points."""
    import astg
    ref = full_scale_run(R, levels, steps, rebuild, astg.PREC_FP64, tol, lattice, device, log=log)
    mix = full_scale_run(R, levels, steps, rebuild, astg.PREC_MIXED, tol, lattice, device, splits=ref["splits"],
                         log=log)
    errors = []
    for a, b in zip(mix["snapshots"], ref["snapshots"]):
        e = {"step": b["step"]}
        for f in ("density", "heat", "velocity"):
            d = a[f] - b[f]
            e[f] = {"rel_l2": float(np.linalg.norm(d) / max(np.linalg.norm(b[f]), 1e-300)),
                    "max_abs": float(np.abs(d).max())}
        errors.append(e)
    return {"fp64": ref, "mixed": mix, "errors": errors}
