"""Benchmark: pressure projection on a 256^3-effective AST octree, FP64 vs mixed.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--seed 0] [--no-cpu]

A step is one full projection (assembly of A and b, MIC(0) factorisation,
PCG to 1e-6 relative residual in the reference's inf-norm test, velocity
correction) of the same synthetic N(0,1) face-velocity frame, restored on the
device before each step. Topology tables and the level schedule are built once
(reported separately as setup). Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ast-grid_amd"))

MI355X_HBM_GBS = 8000.0

KERNEL_NAMES = ("spmv64", "fwd64", "bwd64", "mic", "assemble", "spmv32", "fwd32", "bwd32", "copy")
KERNELS = tuple((nm, i) for i, nm in enumerate(KERNEL_NAMES))

# HBM bytes per launch from rocprofv3 counter runs (FETCH_SIZE + WRITE_SIZE,
# scaled by the copy-kernel calibration) of `bench.py --kernels-only` on this
# workload, as summarised in profiles/traffic_projection_256.md. Constants from
# that file, not measured by this run.
TRAFFIC_SOURCE = "profiles/traffic_projection_256.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)"
TRAFFIC = {"spmv64": 760900000, "fwd64": 629600000, "bwd64": 684900000, "spmv32": 554000000,
           "fwd32": 344800000, "bwd32": 349800000, "copy": 56200000, "mic": 7371200000, "mic_mixed": 6413100000}


def algorithmic_bytes(name, n, nnz_a):
    """Compulsory HBM bytes of one launch, each array streamed once, 4-byte
    indices, value width w: SpMV (w+4) nnz_A + (2w+4) n; forward sweep
    (w+4) nnz_L + (3w+4) n; backward sweep with z.r (w+4) nnz_L + (4w+4) n;
    MIC(0) factorisation (FP64) 12 (nnz_A + nnz_L) + 16 n; copy 16 n. nnz_L is
    the strict lower triangle, (nnz_A - n) / 2. None where no such figure is
    defined (assembly: its own table count is used)."""
    nnz_l = (nnz_a - n) / 2
    w = 4 if name.endswith("32") else 8
    if name.startswith("spmv"):
        return (w + 4) * nnz_a + (2 * w + 4) * n
    if name.startswith("fwd"):
        return (w + 4) * nnz_l + (3 * w + 4) * n
    if name.startswith("bwd"):
        return (w + 4) * nnz_l + (4 * w + 4) * n
    if name == "mic":
        return 12 * (nnz_a + nnz_l) + 16 * n
    if name == "copy":
        return 16 * n
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(ctx, info, prm, steps, warmup):
    import torch
    for _ in range(warmup):
        ctx.snapshot(restore=True)
        ctx.project(info["dt"], info["rho"], prm)
    torch.cuda.synchronize()
    results = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.snapshot(restore=True)
        results.append(ctx.project(info["dt"], info["rho"], prm))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps, results


def advect_timing(ctx, info, prm, reps):
    """One projected frame, then the post-projection chain (constrain, ghosts,
    dual grid) and one advect step of density, heat and velocity, each timed
    over `reps` repetitions (device work bracketed by synchronisation)."""
    import numpy as np
    import torch
    import astg
    L = info["levels"]
    rng = np.random.default_rng(1)
    dofs = 0
    for l in range(L):
        nn, nleaf = ctx.level_info(l)
        for f in (astg.FIELD_DENSITY, astg.FIELD_HEAT):
            ctx.set_scalar(f, l, rng.random(nn), rng.random(nn))
        _, bits, _ = ctx.tilt(l)
        leaf_bits = bits[:nleaf]
        dofs += nleaf * (1 + 3) + int(((leaf_bits & 64) != 0).sum()) + 8 * int(((leaf_bits & 2) == 0).sum())
    nh = len(ctx.half_tilts())
    dofs += nh * (1 + 5)
    ctx.snapshot(restore=True)
    ctx.project(info["dt"], info["rho"], prm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.post_project()
    torch.cuda.synchronize()
    t_post = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.advect(info["dt"])
    torch.cuda.synchronize()
    t_adv = (time.perf_counter() - t0) / reps
    return {"ms_post_project": round(t_post * 1e3, 3), "ms_advect": round(t_adv * 1e3, 3),
            "backtraces": dofs, "backtraces_per_s": round(dofs / t_adv, 1),
            "note": "density+heat on octagon/open-tilt/half-tilt DOFs, velocity on MAC/tilt/half-tilt slots; "
                    "2 velocity samples per back-trace"}


def smoke_timing(steps=50):
    """The plain smoke configuration (64^3, 3 levels, rising source, buoyancy,
    tilt adaptivity): `steps` full steps from rest, each sceneSetup + addForce
    + onUpdate's chain (projection to 1e-6 in FP64, advection); wall time per
    step including the tilt updates (host retopology)."""
    import torch
    import astg
    from astg import scenes
    ctx, info = scenes.smoke_64(device=0)
    prm = astg.default_params(tol=1e-6, max_iterations=3000)
    torch.cuda.synchronize()
    iters = []
    t0 = time.perf_counter()
    for k in range(steps):
        iters.append(scenes.smoke_step(ctx, info, k == 0, prm).iterations)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    open_tilts = 0
    for l in range(info["levels"]):
        _, bits, _ = ctx.tilt(l)
        open_tilts += int(((bits[:ctx.level_info(l)[1]] & 64) != 0).sum())
    ctx.close()
    return {"ms_per_step": round(ms, 3), "steps": steps, "pcg_iterations_mean": round(sum(iters) / len(iters), 1),
            "open_tilts_at_end": open_tilts,
            "note": "64^3, 3 levels; source + addForce + constrain/project/post chain/adjustE/advect per step"}


# Launch shape of the batched solves, frozen from `bench.py --batch-only --tune`
# on seed 0 (the autotuner's choice; profiles/batched_24.md).
BATCH_TUNING = {"merge_levels": 1, "stream_blocks_per_cu": 8, "lanes_per_row": 4}


def batched_timing(seed=0, count=24, tune=False):
    """Batched solves: `count` independent octree systems (leaf counts
    log-uniform in [5e4, 2e6]) solved by one astg_pcg_solve_batch call, FP64
    and mixed, against the same systems solved one after another (each its own
    factor + solve through the single-system path, in the reference numbering
    the batch keeps; the mixed form's block-colour single solves are timed
    too). Device times (factor + solve), second call of each (the first builds
    the stacked pattern)."""
    import numpy as np
    import astg
    from astg import scenes
    # single solves in the reference numbering, the ordering the batch keeps
    prm = {p: astg.default_params(tol=1e-6, max_iterations=3000, precision=p, ordering=astg.ORDER_REFERENCE)
           for p in (astg.PREC_FP64, astg.PREC_MIXED)}
    prm_blk = astg.default_params(tol=1e-6, max_iterations=3000, precision=astg.PREC_MIXED)
    t0 = time.perf_counter()
    systems, leaves, seq, single_its = [], [], {p: 0.0 for p in prm}, {p: [] for p in prm}
    seq_blk = 0.0
    for k, target in enumerate(scenes.batch_targets(seed, count)):
        ctx, info = scenes.batch_scene(seed, k, target)
        ctx.snapshot(restore=False)
        for p in prm:
            ctx.snapshot(restore=True)
            ctx.project(info["dt"], info["rho"], prm[p])  # first call: pattern and buffers
            ctx.snapshot(restore=True)
            r = ctx.project(info["dt"], info["rho"], prm[p])
            seq[p] += r.ms_factor + r.ms_solve
            single_its[p].append(r.iterations)
        for _ in range(2):  # the mixed form's default ordering (block colours), second call timed
            ctx.snapshot(restore=True)
            rb = ctx.project(info["dt"], info["rho"], prm_blk)
        seq_blk += rb.ms_factor + rb.ms_solve
        ctx.snapshot(restore=True)
        r = ctx.project(info["dt"], info["rho"], prm[astg.PREC_FP64])  # the system handed to the batch
        rs, col, val, b, _ = ctx.system(r.n, r.nnz)
        systems.append((rs, col, val, b))
        leaves.append(info["leaves"])
        ctx.close()
    build_s = time.perf_counter() - t0
    c = astg.Context(0)
    out = {"systems": count, "seed": seed, "leaves": leaves, "leaves_total": int(sum(leaves)),
           "n_total": int(sum(len(s[3]) for s in systems)), "nnz_total": int(sum(len(s[1]) for s in systems)),
           "scene_build_s": round(build_s, 1)}
    if tune:
        t1 = time.perf_counter()
        best = c.tune_batch(systems, prm[astg.PREC_MIXED], iterations=20)
        out["tuner"] = dict(best.as_dict(), seconds=round(time.perf_counter() - t1, 1))
        log("tuner: %s" % out["tuner"])
    else:
        c.set_batch_tuning(**BATCH_TUNING)
        out["tuning"] = dict(BATCH_TUNING)
    for p, name in ((astg.PREC_FP64, "fp64"), (astg.PREC_MIXED, "mixed")):
        t1 = time.perf_counter()
        c.pcg_solve_batch(systems, prm[p])
        first_s = time.perf_counter() - t1
        xs, res = c.pcg_solve_batch(systems, prm[p])
        ms = res[0].ms_factor + res[0].ms_solve
        its = [r.iterations for r in res]
        out[name] = {"ms_batch": round(ms, 3), "ms_sequential": round(seq[p], 3),
                     "speedup_vs_sequential": round(seq[p] / ms, 3), "iterations": its,
                     "single_iterations": single_its[p],
                     "max_iteration_diff": int(max(abs(a - b) for a, b in zip(its, single_its[p]))),
                     "all_converged": bool(all(r.converged for r in res)), "depth": res[0].depth,
                     "first_call_s": round(first_s, 2)}
    out["mixed"]["ms_sequential_block_colour"] = round(seq_blk, 3)
    c.close()
    return out


def projected_fields(ctx, res, levels):
    """Pressure and per-level (MAC, tilt) velocities after a projection."""
    _, _, _, _, p = ctx.system(res.n, res.nnz)
    return {"p": p, "u": [ctx.velocity(l) for l in range(levels)]}


def field_errors(fmx, f64):
    """Relative L2 error of the mixed form's fields against the FP64 form's."""
    import numpy as np
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    cat = lambda f, i: np.concatenate([u[i].ravel() for u in f["u"]])
    return {"pressure_rel_l2": rel(fmx["p"], f64["p"]),
            "mac_velocity_rel_l2": rel(cat(fmx, 0), cat(f64, 0)),
            "tilt_velocity_rel_l2": rel(cat(fmx, 1), cat(f64, 1))}


def full_scale_timing(reps=2):
    """The full-scale size (scenes.full_512: 512^3-effective, 5 levels, 18 M
    leaves): set-up, one projection in FP64 and one mixed (best of `reps`
    after a first call that builds the solver pattern), the mixed result's
    error against FP64 per field, the post-projection chain and one advect."""
    import numpy as np
    import torch
    import astg
    from astg import scenes
    t0 = time.perf_counter()
    ctx, info = scenes.full_512(device=0)
    scene_s = time.perf_counter() - t0
    log("full_512: scene %.1f s, %d leaves" % (scene_s, info["leaves"]))
    L = info["levels"]
    out = {"workload": "full_512: 512^3-effective AST octree, 5 levels, %d leaves" % info["leaves"],
           "scene_s": round(scene_s, 1)}
    ctx.snapshot(restore=False)
    fields = {}
    for p, name in ((astg.PREC_FP64, "fp64"), (astg.PREC_MIXED, "mixed")):
        prm = astg.default_params(tol=1e-6, max_iterations=3000, precision=p)
        best = None
        for k in range(reps + 1):
            ctx.snapshot(restore=True)
            t1 = time.perf_counter()
            r = ctx.project(info["dt"], info["rho"], prm)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t1
            if k == 0:
                out["first_call_s_" + name] = round(wall, 2)
            elif best is None or r.ms_total < best.ms_total:
                best = r
        log("full_512 %s: %.1f ms, %d iterations" % (name, best.ms_total, best.iterations))
        out[name] = {"ms_per_step": round(best.ms_total, 2), "iterations": best.iterations, "outer": best.outer,
                     "ms_factor": round(best.ms_factor, 2), "ms_solve": round(best.ms_solve, 2),
                     "rel_residual": best.rel_residual}
        out["n"], out["nnz"], out["depth"] = best.n, best.nnz, best.depth
        fields[name] = projected_fields(ctx, best, L)
    out["mixed_vs_fp64"] = field_errors(fields["mixed"], fields["fp64"])
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.post_project()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ctx.advect(info["dt"])
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out["ms_post_project"] = round((t2 - t1) * 1e3, 2)
    out["ms_advect"] = round((t3 - t2) * 1e3, 2)
    ctx.close()
    return out


def c5_timing(R=512, levels=5, steps=20, rebuild=10, lattice=32):
    """The full-scale adaptive run (scenes.full_scale_compare): `steps` steps
    from rest in FP64, then in mixed precision on the FP64 run's layouts, the
    layout rebuilt from the density every `rebuild` steps; per-step times and
    the mixed run's per-field error against FP64 at a lattice^3 grid every
    `rebuild` steps."""
    import numpy as np
    from astg import scenes
    t0 = time.perf_counter()
    cmp = scenes.full_scale_compare(R=R, levels=levels, steps=steps, rebuild=rebuild, lattice=lattice,
                                    log=lambda m: log("c5 " + m))
    wall = time.perf_counter() - t0
    out = {"workload": "full_scale: %d^3-effective AST octree, %d levels, two r=0.05 sources with +-0.5 lateral "
                       "inflow, tilt adaptivity on, layout rebuilt from the density every %d steps" % (R, levels, rebuild),
           "steps": steps, "rebuild_every": rebuild, "wall_s": round(wall, 1),
           "leaves_initial": int(scenes.leaf_count(scenes.full_scale_initial_split(R, levels))),
           "leaves_rebuilt": [int(scenes.leaf_count(s.astype(np.int32))) for s in cmp["fp64"]["splits"]]}
    for name in ("fp64", "mixed"):
        o = cmp[name]
        out[name] = {"ms_per_step": round(float(np.mean(o["step_ms"])), 2),
                     "ms_projection": round(float(np.mean(o["project_ms"])), 2),
                     "ms_rebuild": round(float(np.mean(o["rebuild_ms"])), 1) if o["rebuild_ms"] else None,
                     "iterations": [int(min(o["iterations"])), int(max(o["iterations"]))],
                     "unknowns": [int(min(o["unknowns"])), int(max(o["unknowns"]))]}
    out["mixed_vs_fp64"] = [{"step": e["step"], **{f: {k: float("%.3e" % v) for k, v in e[f].items()}
                                                   for f in ("density", "heat", "velocity")}} for e in cmp["errors"]]
    return out


def cpu_smoke_baseline(steps=50):
    """The reference's own onUpdate (oracle/_ref) on the smoke configuration,
    the same `steps` steps as smoke_timing, on the host's cores."""
    from oracle import ref
    from astg import scenes
    if not ref.available():
        return None
    R, L = 64, 3
    cx, cy, cz, rad = scenes.SMOKE_SOURCE
    r = ref.RefSolver(scenes.smoke_64_split(R, L), R, L, scenes.CLOSED_XY)
    r.set_source((cx, cy, cz), rad, 1.0, 1.0, 1.0)
    r.set_buoyancy(1.0, 0.0)
    r.set_tilt_adaptivity(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step(0.5 / R, tol=1e-6)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    r.close()
    return {"ms_per_step": round(ms, 2), "steps": steps, "cores": int(ref.lib().ref_num_workers()),
            "kind": "reference"}


def host_cpu():
    """(CPU model, logical CPUs in this process's affinity mask)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return model, n


def cpu_baseline(ctx, info, res64, seed):
    """The reference's own project() (octree_omnigrid_solver3.cpp:350-679,
    compiled from its sources into oracle/_ref) on the same 256^3 scene: the
    split map, opened level-0 tilts and FP32-drawn face velocities of the GPU
    run. One call, timed whole (numbering, assembly, MIC(0) factor, PCG to 1e-6,
    velocity update) and through the reference's own BENCHMARK sections.
    Without the reference library, the oracle's port of its MIC(0)-PCG on the
    device-assembled system, timed once to convergence (solve only)."""
    from oracle import ref
    from astg import scenes
    model, nproc = host_cpu()
    R, L = info["R"], info["levels"]
    n, nnz = res64.n, res64.nnz
    if ref.available():
        t0 = time.perf_counter()
        split, spheres = scenes.projection_256_split(seed, R, L)
        r = ref.RefSolver(split, R, L, scenes.CLOSED_XY)
        del split
        xyz, _ = r.nodes(0)
        nl = r.level_size(0, ghost=False)
        r.set_tilt_open(0, scenes.inside_mask(xyz[:nl], R, spheres))
        sizes = [r.level_size(l, ghost=True) for l in range(L)]
        assert sizes == [ctx.level_info(l)[0] for l in range(L)], "reference scene differs from the GPU scene"
        for l, (mac, til) in enumerate(scenes.velocity_frames(sizes, seed)):
            r.set_velocity(l, mac, til)
        build_s = time.perf_counter() - t0
        ref.benchmark_reset()
        t0 = time.perf_counter()
        r.project(info["dt"], tol=1e-6, maxit=3000)
        wall = time.perf_counter() - t0
        sections = {k: round(v[0] * 1e3, 1) for k, v in ref.benchmark_sections().items()}
        r.close()
        cores = int(ref.lib().ref_num_workers())
        return {"value": round(wall * 1e3, 1), "unit": "ms/step", "cores": cores, "kind": "reference",
                "sample": "one whole project() call of the reference on the same 256^3 scene (n=%d, nnz=%d): "
                          "numbering, assembly, MIC(0) factor, PCG to 1e-6, velocity update; the reference's "
                          "scene build (%.1f s) not included" % (n, nnz, build_s),
                "sections_ms": sections, "cpu_model": model, "nproc": nproc}
    from oracle import oracle as orc
    rs, col, val, b, _ = ctx.system(n, nnz)
    t0 = time.perf_counter()
    orc.pcg(rs, col, val, b, tol=1e-6, maxit=3000)
    wall = time.perf_counter() - t0
    return {"value": round(wall * 1e3, 1), "unit": "ms/step (solve only)", "cores": 1, "kind": "port",
            "sample": "the oracle's MIC(0)-PCG on the device-assembled 256^3 system (n=%d, nnz=%d), factor + "
                      "solve to 1e-6" % (n, nnz), "cpu_model": model, "nproc": nproc}


def c3_timing(fd_coords=8, steps=20, every=5):
    """The differentiable smoke configuration (astg.scenes.diff_smoke: 128^3
    effective, 3 levels, frozen topology; loss 0.5 |rho_T - target|^2 after
    `steps` steps, gradient with respect to the initial MAC velocity through
    the adjoint kernels, checkpoints every `every` steps): wall time per step
    of the forward and of forward + backward, in the mixed form (FP32 inner
    PCG, FP64 refinement, IC(0), 1e-8: the tolerance at which its gradient
    meets the 1e-4 finite-difference bar; at 1e-6 it misses by 1.5e-3) and in
    FP64 (1e-12); the gradients' relative
    L2 difference; central differences (eps 1e-5, FP64 forward) on
    `fd_coords` seeded face coordinates against both gradients."""
    import numpy as np
    import torch
    import astg
    from astg import scenes
    from astg.diff import DiffSim, rollout
    ctx, info = scenes.diff_smoke(R=128, levels=3, seed=0)
    state0 = DiffSim(ctx, info["dt"], info["rho"], astg.default_params()).get_state().clone()
    target = torch.tensor(info["target"], device="cuda")
    mask = torch.tensor(info["leaf_cells"], device="cuda")

    def loss(sim, u, n):
        return 0.5 * torch.sum(mask * (rollout(u, sim, state0, n, every) - target) ** 2)

    out = {"workload": "diff_smoke: 128^3-effective, 3 levels, %d leaves, frozen topology, %d steps, "
                       "checkpoints every %d" % (info["leaves"], steps, every)}
    grads, sims = {}, {}
    for name, prec, tol in (("mixed", astg.PREC_MIXED, 1e-8), ("fp64", astg.PREC_FP64, 1e-12)):
        sim = DiffSim(ctx, info["dt"], info["rho"], astg.default_params(tol=tol, max_iterations=5000,
                                                                          precision=prec))
        u = torch.tensor(info["u0"], device="cuda", requires_grad=True)
        loss(sim, u, 1).backward()  # first call: the transposes' records
        u = torch.tensor(info["u0"], device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = loss(sim, u, steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        grads[name], sims[name] = u.grad.detach().cpu().numpy(), sim
        out[name] = {"tol": tol, "ms_forward_per_step": round((t1 - t0) * 1e3 / steps, 2),
                     "ms_fwd_bwd_per_step": round((t2 - t0) * 1e3 / steps, 2), "loss": float(L.detach())}
    g64, gmx = grads["fp64"], grads["mixed"]
    out["grad_mixed_vs_fp64_rel_l2"] = float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))
    leaf = np.concatenate([info["leaf_cells"]] * 3) > 0
    cand = np.flatnonzero(leaf & (np.abs(g64) > 1e-2 * np.abs(g64).max()))
    coords = np.random.default_rng(8).choice(cand, size=fd_coords, replace=False)
    worst64 = worstmx = 0.0
    eps = 1e-5
    with torch.no_grad():
        for k in coords:
            up = torch.tensor(info["u0"], device="cuda")
            up[k] += eps
            um = torch.tensor(info["u0"], device="cuda")
            um[k] -= eps
            fd = (float(loss(sims["fp64"], up, steps)) - float(loss(sims["fp64"], um, steps))) / (2 * eps)
            worst64 = max(worst64, abs(fd - g64[k]) / abs(g64[k]))
            worstmx = max(worstmx, abs(fd - gmx[k]) / abs(g64[k]))
    out["fd_check"] = {"coords": fd_coords, "eps": eps, "max_rel_err_fp64": worst64, "max_rel_err_mixed": worstmx}
    ctx.close()
    return out


AST_CONFIGS = [(2, (192, 384)), (2, (384, 768)), (3, (64, 64, 64)), (3, (128, 128, 128))]


def ast_uniform_timing(cpu=True, seed=0):
    """The uniform-layout AST projection (astg.uniform: the reference's
    ASTSolver<d>::project) on seeded inputs: 30 % of the tilt nodes open,
    N(0,1) face and tilt velocities, dt 0.025 (the reference's 2D example),
    rho 2. Per configuration: the device projection in FP64 (the reference's
    MIC(0)-PCG in its numbering, absolute 1e-5 residual), in FP64 with block
    colours (same MIC(0), another ordering) and in the mixed form (block
    colours, IC(0), FP32 inner solve), the first call on the layout
    (pattern set-up included) and a repeated one, and the reference's own
    project() on the host cores for the first three."""
    import numpy as np
    import astg
    from astg import uniform
    rows = []
    for ci, (dim, res) in enumerate(AST_CONFIGS):
        rng = np.random.default_rng(seed + ci)
        dx = 1.0 / res[0]
        tn = int(np.prod([r + 1 for r in res]))
        kt = 4 if dim == 2 else 8
        open_ = (rng.random(tn) < 0.3).astype(np.uint8)
        mac = [rng.standard_normal(int(np.prod([r + (a == i) for i, r in enumerate(res)]))) for a in range(dim)]
        tilt = rng.standard_normal((tn, kt))
        s = uniform.ASTSolver(dim, res, dx, device=0)
        s.set_layout(open_)
        row = {"dim": dim, "res": list(res)}
        for name, prm in (("fp64", uniform.default_params()),
                          ("fp64_block_colours", uniform.default_params(ordering=astg.ORDER_BLOCK_COLOUR)),
                          ("mixed", uniform.default_params(precision=astg.PREC_MIXED, ordering=astg.ORDER_AUTO,
                                                           mic_tau=astg.TAU_AUTO))):
            s.set_layout(open_)  # the first call after a layout change sets the pattern up
            s.set_velocity(mac, tilt)
            r0 = s.project(0.025, 2.0, prm)
            s.set_velocity(mac, tilt)
            r = s.project(0.025, 2.0, prm)
            row[name] = {"ms": round(r.ms_total, 3), "ms_first": round(r0.ms_total, 3), "iterations": r.iterations,
                         "ms_solve": round(r.ms_solve, 3), "ms_factor": round(r.ms_factor, 3),
                         "ms_assemble": round(r.ms_assemble, 3), "depth": r.depth}
            row["n"], row["nnz"] = r.n, r.nnz
            if name == "fp64":
                p64 = s.system()[4]
            elif name == "mixed":
                pmx = s.system()[4]
        s.close()
        if cpu and ci < 3:
            try:
                from oracle import ref
                if ref.available():
                    rf = ref.RefAST(dim, res, dx)
                    rf.set_layout(open_)
                    rf.set_velocity(mac, tilt)
                    t0 = time.perf_counter()
                    rf.project(0.025, 2.0, report=False)
                    ms = (time.perf_counter() - t0) * 1e3
                    pr = rf.system()[4]
                    rf.close()
                    row["reference_cpu"] = {"ms": round(ms, 1), "cores": ref.lib().ref_num_workers(),
                                            "kind": "reference",
                                            "fp64_pressure_rel_l2": float(np.linalg.norm(p64 - pr) /
                                                                          np.linalg.norm(pr)),
                                            "mixed_pressure_rel_l2": float(np.linalg.norm(pmx - pr) /
                                                                           np.linalg.norm(pr))}
                    row["speedup_fp64_vs_cpu"] = round(ms / row["fp64"]["ms"], 2)
                    row["speedup_mixed_vs_cpu"] = round(ms / row["mixed"]["ms"], 2)
            except Exception as e:  # the baseline is reported, never required
                log("ast cpu baseline failed: %s" % e)
        log("ast %dD %s: n %d, fp64 %.1f ms (%d it, depth %d), mixed %.1f ms (%d it), cpu %s" % (
            dim, res, row["n"], row["fp64"]["ms"], row["fp64"]["iterations"], row["fp64"]["depth"],
            row["mixed"]["ms"], row["mixed"]["iterations"], row.get("reference_cpu", {}).get("ms")))
        rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-solve workload")
    ap.add_argument("--c5-steps", type=int, default=20,
                    help="steps of the full-scale adaptive run (512^3, 5 levels; 0 skips it)")
    ap.add_argument("--c5-only", action="store_true", help="run only the full-scale adaptive run")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched-solve workload")
    ap.add_argument("--tune", action="store_true", help="with --batch-only: run the batch autotuner first")
    ap.add_argument("--full-scale", action="store_true",
                    help="run only the full-scale size (512^3-effective, 5 levels, 18 M leaves)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="one projection, then per-kernel timings only (for counter collection)")
    ap.add_argument("--ast-only", action="store_true", help="run only the uniform-layout AST projection")
    ap.add_argument("--advect-only", action="store_true",
                    help="one projection, then the post-projection chain and advect, timed --kernel-reps "
                         "times, and the copy kernel (for counter collection)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: runs on one GPU only (the projection's multi-GPU sharding is not built); "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch  # device memory / stream plumbing only
    import astg
    from astg import scenes

    assert torch.cuda.is_available(), "no GPU visible"
    if a.full_scale:
        fs = full_scale_timing()
        log("full_scale: %s" % json.dumps(fs))
        return
    if a.batch_only:
        bt = batched_timing(tune=a.tune)
        log("batched: %s" % json.dumps(bt))
        return
    if a.ast_only:
        print(json.dumps({"ast_uniform": ast_uniform_timing(cpu=not a.no_cpu)}))
        return
    if a.c5_only:
        c5 = c5_timing(steps=a.c5_steps)
        print(json.dumps({"c5": c5}))
        return
    t0 = time.perf_counter()
    ctx, info = scenes.projection_256(device=0, seed=a.seed)
    prm64 = astg.default_params(tol=1e-6, max_iterations=3000, precision=astg.PREC_FP64)
    prmmx = astg.default_params(tol=1e-6, max_iterations=3000, precision=astg.PREC_MIXED)
    ctx.snapshot(restore=False)
    ctx.project(info["dt"], info["rho"], prm64)  # builds device tables + level schedule
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log("setup %.1f s, leaves %d" % (setup_s, info["leaves"]))

    if a.kernels_only:
        # FP64 kernels on the reference-numbered system, FP32 kernels and the
        # mixed form's factorisation on its block-colour system
        for prm_k, names in ((prm64, [(nm, k) for nm, k in KERNELS if not nm.endswith("32")]),
                             (prmmx, [(nm, k) for nm, k in KERNELS if nm.endswith("32")] +
                              [("mic_mixed", dict(KERNELS)["mic"])])):
            ctx.snapshot(restore=True)
            ctx.project(info["dt"], info["rho"], prm_k)
            for name, k in names:
                ms, by = ctx.kernel_time(k, a.kernel_reps)
                log("%-9s %.4f ms  %.1f GB/s" % (name, ms, by / ms / 1e6))
        return

    if a.advect_only:
        adv = advect_timing(ctx, info, prm64, a.kernel_reps)
        log("advect: %s" % json.dumps(adv))
        ms, by = ctx.kernel_time(dict(KERNELS)["copy"], a.kernel_reps)
        log("copy %.4f ms %.0f bytes" % (ms, by))
        return

    ms64, r64 = time_steps(ctx, info, prm64, a.steps, a.warmup)
    log("fp64  %.2f ms/step iters %d depth %d" % (ms64, r64[-1].iterations, r64[-1].depth))
    f64 = projected_fields(ctx, r64[-1], info["levels"])
    msmx, rmx = time_steps(ctx, info, prmmx, a.steps, a.warmup)
    log("mixed %.2f ms/step iters %d outer %d" % (msmx, rmx[-1].iterations, rmx[-1].outer))
    last64, lastmx = r64[-1], rmx[-1]
    errs = field_errors(projected_fields(ctx, lastmx, info["levels"]), f64)
    del f64
    log("mixed vs fp64: %s" % errs)
    # FP64 with the mixed form's preconditioner (block colours, IC(0)): what the
    # preconditioner choice gives without the precision change
    prm64ic = astg.default_params(tol=1e-6, max_iterations=3000, precision=astg.PREC_FP64,
                                  ordering=astg.ORDER_BLOCK_COLOUR, mic_tau=0.0)
    ms64ic, r64ic = time_steps(ctx, info, prm64ic, a.steps, a.warmup)
    log("fp64 in block colours, IC(0): %.2f ms/step iters %d" % (ms64ic, r64ic[-1].iterations))

    # per-kernel device times (HIP events on the context stream), on the solver
    # each form uses: the FP64 kernels on the reference-numbered system after
    # an FP64 projection, the FP32 ones (and the factorisation, as mic_mixed)
    # on the mixed form's block-colour system after a mixed projection
    kern = {}

    def time_kernels(prm_k, names):
        ctx.snapshot(restore=True)
        ctx.project(info["dt"], info["rho"], prm_k)
        for name, k in names:
            ms, by = ctx.kernel_time(k, a.kernel_reps)
            e = {"ms": round(ms, 4), "kernel_bytes": by, "kernel_GBs": round(by / ms / 1e6, 1)}
            alg = algorithmic_bytes("mic" if name == "mic_mixed" else name, last64.n, last64.nnz)
            if alg is not None:
                e["alg_bytes"] = int(alg)
                e["GBs"] = round(alg / ms / 1e6, 1)
            else:
                e["GBs"] = e["kernel_GBs"]
            kern[name] = e

    time_kernels(prm64, [(nm, k) for nm, k in KERNELS if not nm.endswith("32")])
    time_kernels(prmmx, [(nm, k) for nm, k in KERNELS if nm.endswith("32")] + [("mic_mixed", dict(KERNELS)["mic"])])
    # dominant kernel of the timed (mixed) form: largest share of a step
    share = {k: kern[k]["ms"] * lastmx.iterations for k in ("spmv32", "fwd32", "bwd32")}
    share["mic_mixed"] = kern["mic_mixed"]["ms"]  # one FP64 factorisation per step
    dom = max(share, key=share.get)
    ach = kern[dom]["GBs"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / MI355X_HBM_GBS, 4), "traffic": TRAFFIC.get(dom),
                "traffic_source": TRAFFIC_SOURCE,
                "bytes_per_launch": kern[dom].get("alg_bytes"),
                "bytes_basis": "compulsory bytes of one launch from n and nnz, each array streamed once "
                               "(sweeps: (w+4) nnz_L + (3w+4) n forward, (w+4) nnz_L + (4w+4) n backward)",
                "merged_stencil": {"achieved": kern[dom]["kernel_GBs"],
                                   "frac": round(kern[dom]["kernel_GBs"] / MI355X_HBM_GBS, 4),
                                   "bytes_per_launch": kern[dom]["kernel_bytes"],
                                   "note": "bytes of the merged-level stencil tables the sweep streams "
                                           "(more entries than nnz_L), not the compulsory figure"},
                "kernels": kern}

    adv = advect_timing(ctx, info, prm64, a.kernel_reps)
    log("advect %.2f ms (post-project chain %.2f ms), %.1f M back-traces/s" % (
        adv["ms_advect"], adv["ms_post_project"], adv["backtraces_per_s"] / 1e6))

    smoke = smoke_timing()
    log("smoke_64: %.2f ms/step over %d steps (mean PCG iterations %.1f)" % (
        smoke["ms_per_step"], smoke["steps"], smoke["pcg_iterations_mean"]))

    batched = None
    if not a.no_batch:
        batched = batched_timing()
        log("batched (%d systems, %d leaves): mixed %.1f ms vs %.1f ms one after another; fp64 %.1f vs %.1f ms" % (
            batched["systems"], batched["leaves_total"], batched["mixed"]["ms_batch"],
            batched["mixed"]["ms_sequential"], batched["fp64"]["ms_batch"], batched["fp64"]["ms_sequential"]))

    c3 = c3_timing()
    log("c3: mixed %.1f ms per fwd+bwd step, fp64 %.1f; gradient mixed vs fp64 %.2e; fd %s" % (
        c3["mixed"]["ms_fwd_bwd_per_step"], c3["fp64"]["ms_fwd_bwd_per_step"], c3["grad_mixed_vs_fp64_rel_l2"],
        c3["fd_check"]))

    c5 = None
    if a.c5_steps > 0:
        c5 = c5_timing(steps=a.c5_steps)
        log("c5 (%d steps): fp64 %.0f ms/step, mixed %.0f ms/step; last velocity error %.2e" % (
            c5["steps"], c5["fp64"]["ms_per_step"], c5["mixed"]["ms_per_step"],
            c5["mixed_vs_fp64"][-1]["velocity"]["rel_l2"]))

    ast_rows = ast_uniform_timing(cpu=not a.no_cpu)

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(ctx, info, last64, a.seed)
            log("reference project() on the CPU: %.1f ms (%d workers), sections %s" % (
                cpu["value"], cpu["cores"], cpu.get("sections_ms")))
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed: %s" % e)
        try:
            cs = cpu_smoke_baseline()
            if cs is not None and cpu is not None:
                cpu["smoke_64"] = cs
                log("smoke_64 reference CPU: %.1f ms/step (%d cores)" % (cs["ms_per_step"], cs["cores"]))
        except Exception as e:
            log("cpu smoke baseline failed: %s" % e)

    out = {
        "metric": "projection ms/step to 1e-6 rel. residual @256^3 octree (mixed vs FP64); HBM GB/s",
        "value": round(msmx, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(msmx, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner PCG + f64 refinement (value); f64 reference form reported beside it",
        "data": "synthetic (seeded sphere octree, N(0,1) FP32 face velocities)",
        "config": {"workload": "projection_256: 256^3-effective AST octree, 4 levels, %d leaves, n=%d unknowns, "
                               "nnz=%d, 6 seeded spheres, level-0 tilts opened in spheres" % (
                                   info["leaves"], last64.n, last64.nnz),
                   "seed": a.seed, "tol": 1e-6, "norm": "inf (reference)"},
        "fp64": {"ms_per_step": round(ms64, 3), "iterations": last64.iterations,
                 "ms_assemble": round(last64.ms_assemble, 3), "ms_factor": round(last64.ms_factor, 3),
                 "ms_solve": round(last64.ms_solve, 3), "ms_velocity": round(last64.ms_velocity, 3),
                 "rel_residual": last64.rel_residual},
        "mixed": {"ms_per_step": round(msmx, 3), "iterations": lastmx.iterations, "outer": lastmx.outer,
                  "ms_solve": round(lastmx.ms_solve, 3), "rel_residual": lastmx.rel_residual},
        "speedup_mixed_vs_fp64": round(ms64 / msmx, 3),
        "mixed_vs_fp64": errs,
        "sweep_depth": {"fp64": last64.depth, "mixed": lastmx.depth},
        "ordering": {"fp64": "reference numbering", "mixed": "block colours (8 colours of 8^3-cell cubes)"},
        "preconditioner": {"fp64": "MIC(0), tau 0.97 (the reference's)", "mixed": "IC(0) (MIC(0) with tau 0)"},
        "fp64_mixed_preconditioner": {"ms_per_step": round(ms64ic, 3), "iterations": r64ic[-1].iterations,
                                      "note": "FP64 in block colours with IC(0): the mixed form's preconditioner "
                                              "without the precision change"},
        "setup_s": round(setup_s, 2),
        "roofline": roofline,
        "advect": adv,
        "smoke_64": smoke,
        "batched": batched,
        "c3": c3,
        "c5": c5,
        "ast_uniform": ast_rows,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
