"""The differentiable step: adjoint kernels against the forward kernels.

On a small frozen smoke topology (32^3, 3 levels, T-joints and half-tilts
from the refinement, level-0 tilts opened around the source):
  * dot-product tests per linear operator of the step, <A x, y> = <x, A^T y>
    with A x from the forward kernels and A^T y from the recorded transposes:
    wall constraint + velocity ghosts, the post-projection chain (dual grid,
    dual ghosts), the density ghosts (to 1e-12 of the terms' magnitude), and
    the projection with its transposed solve (to 1e-9: the solves' tolerance);
  * the advection (nonlinear in the velocity): <J v, w> from central
    differences of the forward (FP64, eps 1e-8; outputs where the forward is
    not differentiable left out: one-sided differences that disagree mark a
    back-traced point on a kink of the piecewise samplers, such as a trace
    along a closed wall, and central differences at two step sizes that
    disagree mark a jump) against
    <v, J^T w> from the adjoint kernels (interpolation-weight transposes and
    back-trace position derivatives), to 1e-6 of the terms' magnitude;
  * a 4-step rollout through torch.autograd (checkpoints every 2 steps): the
    gradient of 0.5 |rho_T - target|^2 with respect to the initial MAC
    velocity agrees with central differences (eps 1e-5, FP64 projection to
    1e-12) on 12 seeded coordinates to 1e-4 relative, and equals the
    gradient with a checkpoint at every step;
  * the mixed-precision (FP32 inner PCG) gradient stays within 1e-3 of the
    FP64 one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_scene():
    import astg
    from astg import scenes
    from astg.diff import DiffSim
    ctx, info = scenes.diff_smoke(R=32, levels=3, seed=1)
    # open the level-0 tilts within the refined region, then freeze
    xyz, _ = ctx.nodes(0)
    nleaf = ctx.level_info(0)[1]
    ctx.set_tilt_open(0, scenes.inside_mask(xyz[:nleaf], 32, [(0.5, 0.5, 0.12, 0.1)]))
    cx, cy, cz, r = scenes.SMOKE_SOURCE
    ctx.apply_sphere_source((cx, cy, cz), r, density=1.0, heat=0.0, speed=0.0, update_ghost=True)
    prm = astg.default_params(tol=1e-12, max_iterations=3000)
    sim = DiffSim(ctx, info["dt"], info["rho"], prm)
    return ctx, info, sim


def rand_state(sim, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal(sim.n)


def dot_check(sim, op, tol):
    x = rand_state(sim, 10 + op)
    y = rand_state(sim, 20 + op)
    ax = sim.apply(op, x)
    aty = sim.apply(op, y, transpose=True)
    lhs, rhs = float(ax @ y), float(x @ aty)
    scale = float(np.abs(ax * y).sum()) + 1e-300
    assert abs(lhs - rhs) <= tol * scale, (op, lhs, rhs, abs(lhs - rhs) / scale)
    return abs(lhs - rhs) / scale


def test_linear_operator_transposes():
    from astg import diff
    ctx, info, sim = small_scene()
    for op in (diff.OP_CONSTRAIN, diff.OP_POST, diff.OP_SCALAR_GHOST):
        print("dot test op %d" % op, flush=True)
        dot_check(sim, op, 1e-12)
    print("dot test projection", flush=True)
    dot_check(sim, diff.OP_PROJECT, 1e-9)
    ctx.close()


def test_advection_transpose_matches_differences():
    from astg import diff
    ctx, info, sim = small_scene()
    # a projected, smooth-ish state to linearise at: the forward of one step
    import torch
    s0 = sim.get_state()
    a, b = sim.regions["mac"]
    s0[a:b] = torch.tensor(np.random.default_rng(2).standard_normal(b - a) * 0.5, device="cuda")
    s = sim.step(s0).cpu().numpy()  # a state with a consistent dual grid
    rng = np.random.default_rng(3)
    v = rng.standard_normal(sim.n)
    w = rng.standard_normal(sim.n)
    # The reference's samplers are piecewise: trilinear pieces meet with a kink
    # at every cell face, a level change or the switch to the sub-cell sampler
    # near a transition changes the formula, and positions are clamped at the
    # domain walls. The advection is not differentiable at an output whose
    # back-traced points lie on such a boundary. A trace that starts on a
    # closed wall stays on the wall plane (the wall-normal velocity is zero
    # there), which is a kink of the clamped sampler: the adjoint returns one
    # side's derivative there, a central difference the mean of the two.
    # Outputs whose forward and backward differences disagree (a kink at the
    # point) or whose central differences at two step sizes disagree (a jump
    # within the perturbation) are left out of w. eps 1e-8 moves the
    # back-traced points by ~1e-10 (dt = 1/64).
    eps = 1e-8
    f0 = sim.apply(diff.OP_ADVECT, s)
    fp, fm = sim.apply(diff.OP_ADVECT, s + eps * v), sim.apply(diff.OP_ADVECT, s - eps * v)
    fp2, fm2 = sim.apply(diff.OP_ADVECT, s + 2 * eps * v), sim.apply(diff.OP_ADVECT, s - 2 * eps * v)
    j1, j2 = (fp - fm) / (2 * eps), (fp2 - fm2) / (4 * eps)
    fwd, bwd = (fp - f0) / eps, (f0 - fm) / eps
    tol = 1e-4 * (np.abs(j1) + 1.0)
    kink, jump = np.abs(fwd - bwd) > tol, np.abs(j1 - j2) > tol
    smooth = ~(kink | jump)
    for name, (a, b) in sim.regions.items():
        print("advection %-12s outputs at a kink %d, across a jump %d, of %d" %
              (name, int(kink[a:b].sum()), int((jump & ~kink)[a:b].sum()), b - a), flush=True)
    assert (~smooth).sum() <= 0.01 * sim.n
    w = np.where(smooth, w, 0.0)
    sim.set_state_host(s)
    jtw = sim.apply(diff.OP_ADVECT, w, transpose=True)
    lhs, rhs = float(j1 @ w), float(v @ jtw)
    scale = float(np.abs(j1 * w).sum())
    print("advection <Jv,w> %.12e <v,J^T w> %.12e (terms %.3e)" % (lhs, rhs, scale), flush=True)
    assert abs(lhs - rhs) <= 1e-6 * scale
    ctx.close()


def loss_fn(rho, info, torch):
    target = torch.tensor(info["target"], device="cuda")
    mask = torch.tensor(info["leaf_cells"], device="cuda")
    return 0.5 * torch.sum(mask * (rho - target) ** 2)


def rollout_loss(sim, info, state0, u0, steps, every, torch):
    from astg.diff import rollout
    rho = rollout(u0, sim, state0, steps, every)
    return loss_fn(rho, info, torch)


def test_rollout_gradient_matches_finite_differences():
    import torch
    ctx, info, sim = small_scene()
    state0 = sim.get_state()
    steps = 4
    u0 = torch.tensor(info["u0"], device="cuda", requires_grad=True)
    L = rollout_loss(sim, info, state0, u0, steps, 2, torch)
    L.backward()
    g = u0.grad.detach().cpu().numpy()
    # every-step checkpoints give the same gradient
    u1 = torch.tensor(info["u0"], device="cuda", requires_grad=True)
    rollout_loss(sim, info, state0, u1, steps, 1, torch).backward()
    assert np.allclose(u1.grad.cpu().numpy(), g, rtol=1e-10, atol=1e-14 * np.abs(g).max())
    # central differences on leaf MAC faces with the largest gradients and random ones
    G = info["G"]
    leaf = np.concatenate([info["leaf_cells"]] * 3) > 0
    cand = np.flatnonzero(leaf & (np.abs(g) > 1e-3 * np.abs(g).max()))
    rng = np.random.default_rng(5)
    coords = rng.choice(cand, size=min(12, len(cand)), replace=False)
    eps = 1e-5
    worst = 0.0
    with torch.no_grad():
        for k in coords:
            up = torch.tensor(info["u0"], device="cuda")
            up[k] += eps
            um = torch.tensor(info["u0"], device="cuda")
            um[k] -= eps
            fd = (float(rollout_loss(sim, info, state0, up, steps, 2, torch)) -
                  float(rollout_loss(sim, info, state0, um, steps, 2, torch))) / (2 * eps)
            rel = abs(fd - g[k]) / max(abs(g[k]), 1e-300)
            print("coordinate %d: adjoint %.10e fd %.10e rel %.2e" % (k, g[k], fd, rel), flush=True)
            worst = max(worst, rel)
    assert worst < 1e-4, worst
    ctx.close()


def test_mixed_precision_gradient_close_to_fp64():
    import torch
    import astg
    ctx, info, sim = small_scene()
    state0 = sim.get_state()
    grads = []
    for prec in (astg.PREC_FP64, astg.PREC_MIXED):
        sim.params = astg.default_params(tol=1e-8, max_iterations=3000, precision=prec)
        u0 = torch.tensor(info["u0"], device="cuda", requires_grad=True)
        rollout_loss(sim, info, state0, u0, 4, 2, torch).backward()
        grads.append(u0.grad.detach().cpu().numpy())
    rel = np.linalg.norm(grads[1] - grads[0]) / np.linalg.norm(grads[0])
    print("mixed vs fp64 gradient rel L2 %.3e" % rel, flush=True)
    assert rel < 1e-3
    ctx.close()
