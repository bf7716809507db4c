"""Several contexts of different topologies in one process.

The octree kernels read per-topology tables (level resolutions, spacings,
map offsets) from constant memory shared by every context on a GPU; a change
of context re-uploads them. Interleaving two contexts' calls must give the
results each gets alone.
"""
import numpy as np
import pytest

from conftest import ROOT, load_golden, make_context

pytestmark = pytest.mark.gpu


def run(ctx, g, pts, steps):
    """steps: list of operation names applied in order; returns the outputs."""
    import astg
    L = int(g["levels"])
    out = []
    for op in steps:
        if op == "set":
            for l in range(L):
                ctx.set_velocity(l, g["u_mac_in_%d" % l], g["u_tilt_in_%d" % l])
        elif op == "project":
            ctx.project(float(g["dt"]), float(g["rho"]), astg.default_params(tol=1e-8))
        elif op == "post":
            ctx.post_project()
        elif op == "sample":
            out.append(ctx.sample_points(astg.SAMPLE_VELOCITY, pts))
            out.append(ctx.sample_points(astg.SAMPLE_BACKTRACE, pts, float(g["dt"])))
        elif op == "advect":
            ctx.advect(float(g["dt"]))
            out.extend(ctx.velocity(l)[0] for l in range(L))
    return out


def test_interleaved_contexts_match_single_runs():
    a_g = load_golden("%s/tests/golden/s32_two_spheres.npz" % ROOT)
    b_g = load_golden("%s/tests/golden/s16_sphere.npz" % ROOT)
    pts = np.random.default_rng(3).uniform(0.05, 0.95, (2000, 3))
    ops = ["set", "project", "post", "sample", "advect"]
    alone = []
    for g in (a_g, b_g):
        c = make_context(g, 0)
        alone.append(run(c, g, pts, ops))
        c.close()
    a = make_context(a_g, 0)
    b = make_context(b_g, 0)
    got = ([], [])
    for op in ops:  # alternate the two contexts call by call
        print("interleaved op %s" % op, flush=True)
        got[0].extend(run(a, a_g, pts, [op]))
        got[1].extend(run(b, b_g, pts, [op]))
    for k in range(2):
        assert len(got[k]) == len(alone[k])
        for x, y in zip(got[k], alone[k]):
            assert np.array_equal(x, y)
    a.close()
    b.close()
