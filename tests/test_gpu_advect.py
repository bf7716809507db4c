"""Device post-projection chain, point sampling and advection against the
reference's outputs (tests/golden/adv_*.npz, written by
tests/golden/make_golden_advect.py from the reference build).

Through the C ABI on device 0, starting from the projection fixture's
corrected velocity and seeded scalars:
  * onUpdate's chain between project and advect (wall constraint, velocity
    ghosts, dual grid and its ghosts): velocities to 1e-14, dual to 1e-12;
  * OctreeOmniGrid3::updateGhost of density and heat: 1e-14;
  * point queries (omni and MAC velocity, BackTrace, density): 1e-12;
  * one advect(dt): density, heat and velocity per node: 1e-12.
Tolerances are relative to the largest magnitude of the compared array; the
reference is built with -ffast-math, so bit equality is not required.
"""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden, make_context

pytestmark = pytest.mark.gpu

ADV = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "adv_*.npz")))


@pytest.fixture(params=ADV, ids=[os.path.basename(p)[4:-4] for p in ADV])
def adv(request):
    a = load_golden(request.param)
    g = load_golden(os.path.join(ROOT, "tests", "golden", str(a["scene"]) + ".npz"))
    return g, a


def relmax(x, y):
    x, y = np.asarray(x), np.asarray(y)
    if x.size == 0:
        return 0.0
    return np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)


def prepared(g, a):
    c = make_context(g, 0)
    L = int(g["levels"])
    for l in range(L):
        c.set_velocity(l, g["u_mac_out_%d" % l], g["u_tilt_out_%d" % l])
    for f in (0, 1):
        for l in range(L):
            c.set_scalar(f, l, a["s%d_in_oct_%d" % (f, l)], a["s%d_in_tilt_%d" % (f, l)])
    return c, L


def test_post_project_chain_matches_reference(adv):
    g, a = adv
    c, L = prepared(g, a)
    c.post_project()
    for l in range(L):
        m, t = c.velocity(l)
        assert relmax(m, a["vel_post_mac_%d" % l]) < 1e-14, l
        assert relmax(t, a["vel_post_tilt_%d" % l]) < 1e-14, l
        o, v = c.dual(l)
        assert relmax(o, a["dual_oct_%d" % l]) < 1e-12, l
        assert relmax(v, a["dual_tilt_%d" % l]) < 1e-12, l


def test_scalar_ghosts_match_reference(adv):
    g, a = adv
    c, L = prepared(g, a)
    for f in (0, 1):
        c.scalar_update_ghost(f)
        for l in range(L):
            o, t = c.scalar(f, l)
            assert relmax(o, a["s%d_ghost_oct_%d" % (f, l)]) < 1e-14, (f, l)
            assert relmax(t, a["s%d_ghost_tilt_%d" % (f, l)]) < 1e-14, (f, l)


def test_point_samplers_match_reference(adv):
    import astg
    g, a = adv
    c, L = prepared(g, a)
    c.post_project()
    c.scalar_update_ghost(0)
    pts, dt = a["pts"], float(a["dt"])
    for what, key in ((astg.SAMPLE_MAC, "q_mac"), (astg.SAMPLE_VELOCITY, "q_vel"),
                      (astg.SAMPLE_BACKTRACE, "q_bt"), (astg.SAMPLE_DENSITY, "q_den")):
        q = c.sample_points(what, pts, dt)
        assert relmax(q, a[key]) < 1e-12, key


def test_advect_matches_reference(adv):
    g, a = adv
    c, L = prepared(g, a)
    c.post_project()
    c.scalar_update_ghost(0)
    c.scalar_update_ghost(1)
    c.advect(float(a["dt"]))
    for l in range(L):
        for f in (0, 1):
            o, t = c.scalar(f, l)
            assert relmax(o, a["s%d_adv_oct_%d" % (f, l)]) < 1e-12, (f, l)
            assert relmax(t, a["s%d_adv_tilt_%d" % (f, l)]) < 1e-12, (f, l)
        m, t = c.velocity(l)
        assert relmax(m, a["vel_adv_mac_%d" % l]) < 1e-12, l
        assert relmax(t, a["vel_adv_tilt_%d" % l]) < 1e-12, l


def test_step_equals_its_parts(adv):
    """astg_step (constrain + ghosts, project, post-projection chain, advect)
    gives the same fields as the separate calls on the same input."""
    import astg
    g, a = adv
    dt, rho = float(g["dt"]), float(g["rho"])
    prm = astg.default_params(tol=float(g["tol"]), max_iterations=3000)
    outs = []
    for whole in (True, False):
        c = make_context(g, 0)
        L = int(g["levels"])
        for l in range(L):
            c.set_velocity(l, g["u_mac_in_%d" % l], g["u_tilt_in_%d" % l])
        for f in (0, 1):
            for l in range(L):
                c.set_scalar(f, l, a["s%d_in_oct_%d" % (f, l)], a["s%d_in_tilt_%d" % (f, l)])
        if whole:
            res = c.step(dt, rho, prm)
        else:
            c.post_project()  # constrain + ghosts before project (the dual part is recomputed below)
            res = c.project(dt, rho, prm)
            c.post_project()
            c.advect(dt)
        assert res.converged == 1
        outs.append([np.concatenate([np.ravel(x) for x in c.scalar(0, l) + c.velocity(l)]) for l in range(L)])
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
        assert np.all(np.isfinite(x))
