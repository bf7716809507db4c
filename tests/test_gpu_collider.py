"""The collider path on the device against the reference run with its own
sphere collider (tests/golden/col_*.npz, written by make_golden_collider.py).

Through the C ABI on device 0, scenes with the reference's initObstacle sphere
(radius 0.1, and 0.25 in col_s32_big, at (0.5, 0.5, 0.45)):
  * projection with the collider's numbering and cutting-cell weights: A and b
    within 1e-12, iterations within 1, pressure and corrected velocity within
    1e-6 relative L2 (the FP64 parity bar) after the reference's iterations;
  * onUpdate's chain after project with the slip constraint (MAC faces in the
    reference's visiting order, the tilt slots, the dual vertex vectors):
    velocity and dual grid within 1e-12 of the largest magnitude, against the
    harness's serial run of the reference's constraint loop;
  * point queries: boundary SDF, velocity, MAC, BackTrace (with its SDF
    clip), density: 1e-12;
  * one advect(dt) and wipeBoundary: 1e-12.
"""
import numpy as np
import pytest

from test_collider import col, collider_context  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    s = np.linalg.norm(np.asarray(b))
    return d / s if s > 0 else d


def relmax(x, y):
    x, y = np.asarray(x), np.asarray(y)
    if x.size == 0:
        return 0.0
    return np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)


def projected(g):
    import astg
    c = collider_context(g, 0)
    L = int(g["levels"])
    for l in range(L):
        c.set_velocity(l, g["u_mac_in_%d" % l], g["u_tilt_in_%d" % l])
    c.snapshot()
    prm = astg.default_params(tol=float(g["tol"]), max_iterations=3000)
    res = c.project(float(g["dt"]), float(g["rho"]), prm)
    ref_it = int(g["pcg_iters"])
    if res.iterations != ref_it and abs(res.iterations - ref_it) == 1:
        c.snapshot(restore=True)
        c.project(float(g["dt"]), float(g["rho"]), astg.default_params(tol=0.0, max_iterations=ref_it),
                  allow_unconverged=True)
    return c, res


def test_projection_with_collider(col):
    print("collider projection: R=%d levels=%d radius=%.2f" % (int(col["R"]), int(col["levels"]),
                                                                 float(col["collider"][3])), flush=True)
    c, res = projected(col)
    n, nnz = len(col["b"]), len(col["A_col"])
    assert res.n == n and res.nnz == nnz
    rs, cl, val, b, p = c.system(n, nnz)
    assert np.array_equal(rs, col["A_rowstart"]) and np.array_equal(cl, col["A_col"])
    assert relmax(val, col["A_val"]) < 1e-12
    assert relmax(b, col["b"]) < 1e-12
    assert abs(res.iterations - int(col["pcg_iters"])) <= 1
    assert rel(p, col["p"]) < 1e-6
    for l in range(int(col["levels"])):
        m, t = c.velocity(l)
        assert rel(m, col["u_mac_out_%d" % l]) < 1e-6, l
        assert rel(t, col["u_tilt_out_%d" % l]) < 1e-6, l


def prepared(g):
    """The reference's projected velocity and seeded scalars on the device."""
    c = collider_context(g, 0)
    L = int(g["levels"])
    for l in range(L):
        c.set_velocity(l, g["u_mac_out_%d" % l], g["u_tilt_out_%d" % l])
    for f in (0, 1):
        for l in range(L):
            c.set_scalar(f, l, g["s%d_in_oct_%d" % (f, l)], g["s%d_in_tilt_%d" % (f, l)])
    return c, L


def test_slip_chain_matches_reference(col):
    print("collider chain: radius=%.2f" % float(col["collider"][3]), flush=True)
    c, L = prepared(col)
    c.post_project()
    for l in range(L):
        m, t = c.velocity(l)
        assert relmax(m, col["vel_post_mac_%d" % l]) < 1e-12, l
        assert relmax(t, col["vel_post_tilt_%d" % l]) < 1e-12, l
        o, v = c.dual(l)
        assert relmax(o, col["dual_oct_%d" % l]) < 1e-12, l
        assert relmax(v, col["dual_tilt_%d" % l]) < 1e-12, l


def test_point_queries_with_collider(col):
    import astg
    c, L = prepared(col)
    c.post_project()
    c.scalar_update_ghost(0)
    pts, dt = col["pts"], float(col["dt"])
    for what, key in ((astg.SAMPLE_SDF, "q_sdf"), (astg.SAMPLE_MAC, "q_mac"), (astg.SAMPLE_VELOCITY, "q_vel"),
                      (astg.SAMPLE_BACKTRACE, "q_bt"), (astg.SAMPLE_DENSITY, "q_den")):
        q = c.sample_points(what, pts, dt)
        assert relmax(q, col[key]) < 1e-12, key


def test_advect_and_wipe_with_collider(col):
    c, L = prepared(col)
    c.post_project()
    c.scalar_update_ghost(0)
    c.scalar_update_ghost(1)
    c.advect(float(col["dt"]))
    for l in range(L):
        for f in (0, 1):
            o, t = c.scalar(f, l)
            assert relmax(o, col["s%d_adv_oct_%d" % (f, l)]) < 1e-12, (f, l)
            assert relmax(t, col["s%d_adv_tilt_%d" % (f, l)]) < 1e-12, (f, l)
        m, t = c.velocity(l)
        assert relmax(m, col["vel_adv_mac_%d" % l]) < 1e-12, l
        assert relmax(t, col["vel_adv_tilt_%d" % l]) < 1e-12, l
    c.wipe_boundary()
    for l in range(L):
        for f in (0, 1):
            o, t = c.scalar(f, l)
            assert relmax(o, col["s%d_wipe_oct_%d" % (f, l)]) < 1e-12, (f, l)
            assert relmax(t, col["s%d_wipe_tilt_%d" % (f, l)]) < 1e-12, (f, l)
