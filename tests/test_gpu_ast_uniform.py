"""The uniform-layout AST projection on the device (astg_ast_*: numbering,
pattern, assembly, MIC(0)-PCG, velocity correction and wall constraint)
against the reference's own ASTSolver<d>::project (the committed fixtures of
tests/golden/make_golden_ast.py) and the oracle (oracle/ast_uniform.py).

Bars:
  * tilt numbering and CSR pattern bit-exact with the reference;
  * A and b bit-exact with the oracle (the same strict-IEEE operation order)
    and within 1e-12 of their largest magnitude of the reference's (built with
    -ffast-math);
  * the solve: iterations within one of the reference's (the residual test is
    the reference's absolute ||r||_inf <= 1e-5; the device reduces the dots in
    another order), the host-computed residual ||b - A p||_inf within that
    bound; with the same iteration count, pressure and corrected velocities
    within 1e-7 of the reference's (relative L2 / relative to the largest
    velocity: rounding of the dots only), one iteration apart within 1e-5 and
    1e-4 (one iteration's progress, times dt / (rho dx) for the velocities);
    the x walls (and the 3D y walls) exactly zero.
At 512^2, 64^3 and 48x40x56 (sizes the fixtures do not reach) the pattern, A
and b, and the velocity correction applied to the device's own pressure, are
checked bit for bit against the oracle, and the residual bound on the device's
own system. At the reference's example size (192x384) and at 64^3 the live
reference (where built) is held to the fixtures' bars.
"""
import os

import numpy as np
import pytest

from conftest import AST_GOLDEN, load_golden

pytestmark = pytest.mark.gpu
IDS = [os.path.basename(p)[:-4] for p in AST_GOLDEN]


def _inputs(g):
    d = int(g["dim"])
    res = tuple(int(r) for r in g["res"])
    mac = [g["mac_in_%d" % a] for a in range(d)]
    return d, res, float(g["dx"]), g["open"], mac, g["tilt_in"], float(g["dt"]), float(g["rho"])


def _residual(rs, col, val, b, p):
    rows = np.repeat(np.arange(len(rs) - 1), np.diff(rs))
    return np.max(np.abs(b - np.bincount(rows, weights=val * p[col], minlength=len(b))))


def _device(d, res, dx, open_, mac, tilt):
    from astg import uniform
    s = uniform.ASTSolver(d, res, dx, device=0)
    s.set_layout(open_)
    s.set_velocity(mac, tilt)
    return s


@pytest.mark.parametrize("path", AST_GOLDEN, ids=IDS)
def test_device_projection_matches_reference(path):
    from oracle import ast_uniform as au
    g = load_golden(path)
    d, res, dx, open_, mac, tilt, dt, rho = _inputs(g)
    print("case %s: dim %d res %s" % (os.path.basename(path), d, res), flush=True)
    s = _device(d, res, dx, open_, mac, tilt)
    r = s.project(dt, rho)
    rs, col, val, b, p = s.system()
    assert np.array_equal(s.tilt_index(), g["tilt_index"])
    assert np.array_equal(rs, g["rowstart"]) and np.array_equal(col, g["col"])
    ors, ocol, oval, ob = au.assemble(res, dx, open_, mac, tilt, dt, rho)
    assert np.array_equal(val, oval) and np.array_equal(b, ob), "device system differs from the oracle's"
    assert np.max(np.abs(val - g["val"])) <= 1e-12 * np.max(np.abs(g["val"]))
    assert np.max(np.abs(b - g["b"])) <= 1e-12 * np.max(np.abs(g["b"]))
    gp = g["p"]
    ep = np.linalg.norm(p - gp) / np.linalg.norm(gp)
    resid = _residual(rs, col, val, b, p)
    m2, t2 = s.velocity()
    scale = max(max(np.max(np.abs(g["mac_out_%d" % a])) for a in range(d)), np.max(np.abs(g["tilt_out"])))
    em = max(np.max(np.abs(m2[a] - g["mac_out_%d" % a])) for a in range(d)) / scale
    et = np.max(np.abs(t2 - g["tilt_out"])) / scale
    print("n %d nnz %d: iterations %d (reference %d), residual %.2e, pressure %.2e, MAC %.2e, tilt %.2e, "
          "%.3f ms" % (r.n, r.nnz, r.iterations, int(g["iters"]), resid, ep, em, et, r.ms_total), flush=True)
    assert abs(r.iterations - int(g["iters"])) <= 1
    assert resid <= 1e-5
    if r.iterations == int(g["iters"]):
        assert ep <= 1e-7 and em <= 1e-7 and et <= 1e-7
    else:
        assert ep <= 1e-5 and em <= 1e-4 and et <= 1e-4
    nx, ny = res[0], res[1]
    ux = m2[0].reshape(-1, nx + 1)
    assert not ux[:, 0].any() and not ux[:, nx].any()
    if d == 3:
        uy = m2[1].reshape(res[2], ny + 1, nx)
        assert not uy[:, 0, :ny].any() and not uy[:, ny, :ny].any()
    s.close()


@pytest.mark.parametrize("dim,res", [(2, (512, 512)), (3, (64, 64, 64)), (3, (48, 40, 56))])
def test_device_system_at_size_matches_oracle(dim, res):
    from oracle import ast_uniform as au
    rng = np.random.default_rng(5)
    dx = 1.0 / max(res)
    tn = int(np.prod([r + 1 for r in res]))
    open_ = (rng.random(tn) < 0.3).astype(np.uint8)
    mac = [rng.standard_normal(int(np.prod([r + (a == i) for i, r in enumerate(res)]))) for a in range(dim)]
    tilt = rng.standard_normal((tn, 4 if dim == 2 else 8))
    print("size case: dim %d res %s" % (dim, res), flush=True)
    s = _device(dim, res, dx, open_, mac, tilt)
    r = s.project(0.01, 2.0)
    rs, col, val, b, p = s.system()
    ors, ocol, oval, ob = au.assemble(res, dx, open_, mac, tilt, 0.01, 2.0)
    assert np.array_equal(rs, ors) and np.array_equal(col, ocol)
    assert np.array_equal(val, oval) and np.array_equal(b, ob)
    resid = _residual(rs, col, val, b, p)
    print("n %d nnz %d: %d iterations, residual %.2e, %.2f ms (assembly %.2f, factor %.2f, solve %.2f)" % (
        r.n, r.nnz, r.iterations, resid, r.ms_total, r.ms_assemble, r.ms_factor, r.ms_solve), flush=True)
    assert r.converged == 1 and resid <= 1e-5
    # the correction and wall constraint on the device's own pressure, bit for bit
    m2, t2 = s.velocity()
    om, ot = au.correct(res, dx, open_, mac, tilt, p, 0.01, 2.0)
    for a in range(dim):
        assert np.array_equal(m2[a], om[a]), "MAC axis %d" % a
    assert np.array_equal(t2, ot)
    s.close()


def test_layout_change_rebuilds_pattern():
    from oracle import ast_uniform as au
    g = load_golden(AST_GOLDEN[0])
    d, res, dx, open_, mac, tilt, dt, rho = _inputs(g)
    s = _device(d, res, dx, open_, mac, tilt)
    s.project(dt, rho)
    closed = np.zeros_like(open_)
    s.set_layout(closed)
    s.set_velocity(mac, tilt)
    r = s.project(dt, rho)
    rs, col, val, b, p = s.system()
    assert r.n == int(np.prod(res))
    ors, ocol, oval, ob = au.assemble(res, dx, closed, mac, tilt, dt, rho)
    assert np.array_equal(rs, ors) and np.array_equal(col, ocol) and np.array_equal(val, oval)
    assert np.array_equal(b, ob)
    assert (s.tilt_index() == -1).all()
    s.close()


def test_mixed_precision_meets_the_bound():
    import astg
    from astg import uniform
    g = load_golden(AST_GOLDEN[0])
    d, res, dx, open_, mac, tilt, dt, rho = _inputs(g)
    s = _device(d, res, dx, open_, mac, tilt)
    r = s.project(dt, rho, uniform.default_params(precision=astg.PREC_MIXED))
    rs, col, val, b, p = s.system()
    resid = _residual(rs, col, val, b, p)
    ep = np.linalg.norm(p - g["p"]) / np.linalg.norm(g["p"])
    print("mixed: %d iterations (%d outer), residual %.2e, pressure %.2e" % (r.iterations, r.outer, resid, ep))
    assert r.converged == 1 and resid <= 1e-5 and ep <= 1e-4
    s.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_block_colour_solve_meets_the_bound(precision):
    """The block-colour ordering (an added form: another preconditioner, so
    another trajectory): the reference's residual bound on the device's own
    system, pressure within 1e-4 relative L2 of the reference's, on every
    fixture."""
    import astg
    from astg import uniform
    for path in AST_GOLDEN:
        g = load_golden(path)
        d, res, dx, open_, mac, tilt, dt, rho = _inputs(g)
        s = _device(d, res, dx, open_, mac, tilt)
        r = s.project(dt, rho, uniform.default_params(precision=precision, ordering=astg.ORDER_BLOCK_COLOUR,
                                                      mic_tau=astg.TAU_AUTO))
        rs, col, val, b, p = s.system()
        resid = _residual(rs, col, val, b, p)
        ep = np.linalg.norm(p - g["p"]) / np.linalg.norm(g["p"])
        print("%s: %d iterations (reference numbering %d), depth %d, residual %.2e, pressure %.2e" % (
            os.path.basename(path), r.iterations, int(g["iters"]), r.depth, resid, ep), flush=True)
        assert r.converged == 1 and resid <= 1e-5 and ep <= 1e-4
        s.close()


def test_bad_arguments_are_rejected():
    import astg
    from astg import uniform
    with pytest.raises(astg.AstgError) as e:
        uniform.ASTSolver(3, (8, 9, 8), 0.1)  # res_y > res_x in 3D
    assert e.value.code == astg.ASTG_ERR_ARG
    s = uniform.ASTSolver(2, (8, 8), 0.125)
    with pytest.raises(ValueError):
        s.set_layout(np.zeros(5, np.uint8))
    with pytest.raises(astg.AstgError) as e:
        s.project(-1.0)
    assert e.value.code == astg.ASTG_ERR_ARG
    with pytest.raises(astg.AstgError) as e:
        s.project(0.01, 2.0, uniform.default_params(ordering=7))
    assert e.value.code == astg.ASTG_ERR_ARG
    r = s.project(0.01)  # zero field: zero right-hand side, no iteration
    assert r.iterations == 0 and r.converged == 1
    s.close()


@pytest.mark.parametrize("dim,res", [(2, (192, 384)), (3, (64, 64, 64))])
def test_device_matches_live_reference_at_example_size(dim, res):
    """The reference's own example size (example/ast_smoke_2d/main.cpp,
    python/ast_smoke_2d.ipynb: 192x384, dx 1/192, dt 0.025) and a 3D grid,
    against the reference built from its sources (skipped where absent), to
    the fixtures' bars."""
    from oracle import ref
    if not ref.available():
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(21)
    dx = 1.0 / res[0]
    rf = ref.RefAST(dim, res, dx)
    open_ = (rng.random(rf.tilt_count()) < 0.3).astype(np.uint8)
    mac = [rng.standard_normal(rf.mac_count(a)) for a in range(dim)]
    tilt = rng.standard_normal((rf.tilt_count(), rf.kt))
    rf.set_layout(open_)
    rf.set_velocity(mac, tilt)
    rf.project(0.025, 2.0)
    grs, gcol, gval, gb, gp = rf.system()
    git, _ = rf.solve_info()
    gm, gt = rf.velocity()
    rf.close()
    print("size case: dim %d res %s" % (dim, res), flush=True)
    s = _device(dim, res, dx, open_, mac, tilt)
    r = s.project(0.025, 2.0)
    rs, col, val, b, p = s.system()
    assert np.array_equal(rs, grs) and np.array_equal(col, gcol)
    assert np.max(np.abs(val - gval)) <= 1e-12 * np.max(np.abs(gval))
    assert np.max(np.abs(b - gb)) <= 1e-12 * np.max(np.abs(gb))
    ep = np.linalg.norm(p - gp) / np.linalg.norm(gp)
    m2, t2 = s.velocity()
    scale = max(max(np.max(np.abs(gm[a])) for a in range(dim)), np.max(np.abs(gt)))
    em = max(np.max(np.abs(m2[a] - gm[a])) for a in range(dim)) / scale
    et = np.max(np.abs(t2 - gt)) / scale
    print("n %d: iterations %d (reference %d), pressure %.2e, MAC %.2e, tilt %.2e, %.1f ms" % (
        r.n, r.iterations, git, ep, em, et, r.ms_total), flush=True)
    assert git >= 0 and abs(r.iterations - git) <= 1
    assert _residual(rs, col, val, b, p) <= 1e-5
    if r.iterations == git:
        assert ep <= 1e-7 and em <= 1e-7 and et <= 1e-7
    else:
        assert ep <= 1e-5 and em <= 1e-4 and et <= 1e-4
    s.close()
