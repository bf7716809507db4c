"""The uniform-layout AST projection's oracle (oracle/ast_uniform.py) against
the reference's own project() (ast::ASTSolver<d>, include/ast_solver.h:322-498)
on the committed fixtures (tests/golden/make_golden_ast.py): 2D and 3D, square
and non-square, sparse, dense and no open tilts.

Bar: numbering and CSR pattern bit-exact; A and b within 1e-12 of their
largest magnitude (the reference is built with -ffast-math); the same number of
PCG iterations (the oracle's MIC(0)-PCG is the reference's algorithm with the
same parameters); pressure and corrected velocities within 1e-10 (relative L2 /
absolute). A planted mistake in the oracle (a tilt row's sign) must fail the
velocity check, and the live reference (where built) must agree with the
oracle on a fresh seed."""
import os

import numpy as np
import pytest

from conftest import AST_GOLDEN, load_golden

IDS = [os.path.basename(p)[:-4] for p in AST_GOLDEN]


def _inputs(g):
    d = int(g["dim"])
    res = tuple(int(r) for r in g["res"])
    mac = [g["mac_in_%d" % a] for a in range(d)]
    return d, res, float(g["dx"]), g["open"], mac, g["tilt_in"], float(g["dt"]), float(g["rho"])


def _check(g, out, d):
    rel = lambda a, b: np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
    assert np.array_equal(out["rowstart"], g["rowstart"]) and np.array_equal(out["col"], g["col"])
    assert rel(out["val"], g["val"]) <= 1e-12
    assert rel(out["b"], g["b"]) <= 1e-12
    assert np.linalg.norm(out["p"] - g["p"]) <= 1e-10 * np.linalg.norm(g["p"])
    for a in range(d):
        assert np.max(np.abs(out["mac"][a] - g["mac_out_%d" % a])) <= 1e-10
    assert np.max(np.abs(out["tilt"] - g["tilt_out"])) <= 1e-10


@pytest.mark.parametrize("path", AST_GOLDEN, ids=IDS)
def test_oracle_matches_reference_fixture(path):
    from oracle import ast_uniform as au
    g = load_golden(path)
    d, res, dx, open_, mac, tilt, dt, rho = _inputs(g)
    cell, tid, n = au.numbering(res, open_)
    assert np.array_equal(cell, g["cell_index"]) and np.array_equal(tid, g["tilt_index"])
    out = au.project(res, dx, open_, mac, tilt, dt, rho)
    assert out["iters"] == int(g["iters"])
    _check(g, out, d)


def test_planted_sign_mistake_is_caught():
    from oracle import ast_uniform as au
    g = load_golden(AST_GOLDEN[0])
    d, res, dx, open_, mac, tilt, dt, rho = _inputs(g)
    bad = np.array(tilt, copy=True)
    bad[:, 0] = -bad[:, 0]  # as if slot 0's outward sign were flipped
    out = au.project(res, dx, open_, mac, bad, dt, rho)
    with pytest.raises(AssertionError):
        _check(g, out, d)


@pytest.mark.parametrize("dim,res", [(2, (30, 22)), (3, (10, 9, 7))])
def test_oracle_matches_live_reference(dim, res):
    from oracle import ref, ast_uniform as au
    if not ref.available():
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(11)
    r = ref.RefAST(dim, res, 1.0 / max(res))
    open_ = (rng.random(r.tilt_count()) < 0.5).astype(np.uint8)
    mac = [rng.standard_normal(r.mac_count(a)) for a in range(dim)]
    tilt = rng.standard_normal((r.tilt_count(), r.kt))
    r.set_layout(open_)
    r.set_velocity(mac, tilt)
    r.project(0.01, 2.0)
    rs, col, val, b, p = r.system()
    it, _ = r.solve_info()
    m2, t2 = r.velocity()
    g = dict(rowstart=rs, col=col, val=val, b=b, p=p, tilt_out=t2)
    for a in range(dim):
        g["mac_out_%d" % a] = m2[a]
    out = au.project(res, 1.0 / max(res), open_, mac, tilt, 0.01, 2.0)
    assert out["iters"] == it
    _check(g, out, dim)
    r.close()


def test_device_entry_points_fail_loudly_without_a_device():
    """The product path has no host fallback: creating a device solver where
    no GPU is visible is an error, not a silent CPU path."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import astg
    from astg import uniform
    with pytest.raises(astg.AstgError) as e:
        uniform.ASTSolver(2, (8, 8), 0.125, device=0)
    assert e.value.code == astg.ASTG_ERR_HIP
