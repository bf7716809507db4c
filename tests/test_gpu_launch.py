"""Sweep launch mode: ordinary launches of the persistent sweeps (the default)
fall back to cooperative launches when a sweep reports its spin limit, and
the retried solve is the reference's solution.

The failure is forced from the host (ASTG_TEST_SPIN_ERROR_ONCE sets a sweep's
error bit once, as a block that waited for a non-resident one would), so no
kernel is made to wait on purpose.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_spin_error_falls_back_to_cooperative_launches(capfd):
    import astg
    g = load_golden(GOLDEN[0])
    keys = ("ASTG_TEST_SPIN_ERROR_ONCE", "ASTG_COOP")
    old = {k: os.environ.get(k) for k in keys}
    try:
        os.environ["ASTG_TEST_SPIN_ERROR_ONCE"] = "1"
        os.environ.pop("ASTG_COOP", None)
        c = astg.Context(0)
        prm = astg.default_params(tol=float(g["tol"]), max_iterations=3000)
        x, r = c.pcg_solve_csr(g["A_rowstart"], g["A_col"], g["A_val"], g["b"], prm)
        err = capfd.readouterr().err
        assert "retrying with cooperative launches" in err
        assert r.converged == 1 and abs(r.iterations - int(g["pcg_iters"])) <= 1
        assert rel(x, g["p"]) < 1e-6
        # the context keeps cooperative launches: no second notice
        x2, r2 = c.pcg_solve_csr(g["A_rowstart"], g["A_col"], g["A_val"], g["b"], prm)
        assert "retrying" not in capfd.readouterr().err
        assert np.array_equal(x, x2)
    finally:
        for k in keys:
            if old[k] is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = old[k]


def test_cooperative_and_ordinary_launches_agree():
    import astg
    g = load_golden(GOLDEN[-1])
    out = {}
    old = os.environ.get("ASTG_COOP")
    try:
        for mode in ("1", "0"):
            os.environ["ASTG_COOP"] = mode
            c = astg.Context(0)
            prm = astg.default_params(tol=float(g["tol"]), max_iterations=3000)
            out[mode] = c.pcg_solve_csr(g["A_rowstart"], g["A_col"], g["A_val"], g["b"], prm)
    finally:
        if old is None:
            os.environ.pop("ASTG_COOP", None)
        else:
            os.environ["ASTG_COOP"] = old
    (xc, rc), (xo, ro) = out["1"], out["0"]
    assert rc.iterations == ro.iterations and np.array_equal(xc, xo)
