"""Symbolic plans built on the device against the host planners.

The MIC(0) factorisation plan and the merged sweep stencils are built by
device kernels (plan_device.hip) unless ASTG_HOST_PLAN=1 selects the host
planners (mic_plan.cpp, sweep_merge.cpp). Both walk the pattern in the same
order, so the plans agree entry for entry, and every result computed from
them must be bit-identical: the factor, the PCG solution, iteration count and
residual (FP64 and mixed forms), the corrected velocities, a batched solve,
and a run of smoke steps whose tilt adaptivity rebuilds the plans every step.
"""
import numpy as np
import pytest

from conftest import GOLDEN, load_golden, make_context

pytestmark = pytest.mark.gpu


def _project(g, precision):
    import astg
    c = make_context(g, 0)
    L = int(g["levels"])
    for l in range(L):
        c.set_velocity(l, g["u_mac_in_%d" % l], g["u_tilt_in_%d" % l])
    prm = astg.default_params(tol=float(g["tol"]), max_iterations=3000, precision=precision)
    res = c.project(float(g["dt"]), float(g["rho"]), prm)
    n, nnz = res.n, res.nnz
    out = {"iters": res.iterations, "res": res.rel_residual}
    out["p"] = c.system(n, nnz)[4]
    out["factor"] = c.mic_factor(n)
    out["vel"] = [c.velocity(l) for l in range(L)]
    c.close()
    return out


def _both(monkeypatch, f):
    runs = []
    for host in ("1", "0"):
        monkeypatch.setenv("ASTG_HOST_PLAN", host)
        runs.append(f())
    return runs


def _same_arrays(a, b):
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(_same_arrays(x, y) for x, y in zip(a, b))
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_projection_identical_with_host_plans(golden, precision, monkeypatch):
    host, dev = _both(monkeypatch, lambda: _project(golden, precision))
    assert host["iters"] == dev["iters"]
    assert host["res"] == dev["res"]
    assert _same_arrays(host["factor"], dev["factor"])
    assert np.array_equal(host["p"], dev["p"])
    assert _same_arrays(host["vel"], dev["vel"])


def test_batched_solve_identical_with_host_plans(monkeypatch):
    import astg
    gs = [load_golden(p) for p in GOLDEN]
    systems = [(g["A_rowstart"], g["A_col"], g["A_val"], g["b"]) for g in gs]
    prm = astg.default_params(tol=1e-6, max_iterations=3000)

    def run():
        c = astg.Context(0)
        xs, res = c.pcg_solve_batch(systems, prm)
        c.close()
        return xs, [(r.iterations, r.rel_residual) for r in res]

    (xh, rh), (xd, rd) = _both(monkeypatch, run)
    assert rh == rd
    assert _same_arrays(xh, xd)


def test_adaptive_smoke_steps_identical_with_host_plans(monkeypatch):
    import astg
    from astg import scenes
    prm = astg.default_params(tol=1e-6, max_iterations=3000)

    def run():
        ctx, info = scenes.smoke_64(device=0, R=32)
        its = []
        for k in range(4):
            its.append(scenes.smoke_step(ctx, info, k == 0, prm).iterations)
        L = info["levels"]
        out = (its, [ctx.velocity(l) for l in range(L)], [ctx.scalar(0, l) for l in range(L)],
               [ctx.tilt(l) for l in range(L)])
        ctx.close()
        return out

    h, d = _both(monkeypatch, run)
    assert h[0] == d[0]
    for a, b in zip(h[1:], d[1:]):
        assert _same_arrays(a, b)


def _tridiag():
    rs = [0, 2, 5, 8, 10]
    col = [0, 1, 0, 1, 2, 1, 2, 3, 2, 3]
    return rs, col


def _wide():
    # row 17 couples to all 17 rows before it: more lower entries than a sweep row holds
    rs, col = [0], []
    for j in range(17):
        col += [j, 17]
        rs.append(len(col))
    col += list(range(18))
    rs.append(len(col))
    return rs, col


@pytest.mark.parametrize("case,message", [
    ("unsorted", "row 0 is not sorted by column"),
    ("range", "column index out of range"),
    ("asymmetric", "matrix is not structurally symmetric at (1,0)"),
    ("wide", "row 17 has more than 16 entries on one side of the diagonal"),
])
def test_invalid_patterns_are_rejected(case, message):
    """The structural checks run on the device; the host words the error."""
    import astg
    if case == "wide":
        rs, col = _wide()
    else:
        rs, col = _tridiag()
        if case == "unsorted":
            col[0:2] = [1, 0]
        elif case == "range":
            col[9] = 4
        else:
            rs = [0, 1, 4, 7, 9]
            col = [0, 0, 1, 2, 1, 2, 3, 2, 3]
    n = len(rs) - 1
    val = np.where(np.array(col) == np.repeat(np.arange(n), np.diff(rs)), 4.0, -1.0)
    c = astg.Context(0)
    with pytest.raises(astg.AstgError) as e:
        c.pcg_solve_csr(np.array(rs), np.array(col), val, np.ones(n))
    assert e.value.code == astg.ASTG_ERR_ARG
    assert message in str(e.value)
    # the context stays usable: a valid system solves afterwards
    rs, col = _tridiag()
    val = np.where(np.array(col) == np.repeat(np.arange(4), np.diff(rs)), 4.0, -1.0)
    x, res = c.pcg_solve_csr(np.array(rs), np.array(col), val, np.ones(4))
    assert res.converged == 1
    c.close()
