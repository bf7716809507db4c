"""Symbolic plans built on the device against the host planners.

The MIC(0) factorisation plan and the merged sweep stencils are built by
device kernels (plan_device.hip) unless ASTG_HOST_PLAN=1 selects the host
planners (mic_plan.cpp, sweep_merge.cpp). Both walk the pattern in the same
order, so the plans agree entry for entry, and every result computed from
them must be bit-identical: the factor, the PCG solution, iteration count and
residual (FP64 and mixed forms), the corrected velocities, a batched solve,
and a run of smoke steps whose tilt adaptivity rebuilds the plans every step.
"""
import os

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


def _longest_path_depth(rs, col):
    """Dependency depth of tril(A) on the host: level(i) = 1 + max level(k)
    over the row's entries left of the diagonal (0 without any), rows in
    index order (every such k < i)."""
    rs, col = np.asarray(rs), np.asarray(col)
    n = len(rs) - 1
    lev = np.zeros(n, np.int64)
    for i in range(n):
        c = col[rs[i]:rs[i + 1]]
        c = c[c < i]
        if len(c):
            lev[i] = lev[c].max() + 1
    return int(lev.max()) + 1


def test_device_levels_give_the_longest_path_depth():
    """The device level computation (k_levels, one lane per row, sync-free)
    against a plain host longest-path count: the depth the solve reports is
    the DAG's, on every golden system and on one batched scene system."""
    import astg
    from astg import scenes
    c = astg.Context(0)
    systems = []
    for p in GOLDEN:
        g = load_golden(p)
        systems.append((os.path.basename(p), g["A_rowstart"], g["A_col"], g["A_val"], g["b"]))
    sc, info = scenes.batch_scene(0, 0, scenes.batch_targets(0)[0])
    r0 = sc.project(info["dt"], info["rho"], astg.default_params(tol=1e-6, max_iterations=3000))
    rs, col, val, b, _ = sc.system(r0.n, r0.nnz)
    sc.close()
    systems.append(("batch scene 0", rs, col, val, b))
    for name, rs, col, val, b in systems:
        _, r = c.pcg_solve_csr(rs, col, val, b, astg.default_params(tol=1e-6, max_iterations=3000))
        want = _longest_path_depth(rs, col)
        print("%s: n %d depth %d (host %d)" % (name, len(rs) - 1, r.depth, want), flush=True)
        assert r.depth == want, name
    c.close()
