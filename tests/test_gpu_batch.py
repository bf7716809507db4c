"""Batched form (astg_pcg_solve_batch) against the single-system seam and the
reference's solutions.

Every reference system of the golden fixtures is solved in one batch together
with edge cases. Per system the batch runs the same PCG recurrence as
astg_pcg_solve_csr on that system alone (same factor, same sweeps; only the
summation order of the dots differs), so:
  * iterations equal the single solve's (within 1) and the reference's (within 1);
  * x within 1e-6 relative L2 (the FP64 parity bar) of the single solve and
    of the reference (closed boxes modulo the constant null vector); a changed
    summation order can move the stopping test by one iteration, so exact
    agreement is not expected;
  * two copies of one system in a batch give bit-identical solutions (their
    chunks line up, so even the dot partials are the same);
  * the mixed form meets the FP64 residual target b - A x on every system.
"""
import numpy as np
import pytest

from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    s = np.linalg.norm(np.asarray(b))
    return d / s if s > 0 else d


def csr_system(g):
    return (g["A_rowstart"], g["A_col"], g["A_val"], g["b"])


def spmv(rs, col, val, x):
    rows = np.repeat(np.arange(len(rs) - 1), np.diff(rs))
    return np.bincount(rows, weights=val * x[col], minlength=len(rs) - 1)


def closed(g):
    return int(g["closed"]) == 63


def tiny_system():
    # 1x1 system 2 x = 4
    return (np.array([0, 1]), np.array([0]), np.array([2.0]), np.array([4.0]))


@pytest.fixture(scope="module")
def goldens():
    return [load_golden(p) for p in GOLDEN]


def test_batch_matches_single_solves_and_reference(goldens):
    import astg
    c = astg.Context(0)
    prm = astg.default_params(tol=1e-6, max_iterations=3000)
    systems = [csr_system(g) for g in goldens]
    zero = csr_system(goldens[0])
    systems.append((zero[0], zero[1], zero[2], np.zeros_like(zero[3])))  # zero rhs: 0 iterations, x = 0
    systems.append(tiny_system())
    systems.append(csr_system(goldens[1]))  # a second copy of system 1
    print("batch of %d systems, n = %s" % (len(systems), [len(s[0]) - 1 for s in systems]), flush=True)
    xs, res = c.pcg_solve_batch(systems, prm)
    for k, g in enumerate(goldens):
        xs1, r1 = c.pcg_solve_csr(*csr_system(g), prm)
        assert res[k].converged == 1
        assert abs(res[k].iterations - r1.iterations) <= 1, (k, res[k].iterations, r1.iterations)
        assert abs(res[k].iterations - int(g["pcg_iters"])) <= 1
        x, x1, xr = xs[k], xs1, g["p"]
        if closed(g):  # singular, consistent: compare modulo the constant null vector
            x, x1, xr = x - x.mean(), x1 - x1.mean(), xr - xr.mean()
        print("system %d: iterations %d (single %d), rel to single %.2e, rel to reference %.2e"
              % (k, res[k].iterations, r1.iterations, rel(x, x1), rel(x, xr)), flush=True)
        assert rel(x, x1) < 1e-6
        assert rel(x, xr) < 1e-6
    kz = len(goldens)
    assert res[kz].converged == 1 and res[kz].iterations == 0 and not np.any(xs[kz])
    assert res[kz + 1].converged == 1 and abs(xs[kz + 1][0] - 2.0) < 1e-12
    assert np.array_equal(xs[kz + 2], xs[1]) and res[kz + 2].iterations == res[1].iterations


def test_batch_of_one_equals_single_solve(goldens):
    import astg
    c = astg.Context(0)
    prm = astg.default_params(tol=1e-6, max_iterations=3000)
    g = goldens[-1]
    xb, rb = c.pcg_solve_batch([csr_system(g)], prm)
    xs, r1 = c.pcg_solve_csr(*csr_system(g), prm)
    assert abs(rb[0].iterations - r1.iterations) <= 1
    assert rel(xb[0], xs) < 1e-6


def test_batch_mixed_meets_fp64_residual(goldens):
    import astg
    c = astg.Context(0)
    tol = 1e-6
    prm = astg.default_params(tol=tol, max_iterations=3000, precision=astg.PREC_MIXED)
    systems = [csr_system(g) for g in goldens]
    xs, res = c.pcg_solve_batch(systems, prm)
    for k, (rs, col, val, b) in enumerate(systems):
        r = b - spmv(rs, col, val, xs[k])
        assert res[k].converged == 1 and res[k].outer >= 1
        assert np.max(np.abs(r)) <= tol * np.max(np.abs(b)) * (1 + 1e-6), k
        assert res[k].rel_residual <= tol


def test_batch_pattern_cache_and_iteration_limit(goldens):
    """A different batch after a first one (new stacked pattern), the first
    again (cached pattern), and a limit that stops every system unconverged."""
    import astg
    c = astg.Context(0)
    prm = astg.default_params(tol=1e-6, max_iterations=3000)
    a = [csr_system(g) for g in goldens[:2]]
    b = [csr_system(g) for g in goldens[2:]]
    xa, ra = c.pcg_solve_batch(a, prm)
    xb, rb = c.pcg_solve_batch(b, prm)
    xa2, ra2 = c.pcg_solve_batch(a, prm)
    for u, v in zip(xa, xa2):
        assert np.array_equal(u, v)
    assert [r.iterations for r in ra] == [r.iterations for r in ra2]
    lim = astg.default_params(tol=1e-6, max_iterations=5)
    xl, rl = c.pcg_solve_batch(b, lim, allow_unconverged=True)
    assert all(r.converged == 0 and r.iterations == 5 for r in rl)
    with pytest.raises(astg.AstgError):
        c.pcg_solve_batch(b, lim)


def test_batch_rejects_cross_system_columns(goldens):
    import astg
    c = astg.Context(0)
    rs, col, val, b = csr_system(goldens[0])
    bad = col.copy()
    bad[1] = len(rs) - 1 + 5  # column past the system's own rows
    with pytest.raises(astg.AstgError):
        c.pcg_solve_batch([(rs, bad, val, b), csr_system(goldens[1])])


def _scene_systems():
    import astg
    from astg import scenes
    systems = []
    for seed, R, band in ((3, 64, (0.06, 0.12)), (4, 128, (0.03, 0.06))):
        ctx, info = scenes.projection_256(device=0, seed=seed, R=R, levels=3, band=band)
        res = ctx.project(info["dt"], info["rho"], astg.default_params(tol=1e-6))
        rs, col, val, b, _ = ctx.system(res.n, res.nnz)
        systems.append((rs, col, val, b))
        ctx.close()
    return systems


def test_sweep_ordered_solve_matches_row_ordered():
    """Single solves keep their vectors in sweep order (contiguous sweeps);
    ASTG_ROW_ORDER=1 keeps them in row order, as the batched form does. Only
    the summation order of s.q differs, so away from the stopping boundary
    (1e-10) the iteration counts agree within max(2, 1 %) and the solutions to
    1e-8 relative."""
    import os
    import astg
    systems = _scene_systems()
    prm = astg.default_params(tol=1e-10, max_iterations=3000)
    out = []
    for row in ("1", "0"):
        os.environ["ASTG_ROW_ORDER"] = row
        try:
            c = astg.Context(0)
            out.append([c.pcg_solve_csr(*s, prm) for s in systems])
            c.close()
        finally:
            del os.environ["ASTG_ROW_ORDER"]
    for (xr, rr), (xp, rp) in zip(*out):
        assert rr.converged == 1 and rp.converged == 1
        assert abs(rr.iterations - rp.iterations) <= max(2, rr.iterations // 100), (rr.iterations, rp.iterations)
        assert rel(xp, xr) < 1e-8


def test_batch_scene_systems_match_single(monkeypatch):
    """Two device-assembled octree systems (35K and 195K leaves: hundreds of
    chunks each, different depths), batched, against their single solves in
    the batch's own vector layout (row order; the sweep-ordered single solve
    is compared with the row-ordered one above)."""
    import astg
    from astg import scenes
    monkeypatch.setenv("ASTG_ROW_ORDER", "1")
    systems = []
    for seed, R, band in ((3, 64, (0.06, 0.12)), (4, 128, (0.03, 0.06))):
        ctx, info = scenes.projection_256(device=0, seed=seed, R=R, levels=3, band=band)
        res = ctx.project(info["dt"], info["rho"], astg.default_params(tol=1e-6))
        rs, col, val, b, _ = ctx.system(res.n, res.nnz)
        systems.append((rs, col, val, b))
        ctx.close()
    c = astg.Context(0)
    # FP64 at 1e-6 (the solve tolerance) and at 1e-10, away from the stopping
    # boundary, where the two solutions agree to far below the solve
    # tolerance; the summation order of the dots moves the inf-norm stopping
    # test by an iteration or two. Mixed at 1e-6 with the FP64 residual
    # checked on the host.
    for prec, tol in ((astg.PREC_FP64, 1e-6), (astg.PREC_FP64, 1e-10), (astg.PREC_MIXED, 1e-6)):
        prm = astg.default_params(tol=tol, max_iterations=3000, precision=prec)
        xs, res = c.pcg_solve_batch(systems, prm)
        for k, s in enumerate(systems):
            x1, r1 = c.pcg_solve_csr(*s, prm)
            print("prec %d tol %g system %d n=%d iters %d / single %d, rel %.2e" % (
                prec, tol, k, len(s[3]), res[k].iterations, r1.iterations, rel(xs[k], x1)), flush=True)
            assert res[k].converged == 1
            r = s[3] - spmv(s[0], s[1], s[2], xs[k])
            assert np.max(np.abs(r)) <= tol * np.max(np.abs(s[3])) * (1 + 1e-6)
            if prec == astg.PREC_FP64:
                assert abs(res[k].iterations - r1.iterations) <= max(2, r1.iterations // 100)
                assert rel(xs[k], x1) < (1e-8 if tol < 1e-8 else 1e-5)
            else:
                assert rel(xs[k], x1) < 1e-4


def test_batch_tuning_shapes_agree_and_tuner_sets_one(goldens):
    """Every launch shape the autotuner can choose gives the reference's
    solutions; the tuner returns one of its candidates and sets it."""
    import astg
    c = astg.Context(0)
    prm = astg.default_params(tol=1e-6, max_iterations=3000)
    systems = [csr_system(g) for g in goldens]
    for q, blk, lanes in ((1, 4, 4), (3, 16, 16), (2, 8, 8)):
        c.set_batch_tuning(q, blk, lanes)
        xs, res = c.pcg_solve_batch(systems, prm)
        for k, g in enumerate(goldens):
            x, xr = xs[k], g["p"]
            if closed(g):
                x, xr = x - x.mean(), xr - xr.mean()
            assert res[k].converged == 1 and abs(res[k].iterations - int(g["pcg_iters"])) <= 1
            assert rel(x, xr) < 1e-6, (q, blk, lanes, k)
    best = c.tune_batch(systems, astg.default_params(precision=astg.PREC_MIXED), iterations=5)
    print("tuned:", best.as_dict(), flush=True)
    assert best.merge_levels in (1, 2, 3) and best.stream_blocks_per_cu in (4, 8, 16)
    assert best.lanes_per_row in (4, 8, 16) and best.ms_per_iteration > 0
    xs, res = c.pcg_solve_batch(systems, prm)
    assert all(r.converged == 1 for r in res)
    with pytest.raises(astg.AstgError):
        c.set_batch_tuning(2, 8, 6)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_graph_replay_equals_eager(goldens, prec):
    """Iterations replayed from the captured two-iteration hipGraphs (with the
    host's convergence checks between replays, and in the mixed form the
    residual-replacement variant) give the eager loop's results bit for bit."""
    import os
    import astg
    precision = astg.PREC_FP64 if prec == "fp64" else astg.PREC_MIXED
    # a long inner solve (tolerance 1e-12) so that replacement steps occur
    prm = astg.default_params(tol=1e-12, max_iterations=3000, precision=precision)
    systems = [csr_system(g) for g in goldens]
    out = {}
    keys = ("ASTG_BATCH_GRAPH", "ASTG_MIXED_REPLACE")
    old = {k: os.environ.get(k) for k in keys}
    try:
        # mixed: replace the residual every 10 iterations, so the graph
        # variant with a replacement step is replayed too
        os.environ["ASTG_MIXED_REPLACE"] = "10"
        for mode in ("0", "1"):
            os.environ["ASTG_BATCH_GRAPH"] = mode
            c = astg.Context(0)
            out[mode] = c.pcg_solve_batch(systems, prm, allow_unconverged=True)
            c.close()
    finally:
        for k in keys:
            if old[k] is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = old[k]
    (xe, re_), (xg, rg) = out["0"], out["1"]
    print("iterations eager %s graph %s" % ([r.iterations for r in re_], [r.iterations for r in rg]), flush=True)
    for k in range(len(systems)):
        assert re_[k].iterations == rg[k].iterations and re_[k].converged == rg[k].converged
        assert np.array_equal(xe[k], xg[k]), k
