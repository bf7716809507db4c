"""The batched workload at its full size: the 24 seeded octree systems of
astg.scenes.batch_scene (seed 0; leaf counts log-uniform in [5e4, 2e6], 15.4 M
leaves, 17.7 M unknowns in all), solved in one astg_pcg_solve_batch call in the
mixed form to 1e-6, with the launch shape bench.py uses.

Per system, FP64 and mixed: the FP64 relative residual ||b - A x||_inf /
||b||_inf computed on the host from the batch's solution is within 1e-6, and
the batch's iteration count is within max(2, 2 %) of the same system solved
alone through the CSR seam (astg_pcg_solve_csr with the batch's sweep
grouping, one level per group: the same ordering, factor and recurrence; only
the summation order of the dots differs). The bar is 2 % rather than 1 %: the
reference's inf-norm stopping test sits on a residual plateau on the larger
systems, where the dot order alone moves the stop by a few iterations
(measured: at most 4 of 235 in FP64 and 3 of 228 mixed, both on systems of
about a million unknowns). Building the 24 scenes takes about 20 s of host
time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# the launch shape bench.py freezes for this workload
TUNING = {"merge_levels": 1, "stream_blocks_per_cu": 8, "lanes_per_row": 4}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_of_24_scenes(precision, monkeypatch):
    import astg
    from astg import scenes
    # single solves with the batch's sweep grouping (read when a pattern is set up)
    monkeypatch.setenv("ASTG_MERGE", str(TUNING["merge_levels"]))
    systems = []
    for k, target in enumerate(scenes.batch_targets(0, 24)):
        ctx, info = scenes.batch_scene(0, k, target)
        r = ctx.project(info["dt"], info["rho"], astg.default_params(tol=1e-6, max_iterations=3000))
        rs, col, val, b, _ = ctx.system(r.n, r.nnz)
        systems.append((rs, col, val, b))
        ctx.close()
    print("24 systems, %d unknowns, %d entries" % (sum(len(s[3]) for s in systems),
                                                   sum(len(s[1]) for s in systems)), flush=True)
    prm = astg.default_params(tol=1e-6, max_iterations=3000,
                              precision=astg.PREC_MIXED if precision == "mixed" else astg.PREC_FP64)
    c = astg.Context(0)
    c.set_batch_tuning(**TUNING)
    xs, res = c.pcg_solve_batch(systems, prm)
    worst, worst_it = 0.0, 0.0
    ok_it = True
    for k, ((rs, col, val, b), x, rb) in enumerate(zip(systems, xs, res)):
        rows = np.repeat(np.arange(len(rs) - 1), np.diff(rs))
        ax = np.bincount(rows, weights=val * x[col], minlength=len(b))
        resid = np.max(np.abs(b - ax)) / np.max(np.abs(b))
        _, rsg = c.pcg_solve_csr(rs, col, val, b, prm)
        d = abs(rb.iterations - rsg.iterations)
        print("system %2d: n %8d, batch %4d iterations, alone %4d, residual %.3e" % (
            k, len(b), rb.iterations, rsg.iterations, resid), flush=True)
        worst = max(worst, resid)
        worst_it = max(worst_it, d / rsg.iterations)
        assert rb.converged == 1 and resid <= 1e-6
        ok_it = ok_it and d <= max(2, rsg.iterations // 50)
    print("%s: worst relative residual %.3e, largest iteration difference %.1f %%" % (
        precision, worst, 100 * worst_it), flush=True)
    assert ok_it
    c.close()
