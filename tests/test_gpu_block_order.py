"""The block-colour ordering built on the device (reorder.hip BlockOrderDevice:
keys on host threads, stable radix sort, inverse permutation, permuted pattern
per row) against the host construction (reorder.cpp buildBlockOrder).

With ASTG_CHECK_BLOCK_ORDER=1 every block-colour projection also builds the
host form and fails unless permutation, row starts, columns and source map
agree entry for entry. Checked on every golden scene, on a batched scene
system (99 k rows) and over adaptive smoke steps, where the tilt state and
so the pattern change every step. Each projection must also converge.
"""
import numpy as np
import pytest

from conftest import GOLDEN, load_golden, make_context

pytestmark = pytest.mark.gpu


def _mixed_block(tol=1e-6):
    import astg
    return astg.default_params(tol=tol, max_iterations=3000, precision=astg.PREC_MIXED,
                               ordering=astg.ORDER_BLOCK_COLOUR)


@pytest.mark.parametrize("path", GOLDEN, ids=[p.split("/")[-1][:-4] for p in GOLDEN])
def test_block_order_on_golden_scenes(path, monkeypatch):
    monkeypatch.setenv("ASTG_CHECK_BLOCK_ORDER", "1")
    g = load_golden(path)
    c = make_context(g, 0)
    for l in range(int(g["levels"])):
        c.set_velocity(l, g["u_mac_in_%d" % l], g["u_tilt_in_%d" % l])
    r = c.project(float(g["dt"]), float(g["rho"]), _mixed_block(float(g["tol"])))
    assert r.converged == 1
    c.close()


def test_block_order_on_a_batched_scene(monkeypatch):
    from astg import scenes
    monkeypatch.setenv("ASTG_CHECK_BLOCK_ORDER", "1")
    c, info = scenes.batch_scene(0, 0, scenes.batch_targets(0)[0])
    r = c.project(info["dt"], info["rho"], _mixed_block())
    print("n %d, %d iterations" % (r.n, r.iterations))
    assert r.converged == 1
    c.close()


def test_block_order_over_adaptive_smoke_steps(monkeypatch):
    from astg import scenes
    monkeypatch.setenv("ASTG_CHECK_BLOCK_ORDER", "1")
    ctx, info = scenes.smoke_64(device=0)
    ns = set()
    for k in range(6):
        r = scenes.smoke_step(ctx, info, k == 0, _mixed_block())
        assert r.converged == 1
        ns.add(r.n)
    print("unknowns over the steps: %s" % sorted(ns))
    assert len(ns) > 1  # the pattern changed between steps
    ctx.close()
