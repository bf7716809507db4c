"""astg_sample_points_noalloc against astg_sample_points.

The no-alloc entry point stages points and results through the library's
fixed per-GPU device array instead of buffers the context allocates; the
kernel and its inputs are the same, so every result must be bit-identical.
Checked on two calls in a row, the second with fewer points on a smaller
scene (the array keeps the first call's contents), for every query kind, and
the size limit and the set-up requirement are checked from the host.
"""
import numpy as np
import pytest

import astg
from astg import scenes

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


def _scene(R, levels, seed):
    """Smoke-configuration topology, a projected N(0,1) velocity, seeded scalars."""
    split = scenes.smoke_64_split(R, levels)
    c = astg.Context(0)
    c.set_topology(split, R, levels, scenes.CLOSED_XY)
    scenes.fill_velocity(c, levels, seed)
    c.project(0.5 / R, 2.0, astg.default_params(tol=1e-8))
    rng = np.random.default_rng(seed)
    for f in (astg.FIELD_DENSITY, astg.FIELD_HEAT):
        for l in range(levels):
            nn, _ = c.level_info(l)
            c.set_scalar(f, l, rng.random(nn), rng.random(nn))
    c.post_project()
    return c


def _points(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.02, 0.98, size=(n, 3))


def test_noalloc_matches_allocating_entry_point():
    big = _scene(64, 3, 1)
    small = _scene(32, 2, 2)
    dt = 0.5 / 64
    # first call: many points on the larger scene; second: fewer on the smaller
    for ctx, n, seed in ((big, 200000, 3), (small, 5000, 4)):
        xyz = _points(n, seed)
        for what in (astg.SAMPLE_VELOCITY, astg.SAMPLE_MAC, astg.SAMPLE_BACKTRACE, astg.SAMPLE_DENSITY,
                     astg.SAMPLE_HEAT):
            print("noalloc check: n=%d what=%d" % (n, what), flush=True)
            ref = ctx.sample_points(what, xyz, dt)
            got = ctx.sample_points(what, xyz, dt, noalloc=True)
            assert got.shape == ref.shape
            assert np.array_equal(got, ref), (n, what)
    big.close()
    small.close()


def test_noalloc_limits():
    c = _scene(32, 2, 5)
    with pytest.raises(astg.AstgError) as e:
        c.sample_points(astg.SAMPLE_MAC, _points(MAX_POINTS + 1, 6), noalloc=True)
    assert e.value.code == astg.ASTG_ERR_ARG
    # exactly the limit is accepted
    out = c.sample_points(astg.SAMPLE_MAC, _points(MAX_POINTS, 7), noalloc=True)
    assert np.isfinite(out).all()
    c.close()
    # a context whose fields are not on the device yet: refused, nothing set up
    d = astg.Context(0)
    d.set_topology(scenes.smoke_64_split(32, 2), 32, 2, scenes.CLOSED_XY)
    with pytest.raises(astg.AstgError) as e:
        d.sample_points(astg.SAMPLE_MAC, _points(10, 8), noalloc=True)
    assert e.value.code == astg.ASTG_ERR_STATE
    d.close()
