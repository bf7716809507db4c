"""Split maps from distance fields (astg_split_distance), host only.

A split map holds, per finest cell (x fastest), the level of the leaf that
covers it. Checked by brute force on small grids:
  * well formed: every level-l leaf covers an aligned 2^l block whose cells
    all hold l;
  * balanced: the 26 same-level neighbour blocks of every level-l leaf hold
    only levels l-1, l and l+1 (the condition the layer construction needs);
  * a distance field beyond every band gives the coarsest map, one cell at
    distance 0 gives a finest leaf there;
  * wider bands never coarsen a cell;
  * a field on a lattice 2^s cells wide (shift s) gives the same map as the
    same field repeated over the finest cells with shift 0;
  * the full-scale initial layout is finest inside both sources.
"""
import numpy as np
import pytest


def _astg():
    import astg
    return astg


def leaf_ok(split, R, levels):
    s = split.reshape(R, R, R)
    for l in range(levels):
        b = 1 << l
        blocks = s.reshape(R // b, b, R // b, b, R // b, b)
        first = blocks[:, 0, :, 0, :, 0]
        owns = first == l
        same = (blocks == first[:, None, :, None, :, None]).all(axis=(1, 3, 5))
        if not same[owns].all():
            return False
    return True


def balanced(split, R, levels):
    s = split.reshape(R, R, R)
    for l in range(levels):
        b = 1 << l
        n = R // b
        lo = s.reshape(n, b, n, b, n, b).min(axis=(1, 3, 5))
        hi = s.reshape(n, b, n, b, n, b).max(axis=(1, 3, 5))
        own = (s.reshape(n, b, n, b, n, b)[:, 0, :, 0, :, 0] == l) & (lo == l) & (hi == l)
        for z, y, x in np.argwhere(own):
            z0, z1 = max(z - 1, 0), min(z + 2, n)
            y0, y1 = max(y - 1, 0), min(y + 2, n)
            x0, x1 = max(x - 1, 0), min(x + 2, n)
            if lo[z0:z1, y0:y1, x0:x1].min() < l - 1 or hi[z0:z1, y0:y1, x0:x1].max() > l + 1:
                return False
    return True


R, L = 32, 3
BAND = np.array([0.1, 0.2])


def test_far_field_is_coarsest_and_a_point_is_finest():
    astg = _astg()
    far = np.full(R ** 3, 10.0, np.float32)
    s = astg.split_distance(R, L, far, BAND)
    assert (s == L - 1).all()
    d = np.full((R, R, R), 10.0, np.float32)
    d[5, 20, 9] = 0.0
    s = astg.split_distance(R, L, d.ravel(), BAND)
    assert s.reshape(R, R, R)[5, 20, 9] == 0
    assert leaf_ok(s, R, L) and balanced(s, R, L)


def test_random_fields_are_well_formed_balanced_and_monotone_in_band():
    astg = _astg()
    rng = np.random.default_rng(3)
    c = (np.arange(R) + 0.5) / R
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    for trial in range(4):
        p = rng.uniform(0.1, 0.9, 3)
        d = (np.sqrt((X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2) - rng.uniform(0.02, 0.1)).astype(np.float32)
        s1 = astg.split_distance(R, L, d.ravel(), BAND)
        s2 = astg.split_distance(R, L, d.ravel(), 2 * BAND)
        assert leaf_ok(s1, R, L) and balanced(s1, R, L)
        assert leaf_ok(s2, R, L) and balanced(s2, R, L)
        assert (s2 <= s1).all()
        assert (s1.reshape(R, R, R)[d <= 0] == 0).all()


@pytest.mark.parametrize("shift", [1, 2])
def test_coarse_lattice_equals_repeated_field(shift):
    astg = _astg()
    rng = np.random.default_rng(shift)
    n = R >> shift
    c = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    p = rng.uniform(0.2, 0.8, 3)
    d = (np.sqrt((X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2) - 0.05).astype(np.float32)
    b = 1 << shift
    fine = np.repeat(np.repeat(np.repeat(d, b, axis=0), b, axis=1), b, axis=2)
    a = astg.split_distance(R, L, d.ravel(), BAND, shift)
    e = astg.split_distance(R, L, fine.ravel(), BAND, 0)
    assert np.array_equal(a, e)


def test_shift_and_size_are_checked():
    astg = _astg()
    with pytest.raises(astg.AstgError):
        astg.split_distance(R, L, np.zeros(R ** 3 // 8 ** 3, np.float32), BAND, L)
    with pytest.raises(ValueError):
        astg.split_distance(R, L, np.zeros(R ** 3 // 8, np.float32), BAND, 0)


def test_full_scale_initial_layout_is_finest_in_the_sources():
    from astg import scenes
    Rf, Lf = 64, 3
    s = scenes.full_scale_initial_split(Rf, Lf).reshape(Rf, Rf, Rf)
    assert leaf_ok(s.ravel(), Rf, Lf) and balanced(s.ravel(), Rf, Lf)
    c = (np.arange(Rf) + 0.5) / Rf
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    for cx, cy, cz, r, _ in scenes.FULL_SOURCES:
        inside = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 < r * r
        assert inside.any() and (s[inside] == 0).all()
