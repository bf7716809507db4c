"""Octree topology, tilt state, numbering, weights and CSR pattern against the
reference's own outputs (golden fixtures written by tests/golden/make_golden.py).

All comparisons are bit-exact: these are integer tables and exactly
representable spacings."""
import numpy as np

from conftest import make_context


def test_layers_and_tilt_state(golden):
    c = make_context(golden, -1)
    for l in range(int(golden["levels"])):
        xyz, fl = c.nodes(l)
        assert np.array_equal(xyz, golden["nodes_%d" % l]), "node order, level %d" % l
        assert np.array_equal(fl, golden["flags_%d" % l]), "ghost flags, level %d" % l
        assert c.level_info(l)[1] == int(golden["nleaf_%d" % l])
        e, bits, hd = c.tilt(l)
        assert np.array_equal(e, golden["tilt_e_%d" % l]), "tilt e, level %d" % l
        assert np.array_equal(bits, golden["tilt_bits_%d" % l].astype(np.int32)), "tilt flags, level %d" % l
        assert np.array_equal(hd, golden["tilt_hd_%d" % l].astype(np.int32)), "half direction, level %d" % l


def test_half_tilt_list(golden):
    c = make_context(golden, -1)
    assert np.array_equal(c.half_tilts(), golden["half"].astype(np.int32))


def test_numbering(golden):
    c = make_context(golden, -1)
    for l in range(int(golden["levels"])):
        o, t, n = c.numbering(l)
        assert n == len(golden["b"])
        assert np.array_equal(o, golden["pid_oct_%d" % l]), "octagon numbering, level %d" % l
        assert np.array_equal(t, golden["pid_tilt_%d" % l]), "tilt numbering, level %d" % l


def test_cutting_cell_weights(golden):
    c = make_context(golden, -1)
    for l in range(int(golden["levels"])):
        m, t = c.weights(l)
        assert np.array_equal(m, golden["w_mac_%d" % l])
        assert np.array_equal(t, golden["w_tilt_%d" % l])


def test_matrix_pattern(golden):
    c = make_context(golden, -1)
    rs, col = c.pattern()
    assert np.array_equal(rs, golden["A_rowstart"])
    assert np.array_equal(col, golden["A_col"])


def test_matrix_pattern_after_rebuild(golden):
    """A tilt change rebuilds the tables into the previous tables' memory: the
    pattern after a rebuild equals the fresh build's (the golden pattern)."""
    import astg
    c = astg.Context(-1)
    R, L = int(golden["R"]), int(golden["levels"])
    c.set_topology(golden["split"].astype(np.int32), R, L, int(golden["closed"]))
    _, nleaf = c.level_info(0)
    rng = np.random.default_rng(3)
    c.set_tilt_open(0, (rng.random(nleaf) < 0.5).astype(np.uint8))
    c.pattern()  # first build, with a different tilt state (and size)
    final = golden["open_mask_l0"] if int(golden["open_tilts"]) else np.zeros(nleaf, np.uint8)
    c.set_tilt_open(0, final)
    rs, col = c.pattern()
    assert np.array_equal(rs, golden["A_rowstart"])
    assert np.array_equal(col, golden["A_col"])
