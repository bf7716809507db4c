"""The oracle (CPU restatement, oracle/oracle.cpp) pinned against the
reference's own outputs, and against mathematics where those leave room.

* golden fixtures (tests/golden/*.npz, written by the reference): assembled A
  and b, MIC(0) factor and its application, PCG solution and iteration count,
  corrected velocities;
* closed forms: a single-level grid gives h times the 7-point Laplacian with
  Dirichlet rows on open walls; MIC(0) with tau = 0 on a tridiagonal matrix is
  the exact Cholesky factor; tau = 1 preserves row sums (L L^T 1 = A 1);
* a dense solve (numpy) of small systems for the PCG solution;
* a planted mistake (a flipped sign in b) is caught by the velocity check.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import ref


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def max_rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def run_oracle(g):
    lv = orc.levels_from_golden(g)
    return orc.project(int(g["R"]), int(g["levels"]), int(g["closed"]), lv, g["half"].astype(np.int32),
                       len(g["b"]), float(g["dt"]), float(g["rho"]), float(g["tol"]))


def test_oracle_assembly_matches_reference(golden):
    r = run_oracle(golden)
    assert np.array_equal(r["rowstart"], golden["A_rowstart"]) and np.array_equal(r["col"], golden["A_col"])
    assert max_rel(r["val"], golden["A_val"]) < 1e-12
    assert max_rel(r["b"], golden["b"]) < 1e-12


def test_oracle_solution_matches_reference(golden):
    r = run_oracle(golden)
    assert abs(r["iters"] - int(golden["pcg_iters"])) <= 1
    p, pr = r["p"], golden["p"]
    if int(golden["closed"]) == 63:
        p, pr = p - p.mean(), pr - pr.mean()
    assert rel(p, pr) < 1e-6


def test_oracle_velocity_matches_reference(golden):
    r = run_oracle(golden)
    for l, (um, ut) in enumerate(r["u"]):
        assert rel(um, golden["u_mac_out_%d" % l]) < 1e-6
        assert rel(ut, golden["u_tilt_out_%d" % l]) < 1e-6


def test_oracle_mic_matches_reference(golden):
    inv, cs, ri, v = orc.mic_factor(golden["A_rowstart"], golden["A_col"], golden["A_val"])
    assert np.array_equal(cs, golden["mic_colstart"]) and np.array_equal(ri, golden["mic_rowindex"])
    assert max_rel(inv, golden["mic_invdiag"]) < 1e-12
    assert max_rel(v, golden["mic_value"]) < 1e-12
    z = orc.mic_apply(golden["A_rowstart"], golden["A_col"], golden["A_val"], golden["mic_apply_r"])
    assert rel(z, golden["mic_apply_z"]) < 1e-10


def _csr(dense):
    rs, col, val = [0], [], []
    for i in range(dense.shape[0]):
        nz = np.nonzero(dense[i])[0]
        col += list(nz)
        val += list(dense[i, nz])
        rs.append(len(col))
    return np.array(rs, np.int32), np.array(col, np.int32), np.array(val)


def test_mic_tau0_is_exact_cholesky_on_tridiagonal():
    n = 40
    A = np.diag(np.full(n, 2.5)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    rs, col, val = _csr(A)
    inv, cs, ri, v = orc.mic_factor(rs, col, val, tau=0.0)
    Lc = np.linalg.cholesky(A)
    assert np.allclose(1.0 / inv, np.diag(Lc), rtol=1e-13)
    for k in range(n):
        for p in range(cs[k], cs[k + 1]):
            assert abs(v[p] - Lc[ri[p], k]) < 1e-13


def test_mic_tau1_preserves_row_sums():
    rng = np.random.default_rng(3)
    m = 7
    n = m * m
    A = np.zeros((n, n))
    for i in range(m):
        for j in range(m):
            r = i * m + j
            for di, dj in ((1, 0), (0, 1)):
                if i + di < m and j + dj < m:
                    w = rng.uniform(0.5, 2)
                    s = (i + di) * m + j + dj
                    A[r, s] = A[s, r] = -w
    A += np.diag(-A.sum(1) + rng.uniform(0.01, 0.1, n))
    rs, col, val = _csr(A)
    inv, cs, ri, v = orc.mic_factor(rs, col, val, tau=1.0, ratio=0.0)
    Ld = np.diag(1.0 / inv)
    for k in range(n):
        for p in range(cs[k], cs[k + 1]):
            Ld[ri[p], k] = v[p]
    ones = np.ones(n)
    assert np.allclose(Ld @ (Ld.T @ ones), A @ ones, rtol=1e-10, atol=1e-12)


def test_pcg_matches_dense_solve():
    rng = np.random.default_rng(5)
    m = 8
    n = m ** 3
    A = np.zeros((n, n))
    idx = lambda x, y, z: (z * m + y) * m + x
    for z in range(m):
        for y in range(m):
            for x in range(m):
                r = idx(x, y, z)
                for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    q = (x + d[0], y + d[1], z + d[2])
                    if max(q) < m:
                        s = idx(*q)
                        A[r, s] = A[s, r] = -1.0
    A += np.diag(-A.sum(1) + 0.05)
    b = rng.standard_normal(n)
    rs, col, val = _csr(A)
    x, res, it, ok = orc.pcg(rs, col, val, b, tol=1e-12)
    assert ok and rel(x, np.linalg.solve(A, b)) < 1e-9


def test_single_level_assembly_is_scaled_laplacian():
    """One level, no tilts: octagon rows are h * 7-point Laplacian; open walls
    add h to the diagonal (area/h with area = h^2)."""
    import astg
    R = 8
    split = np.zeros(R ** 3, np.int32)
    c = astg.Context(-1)
    c.set_topology(split, R, 1, astg.WALL_XNEG | astg.WALL_XPOS)
    lv, n = orc.levels_from_context(c, [(np.zeros((c.level_info(0)[0], 3)), np.zeros((c.level_info(0)[0], 8)))])
    r = orc.project(R, 1, c.closed, lv, c.half_tilts(), n, 0.1, 2.0)
    h = 1.0 / R
    m = R - 1  # octagons 0..R-2
    A = np.zeros((n, n))
    for i in range(n):
        for p in range(r["rowstart"][i], r["rowstart"][i + 1]):
            A[i, r["col"][p]] = r["val"][p]
    idx = lambda x, y, z: (z * m + y) * m + x
    E = np.zeros((n, n))
    for z in range(m):
        for y in range(m):
            for x in range(m):
                i = idx(x, y, z)
                diag = 0.0
                for d, closed in (((1, 0, 0), True), ((-1, 0, 0), True), ((0, 1, 0), False), ((0, -1, 0), False),
                                  ((0, 0, 1), False), ((0, 0, -1), False)):
                    q = (x + d[0], y + d[1], z + d[2])
                    if 0 <= min(q) and max(q) < m:
                        E[i, idx(*q)] = -h
                        diag += h
                    elif not closed:
                        diag += h
                E[i, i] = diag
    assert np.allclose(A, E, rtol=1e-14, atol=1e-15)


def test_planted_sign_error_is_caught(golden):
    """A flipped sign in the oracle's right-hand side changes the corrected
    velocities well beyond the tolerance, so the velocity check would fail."""
    lv = orc.levels_from_golden(golden)
    for d in lv:
        d["u_mac"] = -d["u_mac"]
        d["u_tilt"] = -d["u_tilt"]
    r = orc.project(int(golden["R"]), int(golden["levels"]), int(golden["closed"]), lv,
                    golden["half"].astype(np.int32), len(golden["b"]), float(golden["dt"]), float(golden["rho"]))
    assert rel(r["b"], golden["b"]) > 1.0


@pytest.mark.skipif(not ref.available(), reason="reference build (oracle/_ref) not present")
def test_oracle_pcg_matches_reference_build_on_fixture_matrices(golden):
    x, res, it, ok = ref.pcg_csr(golden["A_rowstart"], golden["A_col"], golden["A_val"], golden["b"])
    xo, reso, ito, oko = orc.pcg(golden["A_rowstart"], golden["A_col"], golden["A_val"], golden["b"])
    assert ok == oko and abs(it - ito) <= 1
