"""A NaN or infinity in the projection system fails the solve loudly.

The reference's infinity norm (std::max over |r_i|) passes over a NaN entry,
so a right-hand side holding NaN can read as a zero norm and the solve returns
at once with NaN in the solution. The device solvers carry a NaN through the
norm reductions instead and stop with ASTG_ERR_NOT_CONVERGED and a message
that says where the value appeared:
  * single solves (FP64 and mixed) on a golden system with one b entry set
    to NaN or to infinity;
  * a batch where one of three systems holds the NaN: the call fails and
    names that system;
  * a NaN in A: the first preconditioning breaks down (rho NaN), an error too;
  * the context solves a clean system afterwards exactly as a fresh one does.
"""
import numpy as np
import pytest

from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    g = load_golden(GOLDEN[0])
    return (np.asarray(g["A_rowstart"]), np.asarray(g["A_col"]), np.asarray(g["A_val"], np.float64),
            np.asarray(g["b"], np.float64))


def _expect_error(fn, *words):
    import astg
    with pytest.raises(astg.AstgError) as e:
        fn()
    assert e.value.code == astg.ASTG_ERR_NOT_CONVERGED, str(e.value)
    for w in words:
        assert w in str(e.value), str(e.value)
    print(str(e.value), flush=True)


@pytest.mark.parametrize("bad", [np.nan, np.inf])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nonfinite_rhs_fails_the_solve(system, bad, precision):
    import astg
    rs, col, val, b = system
    prec = astg.PREC_FP64 if precision == "fp64" else astg.PREC_MIXED
    prm = astg.default_params(tol=1e-6, max_iterations=3000, precision=prec)
    c = astg.Context(0)
    x0, r0 = c.pcg_solve_csr(rs, col, val, b, prm)
    bb = b.copy()
    bb[len(bb) // 3] = bad
    _expect_error(lambda: c.pcg_solve_csr(rs, col, val, bb, prm), "NaN or infinity", "right-hand side")
    # the context is usable afterwards and solves as before
    x1, r1 = c.pcg_solve_csr(rs, col, val, b, prm)
    assert r1.iterations == r0.iterations and r1.converged == 1
    assert np.array_equal(x1, x0)
    c.close()


def test_nonfinite_system_fails_the_batch(system):
    import astg
    rs, col, val, b = system
    bb = b.copy()
    bb[7] = np.nan
    for prec in (astg.PREC_FP64, astg.PREC_MIXED):
        prm = astg.default_params(tol=1e-6, max_iterations=3000, precision=prec)
        c = astg.Context(0)
        _expect_error(lambda: c.pcg_solve_batch([(rs, col, val, b), (rs, col, val, bb), (rs, col, val, 2 * b)], prm),
                      "system 1", "NaN or infinity")
        xs, res = c.pcg_solve_batch([(rs, col, val, b), (rs, col, val, 2 * b)], prm)
        assert all(r.converged == 1 for r in res)
        c.close()


def test_nan_matrix_entry_fails_the_solve(system):
    import astg
    rs, col, val, b = system
    vv = val.copy()
    vv[len(vv) // 2] = np.nan
    c = astg.Context(0)
    with pytest.raises(astg.AstgError) as e:
        c.pcg_solve_csr(rs, col, vv, b, astg.default_params(tol=1e-6, max_iterations=3000))
    assert e.value.code == astg.ASTG_ERR_NOT_CONVERGED
    c.close()
