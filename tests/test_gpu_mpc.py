"""MPC-CG (hecMESH%mpc with MPC_METHOD = 2): every product of the solve is
T^T A T (hecmw_TtmatTvec_33, hecmw1/src/solver/las/hecmw_solver_las_33.F90:
442-586).  Against the reference build:

* the GPU product is bitwise the reference's (single-threaded: its Ttvec
  scatter then runs in constraint order, which the GPU's gather reproduces);
* hecmw_solve_iterative with MPC_METHOD = 2 on the transformed right-hand
  side T^T b (what fistr1 passes) gives the reference's solution within 1e-6
  and the same convergence flag, for CG and BiCGSTAB with SSOR and DIAG;
* chained constraints are refused (the reference's parallel loops assume none).
"""
import numpy as np
import pytest

import frontistr_amd as fa
from oracle import ref
from tests.mpc_systems import dense_T, tie_constraints

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]


def test_mpc_product_bitwise():
    m = fa.gen_hex8(5, kind="lognormal", seed=1)
    mpc = tie_constraints(5)
    x = np.random.default_rng(2).standard_normal(3 * m.NP)
    fa.set_mpc(m.N, mpc)
    ref.set_mpc(mpc)
    try:
        y = fa.hecmw_matvec_mpc(m, x)
        yr = ref.matvec_mpc(m, x, nthreads=1)
    finally:
        fa.set_mpc(m.N, None)
        ref.set_mpc(None)
    assert np.array_equal(y, yr), np.abs(y - yr).max()


@pytest.mark.parametrize("method,precond", [(1, 1), (1, 3), (2, 1), (2, 3)])
def test_mpc_solve_matches_reference(method, precond):
    n = 6
    m = fa.gen_hex8(n, kind="lognormal", seed=3)
    mpc = tie_constraints(n)
    T = dense_T(m.N, mpc)
    m.B[:] = T.T @ m.B
    m.set_iter(5000)
    m.set_method(method)
    m.set_precond(precond)
    m.set_i(13, 2)  # MPC_METHOD
    ref.set_mpc(mpc)
    fa.set_mpc(m.N, mpc)
    try:
        xr, iout, _ = ref.solve_iterative(m.copy(), nthreads=2)
        g = m.copy()
        res = fa.hecmw_solve_iterative(g)
    finally:
        fa.set_mpc(m.N, None)
        ref.set_mpc(None)
    assert res.error == 0
    assert iout[80] == 1 and g.Iarray[80] == 1  # converged flags (Iarray 81)
    assert iout[82] == 0 and g.Iarray[82] == 0  # the MPC flag is cleared again
    assert np.linalg.norm(g.X - xr) <= 1e-6 * np.linalg.norm(xr)
    # the constrained displacement T x satisfies every tie
    u = T @ g.X
    assert np.allclose(u, T @ u, atol=1e-12 * np.abs(u).max())


def test_chained_constraints_refused():
    m = fa.gen_hex8(3)
    # node 2's x is tied to node 3's x, and node 3's x to node 4's: a chain
    mpc = (np.array([0, 2, 4], np.int32), np.array([2, 3, 3, 4], np.int32), np.array([1, 1, 1, 1], np.int32),
           np.array([1.0, -1.0, 1.0, -1.0]))
    with pytest.raises(fa.HecmwError):
        fa.set_mpc(m.N, mpc)
