"""The Fortran drop-in: fistr1-style Fortran code calls hecmw_solve_33
(frontistr_amd/fortran/hecmw_solver_33.f90) on a hecmwST_matrix and gets the
solution and convergence flags the reference's own hecmw_solve_iterative gives
(oracle/_ref/dropin_test links both; tests/fortran/dropin_driver.f90)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_test")

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in test binary not built")]


@pytest.mark.parametrize("n,method,precond,scaling,fcg,mpc", [
    (6, 1, 1, 0, 0, 0), (6, 1, 3, 0, 0, 0), (5, 2, 1, 0, 0, 0), (5, 2, 3, 0, 0, 0), (9, 1, 1, 0, 0, 0),
    (6, 1, 1, 1, 0, 0), (5, 2, 3, 1, 0, 0), (12, 1, 1, 0, 1, 0), (6, 1, 1, 0, 0, 1), (6, 2, 3, 0, 0, 1)])
def test_fortran_dropin_matches_reference(n, method, precond, scaling, fcg, mpc):
    """fcg = 1: the GPU side runs the mixed-precision flexible CG with the
    two-level correction, coordinates taken from hecMESH%node by the shim.
    mpc = 1: multipoint constraints in hecMESH%mpc with MPC_METHOD = 2, handed
    to the library by the shim (T^T A T products on both sides)."""
    env = dict(os.environ, OMP_NUM_THREADS="2")
    p = subprocess.run([EXE, str(n), str(method), str(precond), str(scaling), str(fcg), str(mpc)],
                       stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, env=env, timeout=120)
    line = [l for l in p.stdout.splitlines() if l.startswith("DROPIN")]
    assert p.returncode == 0 and line, p.stdout
    f = line[-1].split()
    conv_ref, conv_gpu, div_ref, div_gpu, relerr = int(f[4]), int(f[5]), int(f[6]), int(f[7]), float(f[8])
    assert conv_ref == conv_gpu == 1
    assert div_ref == div_gpu == 0
    assert relerr <= 1e-6
