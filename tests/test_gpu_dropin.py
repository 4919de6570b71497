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


def _run(args, cwd=None):
    env = dict(os.environ, OMP_NUM_THREADS="2")
    p = subprocess.run([EXE] + [str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       env=env, timeout=120, cwd=cwd)
    assert p.returncode == 0, p.stdout
    return p.stdout


@pytest.mark.parametrize("method,precond", [(1, 1), (2, 3)])
def test_fortran_dropin_dump_and_verbose_timelog(tmp_path, method, precond):
    """DUMPTYPE=BSR and TIMELOG=VERBOSE on both sides: hecmw_solve_33 writes the
    reference's dump files (hecmw_mat_dump before the solve, hecmw_mat_dump_solution
    after it; hecmw_solver_Iterative.f90:90,160), byte-identical for the matrix
    and right-hand side, and prints hecmw_output_flops' lines (:189-254)."""
    import numpy as np
    out = _run([6, method, precond, 0, 0, 0, 1], cwd=str(tmp_path))
    f = [l for l in out.splitlines() if l.startswith("DROPIN")][-1].split()
    assert int(f[4]) == int(f[5]) == 1 and float(f[8]) <= 1e-6
    for ext in ("bsr", "rhs"):
        a = (tmp_path / ("dump_matrix_1_0." + ext)).read_bytes()
        b = (tmp_path / ("dump_matrix_2_0." + ext)).read_bytes()
        assert a == b and len(a) > 0, ext
    xa = np.loadtxt(tmp_path / "dump_matrix_1_0.sol")
    xb = np.loadtxt(tmp_path / "dump_matrix_2_0.sol")
    assert np.linalg.norm(xa - xb) <= 1e-6 * np.linalg.norm(xa)
    lines = out.splitlines()
    mem = [l for l in lines if l.startswith("memory amount of coef. matrix")]
    assert len(mem) == 2 and mem[0] == mem[1], mem
    for key in ("matvec memory band width", "matvec FLOPs"):
        assert sum(1 for l in lines if l.startswith(key)) == 2, key
    assert sum(1 for l in lines if "### summary of linear solver" in l) == 2


@pytest.mark.parametrize("method,precond", [(1, 1), (1, 3)])
def test_fortran_dropin_newton_sequence(method, precond):
    """Six solves with new values each time and fistr1's recycle flags: after
    every call the recycle counters (Iarray 96-98) agree with the reference's,
    and the solutions are within 1e-6."""
    out = _run([7, method, precond, 0, 0, 0, 2])
    seq = [l.split() for l in out.splitlines() if l.startswith("DROPIN_SEQ")]
    assert len(seq) == 6
    for f in seq:
        assert f[2:5] == f[5:8], f
        assert float(f[8]) <= 1e-6, f
    assert [int(f[2]) for f in seq] == [0, 1, 2, 3, 0, 1]
