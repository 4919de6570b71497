"""frontistr_amd — MI355X solver for hecmw1's 3-DOF-per-node linear systems.

Python mirror of the reference's solver interface over the C ABI of
libhecmw33.so (include/hecmw33.h):

* :class:`HecMatrix` holds the arrays of ``hecmwST_matrix``
  (hecmw1/src/common/hecmw_util_f.F90:458-470) as numpy arrays, with
  ``Iarray``/``Rarray`` initialised as ``hecmw_mat_init`` does
  (hecmw1/src/solver/matrix/hecmw_matrix_misc.F90:171-212) and the same
  named accessors.
* :func:`read_dump` / :func:`write_dump` read and write hecmw_mat_dump's
  MM/CSR/BSR files (hecmw1/src/solver/matrix/hecmw_matrix_dump.f90).
* :func:`hecmw_solve_iterative` / :func:`hecmw_solve_krylov` /
  :func:`hecmw_matvec` / :func:`hecmw_precond_apply` /
  :func:`hecmw_rel_resid_L2` call the GPU library.

The library must be built (``python3 frontistr_amd/build.py``); nothing here
falls back to CPU arithmetic.
"""
from .hecmw import (  # noqa: F401
    HecMatrix,
    SolveResult,
    HecmwError,
    ERRORS,
    lib,
    lib_path,
    hecmw_solve_iterative,
    hecmw_solve_krylov,
    last_history,
    hecmw_matvec,
    hecmw_matvec_mpc,
    stream_read,
    set_mpc,
    hecmw_precond_apply,
    hecmw_rel_resid_L2,
    hecmw_ordering,
    gen_hex8,
    gen_hex8_structure,
    gen_tet10,
    hex8_coords,
    tet10_coords,
    set_coords,
    BatchSolver,
    solve_batch,
    gen_tet10_batch,
    LearnedNet,
    set_net,
    harvest_errors,
    rigid_labels,
    DeviceContext,
)
from .dump import read_dump, write_dump  # noqa: F401
