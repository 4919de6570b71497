"""Loader for oracle/_ref/libhecmw_ref.so: the reference's own hecmw1 solver
compiled from its sources by oracle/build_ref.py.

Test infrastructure only (the checker): imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
product package.

The reference keeps module-level state between calls (the SSOR factorisation,
the SpMV thread partition); every function here forces a fresh preconditioner
setup unless told otherwise.  Caveat of the reference: its SpMV keeps the
previous thread partition (hecmw_solver_las_33.F90:164-171) whenever the last
block ends at N-1, which is always the case for systems with fewer rows than
100 x threads, so a change of the OpenMP thread count between two systems of
the same N leaves rows uncomputed.  Every call here therefore uses the same
thread count (NTHREADS) unless a caller deliberately passes another one (the
CPU baseline, in its own process).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libhecmw_ref.so")
_lib = None
NTHREADS = 2  # >= 2 selects the RCM + multicolour SSOR ordering (hecmw_precond_SSOR_33.F90:115-127)


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref/libhecmw_ref.so missing: run python3 oracle/build_ref.py")
        L = C.CDLL(LIB)
        P = C.c_void_p
        I = C.c_int32
        L.ref_solve_krylov.argtypes = [I, I, I, I] + [P] * 9 + [P, P, I, P, P]
        L.ref_solve_krylov.restype = I
        L.ref_solve_iterative.argtypes = [I, I, I, I] + [P] * 9 + [P, P, I, P, P]
        L.ref_solve_iterative.restype = I
        L.ref_ordering.argtypes = [I, P, P, P, P, I, P, P, C.POINTER(I), P, P, P]
        L.ref_ordering.restype = None
        L.ref_matvec.argtypes = [I, I, I, I] + [P] * 9 + [I]
        L.ref_matvec.restype = None
        L.ref_precond.argtypes = [I, I, I, I] + [P] * 9 + [P, P, I]
        L.ref_precond.restype = None
        L.ref_rel_resid.argtypes = [I, I, I, I] + [P] * 9 + [I]
        L.ref_rel_resid.restype = C.c_double
        L.ref_matvec_mpc.argtypes = [I, I, I, I] + [P] * 9 + [I]
        L.ref_matvec_mpc.restype = None
        L.ref_set_mpc.argtypes = [I, P, P, P, P]
        L.ref_set_mpc.restype = None
        L.ref_mat_dump.argtypes = [I, I, I, I] + [P] * 9 + [I]
        L.ref_mat_dump.restype = None
        L.ref_flush.argtypes = []
        L.ref_flush.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _mat_args(m):
    # keep at least one element behind every pointer (empty AL/AU on 1-row systems)
    AL = m.AL if m.AL.size else np.zeros(9)
    AU = m.AU if m.AU.size else np.zeros(9)
    tL = m.itemL if m.itemL.size else np.ones(1, np.int32)
    tU = m.itemU if m.itemU.size else np.ones(1, np.int32)
    return [m.N, m.NP, m.NPL, m.NPU, _p(m.D), _p(AL), _p(AU), _p(m.indexL), _p(tL), _p(m.indexU), _p(tU)]


def _capture_stdout(fn):
    """Run fn() with file descriptor 1 redirected to a temporary file; return
    (fn's result, captured text).  Used to read the reference's ITERLOG lines."""
    import sys
    import tempfile
    sys.stdout.flush()
    lib().ref_flush()  # what earlier calls left in Fortran's buffer goes to the real stdout
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tf:
        os.dup2(tf.fileno(), 1)
        try:
            res = fn()
            lib().ref_flush()
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tf.seek(0)
        text = tf.read().decode(errors="replace")
    return res, text


def solve_krylov(m, x0=None, nthreads=NTHREADS, iarray=None, rarray=None, history=False):
    """hecmw_solve_CG / hecmw_solve_BiCGSTAB on a copy of the system.
    Returns (x, info) with info = dict(iter, resid, error, Tset, Tsol, Tcomm, time_Ax, time_precond, iarray
    [, history]).  history=True turns on ITERLOG and returns the printed
    per-iteration relative residuals (format (i7,1pe16.6), hecmw_solver_CG.f90:253)."""
    x = (np.zeros(3 * m.NP) if x0 is None else np.array(x0, dtype=np.float64)).copy()
    b = m.B.copy()
    ia = np.array(m.Iarray if iarray is None else iarray, dtype=np.int32).copy()
    ra = np.array(m.Rarray if rarray is None else rarray, dtype=np.float64).copy()
    ia[96] = 1  # Iarray(97): numeric setup
    ia[97] = 1  # Iarray(98): symbolic setup
    if history:
        ia[20] = 1  # Iarray(21): ITERLOG
    iout = np.zeros(100, np.int32)
    out = np.zeros(7)

    def run():
        return lib().ref_solve_krylov(*_mat_args(m), _p(b), _p(x), _p(ia), _p(ra), nthreads, _p(iout), _p(out))

    if history:
        err, text = _capture_stdout(run)
    else:
        err, text = run(), ""
    info = dict(iter=int(out[0]), resid=float(out[1]), error=int(err), Tset=float(out[2]), Tsol=float(out[3]),
                Tcomm=float(out[4]), time_Ax=float(out[5]), time_precond=float(out[6]), iarray=iout)
    if history:
        hist = {}
        for line in text.splitlines():
            parts = line.split()
            if len(parts) == 2 and parts[0].isdigit():
                hist[int(parts[0])] = float(parts[1])
        info["history"] = np.array([hist[k] for k in sorted(hist)])
    return x, info


def solve_iterative(m, x0=None, nthreads=NTHREADS, iarray=None, rarray=None):
    """hecmw_solve_iterative (full dispatcher).  Returns (x, iarray_out, rarray_out).
    iarray/rarray (default m's) are passed as they are, so a sequence of calls
    can carry the recycle counters (Iarray 96-98) the way fistr1 does; the
    reference keeps its preconditioner between calls unless they ask for a
    new one."""
    x = (np.zeros(3 * m.NP) if x0 is None else np.array(x0, dtype=np.float64)).copy()
    b = m.B.copy()
    ia = np.array(m.Iarray if iarray is None else iarray, dtype=np.int32).copy()
    ra = np.array(m.Rarray if rarray is None else rarray, dtype=np.float64).copy()
    iout = np.zeros(100, np.int32)
    rout = np.zeros(100)
    lib().ref_solve_iterative(*_mat_args(m), _p(b), _p(x), _p(ia), _p(ra), nthreads, _p(iout), _p(rout))
    return x, iout, rout


def solve_iterative_iter(m, x0=None, nthreads=NTHREADS, iarray=None, rarray=None):
    """solve_iterative with TIMELOG=YES, returning (x, iarray_out, rarray_out,
    iterations) where the iterations are read from the summary the reference
    prints ('ITER iterations RESID', hecmw_solver_Iterative.f90:173)."""
    ia = np.array(m.Iarray if iarray is None else iarray, dtype=np.int32).copy()
    ia[21] = 1  # Iarray(22): TIMELOG=YES
    (x, iout, rout), text = _capture_stdout(lambda: solve_iterative(m, x0, nthreads, ia, rarray))
    iters = [int(l.split()[0]) for l in text.splitlines() if len(l.split()) == 3 and l.split()[1] == "iterations"]
    if not iters:
        raise RuntimeError("no iteration summary in the reference's output:\n" + text[-2000:])
    iout[21] = 0 if iarray is None else np.asarray(iarray)[21]
    return x, iout, rout, iters[-1]


def ordering(m, ncolor_in=10):
    """(perm_rcm, iperm_rcm, perm, iperm, colorindex) from the reference routines, 1-based."""
    N = m.N
    pr = np.zeros(N, np.int32)
    ipr = np.zeros(N, np.int32)
    p = np.zeros(N, np.int32)
    ip = np.zeros(N, np.int32)
    ci = np.zeros(N + 1, np.int32)
    nc = C.c_int32(0)
    tL = m.itemL if m.itemL.size else np.ones(1, np.int32)
    tU = m.itemU if m.itemU.size else np.ones(1, np.int32)
    lib().ref_ordering(N, _p(m.indexL), _p(tL), _p(m.indexU), _p(tU), ncolor_in, _p(pr), _p(ipr), C.byref(nc),
                       _p(ci), _p(p), _p(ip))
    return pr, ipr, p, ip, ci[: nc.value + 1].copy()


def matvec(m, x, nthreads=NTHREADS):
    x = np.array(x, dtype=np.float64)
    y = np.zeros(3 * m.NP)
    lib().ref_matvec(*_mat_args(m), _p(x), _p(y), nthreads)
    return y[: 3 * m.N]


def precond(m, r, nthreads=NTHREADS, iarray=None, rarray=None):
    r = np.array(r, dtype=np.float64)
    z = np.zeros(3 * m.NP)
    ia = np.array(m.Iarray if iarray is None else iarray, dtype=np.int32).copy()
    ra = np.array(m.Rarray if rarray is None else rarray, dtype=np.float64).copy()
    lib().ref_precond(*_mat_args(m), _p(r), _p(z), _p(ia), _p(ra), nthreads)
    return z[: 3 * m.N]


def rel_resid(m, x, nthreads=NTHREADS):
    x = np.array(x, dtype=np.float64)
    b = m.B.copy()
    return lib().ref_rel_resid(*_mat_args(m), _p(b), _p(x), nthreads)


def mat_dump(m, kind, directory):
    """The reference's hecmw_mat_dump (DUMPTYPE kind: "mm", "csr" or "bsr") and
    hecmw_mat_dump_solution of m, run in `directory` (the Fortran writes
    dump_matrix_<call>_0.* into the current directory).  Returns the path of
    the matrix file; .rhs and .sol lie beside it."""
    import glob
    code = {"mm": 1, "csr": 2, "bsr": 3}[kind]
    b = m.B.copy()
    x = m.X.copy()
    here = os.getcwd()
    os.chdir(directory)
    try:
        before = set(glob.glob("dump_matrix_*." + kind))
        lib().ref_mat_dump(*_mat_args(m), _p(b), _p(x), code)
        new = sorted(set(glob.glob("dump_matrix_*." + kind)) - before)
    finally:
        os.chdir(here)
    if len(new) != 1:
        raise RuntimeError("reference dump wrote %s" % new)
    return os.path.join(directory, new[0])


_mpc_keep = None


def set_mpc(mpc):
    """Bind multipoint constraints for the reference calls that follow:
    mpc = (index[0:n+1], item, dof, val) as in hecMESH%mpc (1-based items and
    dofs), or None to clear."""
    global _mpc_keep
    if mpc is None:
        _mpc_keep = None
        lib().ref_set_mpc(0, None, None, None, None)
        return
    idx, item, dof, val = (np.ascontiguousarray(mpc[0], np.int32), np.ascontiguousarray(mpc[1], np.int32),
                           np.ascontiguousarray(mpc[2], np.int32), np.ascontiguousarray(mpc[3], np.float64))
    _mpc_keep = (idx, item, dof, val)
    lib().ref_set_mpc(len(idx) - 1, _p(idx), _p(item), _p(dof), _p(val))


def matvec_mpc(m, x, nthreads=NTHREADS):
    """y = T^T A T x (hecmw_TtmatTvec_33) with the constraints of set_mpc."""
    xin = np.array(x, dtype=np.float64)
    y = np.zeros(3 * m.NP)
    lib().ref_matvec_mpc(*_mat_args(m), _p(xin), _p(y), nthreads)
    return y
