"""ctypes loader of oracle/_build/librestate.so, the plain-C restatement of
hecmw1's 3x3-block CG/BiCGSTAB with SSOR/DIAG (oracle/restate.c).

Test infrastructure only (the checker): used by tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg; the product package never imports it.
Same call shapes as oracle/ref.py, so a test can run either oracle.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "librestate.so")
_lib = None


class _Sys(C.Structure):
    _fields_ = [("N", C.c_int32), ("NP", C.c_int32), ("D", C.c_void_p), ("AL", C.c_void_p), ("AU", C.c_void_p),
                ("indexL", C.c_void_p), ("itemL", C.c_void_p), ("indexU", C.c_void_p), ("itemU", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        I = C.c_int32
        L.rs_ordering_rcm.argtypes = [C.POINTER(_Sys), P, P]
        L.rs_ordering_rcm.restype = None
        L.rs_ordering_mc.argtypes = [C.POINTER(_Sys), P, I, P, P, P]
        L.rs_ordering_mc.restype = I
        L.rs_matvec.argtypes = [C.POINTER(_Sys), P, P]
        L.rs_matvec.restype = None
        L.rs_precond.argtypes = [C.POINTER(_Sys), I, I, C.c_double, I, P, P, C.POINTER(I)]
        L.rs_precond.restype = I
        for f in (L.rs_solve_cg, L.rs_solve_bicgstab):
            f.argtypes = [C.POINTER(_Sys), P, P, P, P, C.POINTER(I), C.POINTER(C.c_double), P]
            f.restype = I
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _sys(m):
    keep = [m.D, m.AL if m.AL.size else np.zeros(9), m.AU if m.AU.size else np.zeros(9), m.indexL,
            m.itemL if m.itemL.size else np.ones(1, np.int32), m.indexU,
            m.itemU if m.itemU.size else np.ones(1, np.int32)]
    s = _Sys(m.N, m.NP, *[_p(a) for a in keep])
    return s, keep


def ordering(m, ncolor_in=10):
    """(perm_rcm, perm, iperm, colorindex), 1-based ids."""
    s, keep = _sys(m)
    N = m.N
    pr = np.zeros(N, np.int32)
    ipr = np.zeros(N, np.int32)
    lib().rs_ordering_rcm(C.byref(s), _p(pr), _p(ipr))
    p = np.zeros(N, np.int32)
    ip = np.zeros(N, np.int32)
    ci = np.zeros(N + 1, np.int32)
    nc = lib().rs_ordering_mc(C.byref(s), _p(pr), ncolor_in, _p(ci), _p(p), _p(ip))
    return pr, p, ip, ci[: nc + 1].copy()


def matvec(m, x):
    s, keep = _sys(m)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(3 * m.N)
    lib().rs_matvec(C.byref(s), _p(x), _p(y))
    return y


def precond(m, r, precond=None, iterpremax=None, sigma=None, ncolor_in=None):
    s, keep = _sys(m)
    pc = m.get_i(m.I_PRECOND) if precond is None else precond
    ipm = m.get_i(m.I_ITERPREMAX) if iterpremax is None else iterpremax
    sg = float(m.Rarray[m.R_SIGMA_DIAG - 1]) if sigma is None else sigma
    nci = m.get_i(m.I_NCOLOR_IN) if ncolor_in is None else ncolor_in
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros(3 * m.NP)
    nc = C.c_int32(0)
    e = lib().rs_precond(C.byref(s), pc, ipm, sg, nci, _p(r), _p(z), C.byref(nc))
    if e:
        raise ValueError("rs_precond error %d" % e)
    return z[: 3 * m.N]


def solve_krylov(m, x0=None):
    """CG (METHOD 1) or BiCGSTAB (METHOD 2) as m.Iarray/Rarray select.
    Returns (x, info) with info = dict(iter, resid, error, history)."""
    s, keep = _sys(m)
    x = (np.zeros(3 * m.NP) if x0 is None else np.array(x0, dtype=np.float64)).copy()
    b = np.ascontiguousarray(m.B, dtype=np.float64)
    maxit = m.get_i(m.I_ITER)
    params = np.array([maxit, m.get_i(m.I_PRECOND), m.get_i(m.I_ITERPREMAX), m.get_i(m.I_NCOLOR_IN)], np.int32)
    rparams = np.array([m.Rarray[m.R_RESID - 1], m.Rarray[m.R_SIGMA_DIAG - 1]])
    hist = np.zeros(max(1, maxit))
    it = C.c_int32(0)
    res = C.c_double(0.0)
    f = lib().rs_solve_bicgstab if m.get_i(m.I_METHOD) == 2 else lib().rs_solve_cg
    err = f(C.byref(s), _p(b), _p(x), _p(params), _p(rparams), C.byref(it), C.byref(res), _p(hist))
    n = min(it.value, maxit)
    return x, dict(iter=it.value, resid=res.value, error=int(err), history=hist[:n].copy())
