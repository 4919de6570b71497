"""ctypes binding of libhecmw33.so and a numpy mirror of hecmwST_matrix."""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "lib", "libhecmw33.so")

# hecmw1/src/solver/init/hecmw_solve_error.f90:9-16
ERRORS = {
    1001: "INCONS_PC",
    2001: "ZERO_DIAG",
    2002: "ZERO_RHS",
    3001: "NOCONV_MAXIT",
    3002: "DIVERGE_MAT",
    3003: "DIVERGE_PC",
    3004: "DIVERGE_NAN",
}


class HecmwError(RuntimeError):
    """A library failure (negative return code); reference error codes are
    returned in SolveResult.error instead, as the reference returns them."""


class _Sys(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("NP", C.c_int32), ("NPL", C.c_int32), ("NPU", C.c_int32),
        ("D", C.c_void_p), ("AL", C.c_void_p), ("AU", C.c_void_p),
        ("indexL", C.c_void_p), ("indexU", C.c_void_p), ("itemL", C.c_void_p), ("itemU", C.c_void_p),
    ]


class _Out(C.Structure):
    _fields_ = [
        ("iter", C.c_int32), ("error", C.c_int32), ("method", C.c_int32), ("ncolor", C.c_int32),
        ("resid", C.c_double), ("resid_true", C.c_double),
        ("Tset", C.c_double), ("Tsol", C.c_double), ("Tcomm", C.c_double),
        ("time_Ax", C.c_double), ("time_precond", C.c_double), ("t_upload", C.c_double),
    ]


_lib = None


def lib():
    """Load libhecmw33.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path):
        raise HecmwError("libhecmw33.so is not built: run python3 frontistr_amd/build.py")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # soname), and a process that loads ROCm's first and torch's second holds
    # two runtimes, after which torch sees no device.  Binding torch's first
    # makes libhecmw33 resolve to it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(lib_path)
    P = C.c_void_p
    I = C.c_int32
    L.hecmw33_solve.argtypes = [C.POINTER(_Sys), P, P, P, P, C.POINTER(_Out)]
    L.hecmw33_solve_krylov.argtypes = [C.POINTER(_Sys), P, P, P, P, C.POINTER(_Out)]
    L.hecmw33_matvec.argtypes = [C.POINTER(_Sys), P, P]
    L.hecmw33_precond_apply.argtypes = [C.POINTER(_Sys), P, P, P, P]
    L.hecmw33_rel_resid_L2.argtypes = [C.POINTER(_Sys), P, P, C.POINTER(C.c_double)]
    L.hecmw33_rel_resid_L2_device_noalloc.argtypes = [C.POINTER(_Sys), P, P, P, P]
    L.hecmw33_ordering.argtypes = [I, P, P, P, P, I, P, P, P, C.POINTER(I), P]
    L.hecmw33_ctx_create.restype = P
    L.hecmw33_ctx_destroy.argtypes = [P]
    L.hecmw33_ctx_load.argtypes = [P, C.POINTER(_Sys), P, P]
    L.hecmw33_ctx_solve_device.argtypes = [P, P, P, P, P, C.POINTER(_Out)]
    L.hecmw33_ctx_matvec_device.argtypes = [P, P, P, I, C.POINTER(C.c_double)]
    L.hecmw33_ctx_matvec_device_f32.argtypes = [P, P, P, I, C.POINTER(C.c_double)]
    L.hecmw33_stream_read.argtypes = [P, C.c_int64, P, I, C.POINTER(C.c_double)]
    L.hecmw33_ctx_ssor_device.argtypes = [P, P, P, I, I, C.POINTER(C.c_double)]
    L.hecmw33_ctx_history.argtypes = [P, P, I]
    L.hecmw33_ctx_info.argtypes = [P, P, I]
    L.hecmw33_gen_hex8_count.argtypes = [I, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
    L.hecmw33_gen_hex8.argtypes = [I, I, P, C.c_uint64, P, P, P, P, P, P, P, P, P]
    L.hecmw33_gen_tet10_count.argtypes = [I, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
    L.hecmw33_gen_tet10.argtypes = [I, I, I, P, C.c_double, C.c_double, I, C.c_uint64, P, P, P, P, P, P, P, P]
    L.hecmw33_last_history.argtypes = [P, I]
    L.hecmw33_batch_create.restype = P
    L.hecmw33_batch_destroy.argtypes = [P]
    L.hecmw33_batch_load.argtypes = [P, I, P, I, P, P, P]
    L.hecmw33_batch_solve.argtypes = [P, P, P, P, P, P, P]
    L.hecmw33_batch_get_x.argtypes = [P, I, P]
    L.hecmw33_batch_spmm.argtypes = [P, I, P, P]
    L.hecmw33_batch_info.argtypes = [P, P, I]
    L.hecmw33_batch_variant_name.argtypes = [I]
    L.hecmw33_batch_variant_name.restype = C.c_char_p
    L.hecmw33_batch_variant64_name.argtypes = [I]
    L.hecmw33_batch_variant64_name.restype = C.c_char_p
    L.hecmw33_batch_last_error.restype = C.c_char_p
    L.hecmw33_solve_batch.argtypes = [I, P, I, P, P, P, P, P, P, P]
    L.hecmw33_net_create.argtypes = [C.c_uint64]
    L.hecmw33_net_create.restype = P
    L.hecmw33_net_destroy.argtypes = [P]
    L.hecmw33_net_get_weights.argtypes = [P, P, I]
    L.hecmw33_net_set_weights.argtypes = [P, P, I]
    L.hecmw33_net_eval.argtypes = [P, C.c_int64, P, P, P, P]
    L.hecmw33_net_add_sample.argtypes = [P, P, P, C.c_double]
    L.hecmw33_net_clear_samples.argtypes = [P]
    L.hecmw33_net_train.argtypes = [P, I, C.c_double, C.c_uint64, P]
    L.hecmw33_net_score.argtypes = [P, P, P]
    L.hecmw33_rigid_labels.argtypes = [I, P, P, P, P, P, I, P, C.c_double, P]
    L.hecmw33_ctx_set_net.argtypes = [P, P]
    L.hecmw33_set_net.argtypes = [P]
    L.hecmw33_set_coords.argtypes = [I, P]
    L.hecmw33_set_mpc.argtypes = [I, I, P, P, P, P]
    L.hecmw33_ctx_set_mpc.argtypes = [P, I, I, P, P, P, P]
    L.hecmw33_matvec_mpc.argtypes = [C.POINTER(_Sys), P, P]
    L.hecmw33_ctx_set_coords.argtypes = [P, I, P]
    L.hecmw33_ctx_coarse_get.argtypes = [P, P, P, P, P, P, P, P]
    L.hecmw33_ctx_modes_get.argtypes = [P, P, P, P]
    L.hecmw33_ctx_coarse_inverse_get.argtypes = [P, P, P, P, P]
    L.hecmw33_last_error.restype = C.c_char_p
    for f in ("hecmw33_solve", "hecmw33_solve_krylov", "hecmw33_matvec", "hecmw33_precond_apply",
              "hecmw33_rel_resid_L2", "hecmw33_ordering", "hecmw33_ctx_load", "hecmw33_ctx_solve_device",
              "hecmw33_ctx_matvec_device", "hecmw33_ctx_matvec_device_f32", "hecmw33_stream_read", "hecmw33_ctx_history",
              "hecmw33_ctx_info", "hecmw33_gen_hex8_count", "hecmw33_gen_hex8",
              "hecmw33_gen_tet10_count", "hecmw33_gen_tet10", "hecmw33_device_count",
              "hecmw33_last_history", "hecmw33_batch_load", "hecmw33_batch_solve", "hecmw33_batch_get_x",
              "hecmw33_batch_spmm", "hecmw33_batch_info", "hecmw33_solve_batch", "hecmw33_net_weight_count",
              "hecmw33_net_get_weights", "hecmw33_net_set_weights", "hecmw33_net_eval", "hecmw33_net_add_sample",
              "hecmw33_net_clear_samples", "hecmw33_net_train", "hecmw33_net_score", "hecmw33_ctx_set_net",
              "hecmw33_set_net", "hecmw33_set_coords", "hecmw33_ctx_set_coords", "hecmw33_ctx_coarse_get",
              "hecmw33_ctx_coarse_inverse_get", "hecmw33_set_mpc", "hecmw33_ctx_set_mpc", "hecmw33_matvec_mpc",
              "hecmw33_rigid_labels", "hecmw33_ctx_ssor_device"):
        getattr(L, f).restype = C.c_int32
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def _check(rc):
    if rc < 0:
        raise HecmwError("libhecmw33 error %d: %s" % (rc, lib().hecmw33_last_error().decode()))
    return rc


class HecMatrix:
    """The arrays of hecmwST_matrix for NDOF = 3, with hecmw_mat_init defaults.

    index arrays are 0-based cumulative (length NP+1), items 1-based, blocks
    row-major, B/X xyz interleaved -- exactly the reference layout."""

    # Iarray / Rarray slots, hecmw1/src/solver/matrix/hecmw_matrix_misc.F90:107-150
    I_ITER, I_METHOD, I_PRECOND, I_NSET, I_ITERPREMAX, I_NREST, I_SCALING, I_METHOD2 = 1, 2, 3, 4, 5, 6, 7, 8
    I_ITERLOG, I_TIMELOG, I_NCOLOR_IN, I_MAXRECYCLE = 21, 22, 34, 35
    I_FLAG_CONVERGED, I_FLAG_DIVERGED = 81, 82
    I_NRECYCLE, I_FLAG_NUMFACT, I_FLAG_SYMBFACT, I_SOLVER_TYPE = 96, 97, 98, 99
    R_RESID, R_SIGMA_DIAG, R_SIGMA, R_THRESH, R_FILTER = 1, 2, 3, 4, 5
    # SOLVER_OPT block (hecmw_matrix_misc.F90:138-150), unused by CG with PRECOND 1-3
    I_SOLVER_OPT, R_SOLVER_OPT = 41, 41

    def __init__(self, N, D, AL, AU, indexL, itemL, indexU, itemU, B=None, X=None, NP=None):
        self.N = int(N)
        self.NP = int(NP if NP is not None else N)
        self.D = np.ascontiguousarray(D, dtype=np.float64)
        self.AL = np.ascontiguousarray(AL, dtype=np.float64)
        self.AU = np.ascontiguousarray(AU, dtype=np.float64)
        self.indexL = np.ascontiguousarray(indexL, dtype=np.int32)
        self.indexU = np.ascontiguousarray(indexU, dtype=np.int32)
        self.itemL = np.ascontiguousarray(itemL, dtype=np.int32)
        self.itemU = np.ascontiguousarray(itemU, dtype=np.int32)
        self.NPL = int(self.indexL[self.NP])
        self.NPU = int(self.indexU[self.NP])
        self.B = np.zeros(3 * self.NP) if B is None else np.ascontiguousarray(B, dtype=np.float64)
        self.X = np.zeros(3 * self.NP) if X is None else np.ascontiguousarray(X, dtype=np.float64)
        self.Iarray = np.zeros(100, dtype=np.int32)
        self.Rarray = np.zeros(100, dtype=np.float64)
        self.mat_init()

    # hecmw_mat_init (hecmw_matrix_misc.F90:171-212)
    def mat_init(self):
        self.Iarray[:] = 0
        self.Rarray[:] = 0.0
        self.set_i(self.I_ITER, 100)
        self.set_i(self.I_METHOD, 1)
        self.set_i(self.I_PRECOND, 1)
        self.set_i(self.I_ITERPREMAX, 1)
        self.set_i(self.I_NREST, 10)
        self.set_i(self.I_NCOLOR_IN, 10)
        self.set_i(self.I_MAXRECYCLE, 3)
        self.set_i(24, -1)
        self.Rarray[self.R_RESID - 1] = 1.0e-8
        self.Rarray[self.R_SIGMA_DIAG - 1] = 1.0
        self.Rarray[self.R_THRESH - 1] = 0.10
        self.Rarray[self.R_FILTER - 1] = 0.10
        self.Rarray[11 - 1] = 1.0e4
        self.set_i(self.I_NRECYCLE, 0)
        self.set_i(self.I_FLAG_NUMFACT, 1)
        self.set_i(self.I_FLAG_SYMBFACT, 1)
        self.set_i(self.I_SOLVER_TYPE, 1)

    def set_i(self, k, v):
        self.Iarray[k - 1] = v

    def get_i(self, k):
        return int(self.Iarray[k - 1])

    # named setters mirroring hecmw_mat_set_*
    def set_iter(self, v): self.set_i(self.I_ITER, v)
    def set_method(self, v): self.set_i(self.I_METHOD, v)
    def set_method2(self, v): self.set_i(self.I_METHOD2, v)
    def set_precond(self, v): self.set_i(self.I_PRECOND, v)
    def set_iterpremax(self, v): self.set_i(self.I_ITERPREMAX, v)
    def set_ncolor_in(self, v): self.set_i(self.I_NCOLOR_IN, v)
    def set_resid(self, v): self.Rarray[self.R_RESID - 1] = v

    def set_mixed(self, on=True, delta=None, spmv=0, ssor_bf16=True):
        """Flexible CG with FP32 operators and FP64 residual replacement
        (Iarray 41 = 1, METHOD 1 only); delta: reliable-update factor on |r|
        (Rarray 41, default 1e-2); spmv (Iarray 42): 1 FP32 values in q = A p,
        2 FP64 values with FP32 vectors, 0 chosen from the stiffness contrast;
        ssor_bf16 (Iarray 47 = 0, else 1): the FP32 SSOR sweeps read BF16
        copies of the L'/U' values."""
        self.set_i(self.I_SOLVER_OPT, 1 if on else 0)
        self.set_i(self.I_SOLVER_OPT + 1, spmv)
        self.set_i(self.I_SOLVER_OPT + 6, 0 if ssor_bf16 else 1)
        self.Rarray[self.R_SOLVER_OPT - 1] = 0.0 if delta is None else float(delta)

    def set_coarse(self, n=1, additive=False, deflation=False):
        """Two-level correction of the flexible CG (Iarray 43): 0 off, 1 about
        1000 aggregates, larger values that many (needs node coordinates:
        set_coords / DeviceContext.set_coords).  Iarray 44: the symmetric
        coarse-SSOR-coarse step (default), with additive=True M^-1 r + Q r,
        with deflation=True y = M^-1 r, z = y + Q (r - A y)."""
        self.set_i(43, int(n))
        self.set_i(44, 1 if additive else (2 if deflation else 0))

    def set_sigma_diag(self, v):
        # hecmw_mat_set_sigma_diag clamps to [1, 2], negative = automatic (-1)
        self.Rarray[self.R_SIGMA_DIAG - 1] = -1.0 if v < 0 else min(2.0, max(1.0, v))

    def _sys(self):
        s = _Sys(self.N, self.NP, self.NPL, self.NPU, _ptr(self.D), _ptr(self.AL), _ptr(self.AU),
                 _ptr(self.indexL), _ptr(self.indexU), _ptr(self.itemL), _ptr(self.itemU))
        return s

    def copy(self):
        m = HecMatrix(self.N, self.D.copy(), self.AL.copy(), self.AU.copy(), self.indexL.copy(), self.itemL.copy(),
                      self.indexU.copy(), self.itemU.copy(), self.B.copy(), self.X.copy(), NP=self.NP)
        m.Iarray[:] = self.Iarray
        m.Rarray[:] = self.Rarray
        return m


@dataclass
class SolveResult:
    iter: int
    error: int
    method: int
    ncolor: int
    resid: float
    resid_true: float
    Tset: float
    Tsol: float
    time_Ax: float
    time_precond: float
    t_upload: float

    @classmethod
    def _from(cls, o):
        return cls(o.iter, o.error, o.method, o.ncolor, o.resid, o.resid_true, o.Tset, o.Tsol, o.time_Ax,
                   o.time_precond, o.t_upload)


def hecmw_solve_iterative(mat):
    """hecmw_solve_iterative for NDOF=3 on the GPU; mat.X is updated and
    mat.Iarray/Rarray receive the flags the reference writes."""
    out = _Out()
    s = mat._sys()
    rc = lib().hecmw33_solve(C.byref(s), _ptr(mat.B), _ptr(mat.X), _ptr(mat.Iarray), _ptr(mat.Rarray), C.byref(out))
    _check(rc)
    if rc == 2001:
        out.error = 2001
    return SolveResult._from(out)


def hecmw_solve_krylov(mat):
    """One hecmw_solve_CG / hecmw_solve_BiCGSTAB call (METHOD from Iarray)."""
    out = _Out()
    s = mat._sys()
    rc = lib().hecmw33_solve_krylov(C.byref(s), _ptr(mat.B), _ptr(mat.X), _ptr(mat.Iarray), _ptr(mat.Rarray),
                                    C.byref(out))
    _check(rc)
    return SolveResult._from(out)


def last_history(cap=100000):
    """ITERLOG residuals of the last hecmw_solve_iterative / hecmw_solve_krylov call."""
    h = np.zeros(cap)
    n = _check(lib().hecmw33_last_history(_ptr(h), cap))
    return h[:n]


def stream_read(buf_ptr, n, part_ptr, reps=10):
    """HBM read bandwidth (GB/s) of a plain streaming kernel over n doubles of
    device memory at buf_ptr; part_ptr: device scratch of 65536 doubles."""
    g = C.c_double(0.0)
    _check(lib().hecmw33_stream_read(buf_ptr, n, part_ptr, reps, C.byref(g)))
    return g.value


def hecmw_matvec(mat, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(3 * mat.N)
    s = mat._sys()
    _check(lib().hecmw33_matvec(C.byref(s), _ptr(x), _ptr(y)))
    return y


def _mpc_arrays(mpc):
    idx, item, dof, val = mpc
    return (np.ascontiguousarray(idx, np.int32), np.ascontiguousarray(item, np.int32),
            np.ascontiguousarray(dof, np.int32), np.ascontiguousarray(val, np.float64))


_mpc_keep = None


def set_mpc(N, mpc):
    """hecMESH%mpc for hecmw_solve_iterative / hecmw_solve_krylov's MPC-CG
    (T^T A T products when MPC_METHOD, Iarray 13, is 2): mpc = (index[0:n+1],
    item, dof, val) with 1-based items and dofs, the first term of each
    constraint its slave; None clears."""
    global _mpc_keep
    if mpc is None:
        _mpc_keep = None
        _check(lib().hecmw33_set_mpc(N, 0, None, None, None, None))
        return
    a = _mpc_arrays(mpc)
    _check(lib().hecmw33_set_mpc(N, len(a[0]) - 1, *[_ptr(x) for x in a]))
    _mpc_keep = a


def hecmw_matvec_mpc(mat, x):
    """y = T^T A T x (hecmw_TtmatTvec_33) with the constraints of set_mpc."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(3 * mat.N)
    s = mat._sys()
    _check(lib().hecmw33_matvec_mpc(C.byref(s), _ptr(x), _ptr(y)))
    return y


def hecmw_precond_apply(mat, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros(3 * mat.N)
    s = mat._sys()
    _check(lib().hecmw33_precond_apply(C.byref(s), _ptr(mat.Iarray), _ptr(mat.Rarray), _ptr(r), _ptr(z)))
    return z


def hecmw_rel_resid_L2(mat):
    s = mat._sys()
    v = C.c_double(0.0)
    _check(lib().hecmw33_rel_resid_L2(C.byref(s), _ptr(mat.B), _ptr(mat.X), C.byref(v)))
    return v.value


def hecmw_ordering(mat, ncolor_in=10):
    """(perm_rcm, perm, iperm, colorindex) of the SSOR setup, 1-based ids."""
    N = mat.N
    pr = np.zeros(N, np.int32)
    p = np.zeros(N, np.int32)
    ip = np.zeros(N, np.int32)
    ci = np.zeros(N + 1, np.int32)
    nc = C.c_int32(0)
    _check(lib().hecmw33_ordering(N, _ptr(mat.indexL), _ptr(mat.itemL), _ptr(mat.indexU), _ptr(mat.itemU),
                                  ncolor_in, _ptr(pr), _ptr(p), _ptr(ip), C.byref(nc), _ptr(ci)))
    return pr, p, ip, ci[: nc.value + 1].copy()


def gen_hex8(n, kind="uniform", E=210000.0, nu=0.3, seed=0, emin=1e3, emax=1e6, numin=0.2, numax=0.4,
             contrast=1e4, rmin=2, rmax=6, vfrac=0.15, return_E=False):
    """Structured n^3 HEX8 unit cube assembled like fistr1 (see gen_hex8.cpp).

    kind: "uniform" (E, nu), "lognormal" per-element E log-uniform in
    [emin, emax] with nu uniform in [numin, numax], or "inclusions"
    (two-phase spheres of E*contrast, vfrac volume fraction)."""
    L = lib()
    N = C.c_int32()
    NPL = C.c_int32()
    NPU = C.c_int32()
    _check(L.hecmw33_gen_hex8_count(n, C.byref(N), C.byref(NPL), C.byref(NPU)))
    N, NPL, NPU = N.value, NPL.value, NPU.value
    D = np.empty(9 * N)
    AL = np.empty(9 * NPL)
    AU = np.empty(9 * NPU)
    iL = np.empty(N + 1, np.int32)
    iU = np.empty(N + 1, np.int32)
    tL = np.empty(NPL, np.int32)
    tU = np.empty(NPU, np.int32)
    B = np.empty(3 * N)
    Ee = np.empty(n ** 3, np.float32) if return_E else None
    if kind == "uniform":
        k, p = 0, [E, nu]
    elif kind == "lognormal":
        k, p = 1, [emin, emax, numin, numax]
    elif kind == "inclusions":
        k, p = 2, [E, nu, contrast, rmin, rmax, vfrac]
    else:
        raise ValueError(kind)
    pa = np.array(p + [0.0] * (8 - len(p)), dtype=np.float64)
    rc = L.hecmw33_gen_hex8(n, k, _ptr(pa), seed, _ptr(D), _ptr(AL), _ptr(AU), _ptr(iL), _ptr(tL), _ptr(iU),
                            _ptr(tU), _ptr(B), _ptr(Ee) if Ee is not None else None)
    if rc != 0:
        raise HecmwError("hecmw33_gen_hex8 failed (%d)" % rc)
    m = HecMatrix(N, D, AL, AU, iL, tL, iU, tU, B=B)
    if return_E:
        return m, Ee
    return m


def gen_hex8_structure(n):
    """The profile alone of gen_hex8(n) (no values, loads or boundary
    conditions): an object with N, NP, NPL, NPU, indexL, itemL, indexU, itemU,
    for the orderings of large cubes."""
    from types import SimpleNamespace
    L = lib()
    N, NPL, NPU = C.c_int32(), C.c_int32(), C.c_int32()
    _check(L.hecmw33_gen_hex8_count(n, C.byref(N), C.byref(NPL), C.byref(NPU)))
    N, NPL, NPU = N.value, NPL.value, NPU.value
    iL = np.empty(N + 1, np.int32)
    iU = np.empty(N + 1, np.int32)
    tL = np.empty(NPL, np.int32)
    tU = np.empty(NPU, np.int32)
    pa = np.array([210000.0, 0.3] + [0.0] * 6)
    rc = L.hecmw33_gen_hex8(n, 0, _ptr(pa), 0, None, None, None, _ptr(iL), _ptr(tL), _ptr(iU), _ptr(tU), None, None)
    if rc != 0:
        raise HecmwError("hecmw33_gen_hex8 (structure) failed (%d)" % rc)
    return SimpleNamespace(N=N, NP=N, NPL=NPL, NPU=NPU, indexL=iL, itemL=tL, indexU=iU, itemU=tU)


def hex8_coords(n):
    """Node coordinates of gen_hex8(n) (node i + (n+1)(j + (n+1)k) at (i, j, k)/n), 3*(n+1)^3."""
    nn = n + 1
    idx = np.arange(nn ** 3)
    return (np.stack([idx % nn, (idx // nn) % nn, idx // (nn * nn)], 1) / float(n)).ravel()


def tet10_coords(nx, ny, nz, box=(1.0, 1.0, 1.0)):
    """Node coordinates of gen_tet10(nx, ny, nz, box): the half-cell lattice,
    node x + X(y + Y z) at (x/(2nx) Lx, y/(2ny) Ly, z/(2nz) Lz)."""
    X, Y, Z = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    idx = np.arange(X * Y * Z)
    return np.stack([(idx % X) * (box[0] / (2 * nx)), ((idx // X) % Y) * (box[1] / (2 * ny)),
                     (idx // (X * Y)) * (box[2] / (2 * nz))], 1).ravel()


def set_coords(xyz):
    """Node coordinates for hecmw_solve_iterative / hecmw_solve_krylov's two-level correction."""
    global _coords
    _coords = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
    _check(lib().hecmw33_set_coords(_coords.size // 3, _ptr(_coords)))


_coords = None


def gen_tet10(nx, ny, nz, box=(1.0, 1.0, 1.0), E=210000.0, nu=0.3, nrhs=1, seed=0):
    """TET10 (fistr1 type 342) box on an nx x ny x nz cell grid, 6 tetrahedra per
    cell, bottom face fixed, nrhs columns of N(0,1) nodal loads on the top face
    (see gen_tet10.cpp).  Returns (HecMatrix with B = column 0, B_all of shape
    (nrhs, 3N))."""
    L = lib()
    N = C.c_int32()
    NPL = C.c_int32()
    NPU = C.c_int32()
    _check(L.hecmw33_gen_tet10_count(nx, ny, nz, C.byref(N), C.byref(NPL), C.byref(NPU)))
    N, NPL, NPU = N.value, NPL.value, NPU.value
    D = np.zeros(9 * N)
    AL = np.zeros(9 * max(NPL, 1))
    AU = np.zeros(9 * max(NPU, 1))
    iL = np.zeros(N + 1, np.int32)
    iU = np.zeros(N + 1, np.int32)
    tL = np.zeros(max(NPL, 1), np.int32)
    tU = np.zeros(max(NPU, 1), np.int32)
    Ball = np.zeros((max(nrhs, 1), 3 * N))
    bx = np.array(box, dtype=np.float64)
    rc = L.hecmw33_gen_tet10(nx, ny, nz, _ptr(bx), E, nu, nrhs, seed, _ptr(D), _ptr(AL), _ptr(AU), _ptr(iL),
                             _ptr(tL), _ptr(iU), _ptr(tU), _ptr(Ball))
    if rc != 0:
        raise HecmwError("hecmw33_gen_tet10 failed (%d)" % rc)
    m = HecMatrix(N, D, AL[:9 * NPL], AU[:9 * NPU], iL, tL[:NPL], iU, tU[:NPU], B=Ball[0].copy())
    return m, Ball[:nrhs]


class DeviceContext:
    """A context that keeps the system resident in HBM (hecmw33_ctx_*)."""

    def __init__(self):
        self._h = lib().hecmw33_ctx_create()
        if not self._h:
            raise HecmwError("hecmw33_ctx_create failed: %s" % lib().hecmw33_last_error().decode())

    def close(self):
        if self._h:
            lib().hecmw33_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, mat):
        s = mat._sys()
        _check(lib().hecmw33_ctx_load(self._h, C.byref(s), _ptr(mat.Iarray), _ptr(mat.Rarray)))

    def solve_device(self, b_ptr, x_ptr, iarray, rarray):
        out = _Out()
        ia = np.ascontiguousarray(iarray, dtype=np.int32)
        ra = np.ascontiguousarray(rarray, dtype=np.float64)
        rc = lib().hecmw33_ctx_solve_device(self._h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), _ptr(ia), _ptr(ra),
                                            C.byref(out))
        _check(rc)
        return SolveResult._from(out)

    def matvec_device(self, x_ptr, y_ptr, count=1, fp32=False):
        ms = C.c_double(0.0)
        f = lib().hecmw33_ctx_matvec_device_f32 if fp32 else lib().hecmw33_ctx_matvec_device
        _check(f(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), count, C.byref(ms)))
        return ms.value

    def ssor_device(self, r_ptr, z_ptr, count=1, fp32=False, bf16=True):
        """count SSOR applications on device vectors; mean ms per application.
        fp32: float vectors and the FP32 image, with BF16 (bf16) or FP32 L'/U' values"""
        ms = C.c_double(0.0)
        mode = (1 if bf16 else 2) if fp32 else 0
        _check(lib().hecmw33_ctx_ssor_device(self._h, C.c_void_p(r_ptr), C.c_void_p(z_ptr), mode, count,
                                             C.byref(ms)))
        return ms.value

    def history(self, cap):
        h = np.zeros(cap)
        n = _check(lib().hecmw33_ctx_history(self._h, _ptr(h), cap))
        return h[:n]

    def info(self):
        v = np.zeros(21, np.int64)
        _check(lib().hecmw33_ctx_info(self._h, _ptr(v), 21))
        return {"nslots": int(v[0]), "N": int(v[1]), "NPL": int(v[2]), "NPU": int(v[3]), "ncolor": int(v[4]),
                "ssor_slots": int(v[5]), "nreliable": int(v[6]), "fallback_iter": int(v[7]), "learned_off": int(v[8]),
                "spmv64": int(v[9]), "learned_modes": int(v[10]), "coarse": int(v[11]), "coarse_pairs": int(v[12]),
                "az_entries": int(v[13]), "az_entries_symbolic": int(v[14]), "modes_per_aggregate": int(v[15]),
                "mode_samples": int(v[16]), "mode_sample_iters": int(v[17]), "mode_harvest_s": v[18] * 1e-6,
                "mode_fit_s": v[19] * 1e-6, "coarse_unknowns": int(v[20])}

    def coarse(self, with_E=True):
        """The last two-level set-up: dict(n_agg, n_pair, agg [N], dv [N,3], pair_a, pair_b [n_pair],
        Eblk [n_pair,6,6] of the rigid-body modes, or None without with_E)."""
        na, npair = C.c_int32(0), C.c_int32(0)
        _check(lib().hecmw33_ctx_coarse_get(self._h, C.byref(na), C.byref(npair), None, None, None, None, None))
        na, npair = na.value, npair.value
        if na == 0:
            return None
        N = self.info()["N"]
        agg = np.zeros(N, np.int32)
        dv = np.zeros(3 * N, np.float32)
        pa = np.zeros(npair, np.int32)
        pb = np.zeros(npair, np.int32)
        E = np.zeros(36 * npair) if with_E else None
        n2, p2 = C.c_int32(0), C.c_int32(0)
        _check(lib().hecmw33_ctx_coarse_get(self._h, C.byref(n2), C.byref(p2), _ptr(agg), _ptr(dv), _ptr(pa),
                                            _ptr(pb), _ptr(E) if with_E else None))
        return {"n_agg": na, "n_pair": npair, "agg": agg, "dv": dv.reshape(N, 3), "pair_a": pa, "pair_b": pb,
                "Eblk": E.reshape(npair, 6, 6) if with_E else None}

    def coarse_inverse(self):
        """The coarse operator the iterations read: (W [nc, nc], fp64).  fp64 False
        (six modes per aggregate): W = E^-1 in FP32; True (twelve): W holds L^-1 on
        and below the diagonal and L^-T above it, E = L L^T."""
        nc, ldw, f64 = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        _check(lib().hecmw33_ctx_coarse_inverse_get(self._h, C.byref(nc), C.byref(ldw), C.byref(f64), None))
        W = np.zeros(nc.value * ldw.value, np.float64 if f64.value else np.float32)
        _check(lib().hecmw33_ctx_coarse_inverse_get(self._h, C.byref(nc), C.byref(ldw), C.byref(f64), _ptr(W)))
        return W.reshape(nc.value, ldw.value)[:, :nc.value], bool(f64.value)

    def modes(self, with_E=True):
        """The learned coarse modes of the last fit: dict(xm [N,6,3], Eblk [n_pair,12,12] or None,
        R [N,3,16]: the error samples beyond the rigid-body modes that the fit used)."""
        N = self.info()["N"]
        xm = np.zeros(18 * N, np.float32)
        R = np.zeros(48 * N, np.float32)
        E = None
        if with_E:
            na, npair = C.c_int32(0), C.c_int32(0)
            _check(lib().hecmw33_ctx_coarse_get(self._h, C.byref(na), C.byref(npair), None, None, None, None, None))
            E = np.zeros(144 * npair.value)
        _check(lib().hecmw33_ctx_modes_get(self._h, _ptr(xm), _ptr(E) if E is not None else None, _ptr(R)))
        return {"xm": xm.reshape(N, 6, 3), "Eblk": E.reshape(-1, 12, 12) if E is not None else None,
                "R": R.reshape(N, 3, 16)}

    def set_mpc(self, N, mpc):
        """Multipoint constraints (as frontistr_amd.set_mpc) for this context's solves."""
        if mpc is None:
            _check(lib().hecmw33_ctx_set_mpc(self._h, N, 0, None, None, None, None))
            self._mpc = None
            return
        self._mpc = _mpc_arrays(mpc)
        _check(lib().hecmw33_ctx_set_mpc(self._h, N, len(self._mpc[0]) - 1, *[_ptr(x) for x in self._mpc]))

    def set_coords(self, xyz):
        """Node coordinates (3*NP, xyz interleaved) for the two-level correction."""
        self._coords = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
        _check(lib().hecmw33_ctx_set_coords(self._h, self._coords.size // 3, _ptr(self._coords)))

    def set_net(self, net):
        """Attach a LearnedNet for solves with Iarray(41) = 2 (None detaches)."""
        self._net = net
        _check(lib().hecmw33_ctx_set_net(self._h, net._h if net is not None else None))


def _bcheck(rc):
    if rc < 0:
        raise HecmwError("libhecmw33 batch error %d: %s" % (rc, lib().hecmw33_batch_last_error().decode()))
    return rc


class BatchSolver:
    """Many systems x up to 16 right-hand sides in one block-diagonal SpMM
    (hecmw33_batch_*): DIAG-preconditioned flexible CG per (system, rhs),
    FP32 operators, FP64 x / inner products / residual replacement."""

    def __init__(self, mats, Bs, iarray=None, rarray=None):
        self._mats = [m for m in mats]
        self.nrhs = int(Bs[0].shape[0])
        self._B = [np.ascontiguousarray(b, dtype=np.float64) for b in Bs]
        self._sys = (_Sys * len(mats))(*[m._sys() for m in mats])
        bp = (C.c_void_p * len(mats))(*[b.ctypes.data for b in self._B])
        ia = np.ascontiguousarray(mats[0].Iarray if iarray is None else iarray, dtype=np.int32)
        ra = np.ascontiguousarray(mats[0].Rarray if rarray is None else rarray, dtype=np.float64)
        self._h = lib().hecmw33_batch_create()
        if not self._h:
            raise HecmwError("hecmw33_batch_create failed")
        _bcheck(lib().hecmw33_batch_load(self._h, len(mats), C.cast(self._sys, C.c_void_p), self.nrhs,
                                         C.cast(bp, C.c_void_p), _ptr(ia), _ptr(ra)))

    def close(self):
        if getattr(self, "_h", None):
            lib().hecmw33_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, iarray, rarray):
        """Returns (iter, error, resid) arrays of shape (nsys, nrhs) and the
        times (setup s, solve s, tuned SpMM ms, iterations enqueued)."""
        ns = len(self._mats)
        it = np.zeros(ns * self.nrhs, np.int32)
        er = np.zeros(ns * self.nrhs, np.int32)
        rs = np.zeros(ns * self.nrhs)
        tm = np.zeros(4)
        ia = np.ascontiguousarray(iarray, dtype=np.int32)
        ra = np.ascontiguousarray(rarray, dtype=np.float64)
        _bcheck(lib().hecmw33_batch_solve(self._h, _ptr(ia), _ptr(ra), _ptr(it), _ptr(er), _ptr(rs), _ptr(tm)))
        sh = (ns, self.nrhs)
        return it.reshape(sh), er.reshape(sh), rs.reshape(sh), tm

    def x(self, s):
        X = np.zeros((self.nrhs, 3 * self._mats[s].NP))
        _bcheck(lib().hecmw33_batch_get_x(self._h, s, _ptr(X)))
        return X

    def spmm(self, Xs, fp64=True, shape64=None):
        """Y = A X per system; fp64 with shape64=None runs the tuned FP64 shape,
        shape64=v FP64 shape v (all sum in the reference's order)."""
        Xs = [np.ascontiguousarray(x, dtype=np.float64) for x in Xs]
        Ys = [np.zeros_like(x) for x in Xs]
        xp = (C.c_void_p * len(Xs))(*[x.ctypes.data for x in Xs])
        yp = (C.c_void_p * len(Ys))(*[y.ctypes.data for y in Ys])
        mode = (1 if shape64 is None else 2 + int(shape64)) if fp64 else 0
        _bcheck(lib().hecmw33_batch_spmm(self._h, mode, C.cast(xp, C.c_void_p), C.cast(yp, C.c_void_p)))
        return Ys

    def info(self):
        v = np.zeros(128, np.int64)
        _bcheck(lib().hecmw33_batch_info(self._h, _ptr(v), 128))
        n32 = int(v[4])
        names = [lib().hecmw33_batch_variant_name(i).decode() for i in range(n32)]
        o = 5 + n32
        n64 = int(v[o + 1])
        names64 = [lib().hecmw33_batch_variant64_name(i).decode() for i in range(n64)]
        e = o + 2 + n64
        return {"rows": int(v[0]), "blocks": int(v[1]), "tiles": int(v[2]), "variant": names[int(v[3])],
                "variant_us": dict(zip(names, [int(t) for t in v[5:5 + n32]])),
                "variant64": names64[int(v[o])],
                "variant64_us": dict(zip(names64, [int(t) for t in v[o + 2:o + 2 + n64]])),
                "tune_cached": bool(v[e]), "cus": int(v[e + 1]),
                "grids": dict(zip(("p", "q", "update", "res", "diag"), [int(t) for t in v[e + 2:e + 7]]))}


def solve_batch(mats, Bs, iarray, rarray):
    """hecmw33_solve_batch: load + solve + copy back in one call.
    Returns (X list of (nrhs, 3N) arrays, iter, error, resid of shape (nsys, nrhs))."""
    ns, nrhs = len(mats), int(Bs[0].shape[0])
    Bs = [np.ascontiguousarray(b, dtype=np.float64) for b in Bs]
    Xs = [np.zeros_like(b) for b in Bs]
    sysa = (_Sys * ns)(*[m._sys() for m in mats])
    bp = (C.c_void_p * ns)(*[b.ctypes.data for b in Bs])
    xp = (C.c_void_p * ns)(*[x.ctypes.data for x in Xs])
    it = np.zeros(ns * nrhs, np.int32)
    er = np.zeros(ns * nrhs, np.int32)
    rs = np.zeros(ns * nrhs)
    ia = np.ascontiguousarray(iarray, dtype=np.int32)
    ra = np.ascontiguousarray(rarray, dtype=np.float64)
    _bcheck(lib().hecmw33_solve_batch(ns, C.cast(sysa, C.c_void_p), nrhs, C.cast(bp, C.c_void_p),
                                      C.cast(xp, C.c_void_p), _ptr(ia), _ptr(ra), _ptr(it), _ptr(er), _ptr(rs)))
    sh = (ns, nrhs)
    return Xs, it.reshape(sh), er.reshape(sh), rs.reshape(sh)


def gen_tet10_batch(nsys=48, nrhs=16, dof_min=5000, dof_max=600000, seed=0, E=210000.0, nu=0.3):
    """The batched workload: nsys TET10 boxes, sizes log-uniform in
    [dof_min, dof_max] DOF, aspect ratios y/x and z/x uniform in [0.5, 2],
    nrhs N(0,1) top-face load columns each (seeded).  Returns (mats, Bs, specs)."""
    rng = np.random.default_rng(seed)
    mats, Bs, specs = [], [], []
    for s in range(nsys):
        dof = float(np.exp(rng.uniform(np.log(dof_min), np.log(dof_max))))
        ay, az = rng.uniform(0.5, 2.0, size=2)
        # nodes = (2nx+1)(2ny+1)(2nz+1) ~ 8 nx^3 ay az
        nx = max(1, int(round((dof / 3.0 / (8.0 * ay * az)) ** (1.0 / 3.0))))
        ny, nz = max(1, int(round(nx * ay))), max(1, int(round(nx * az)))
        m, B = gen_tet10(nx, ny, nz, box=(1.0, ny / nx, nz / nx), E=E, nu=nu, nrhs=nrhs, seed=seed * 1000 + s)
        mats.append(m)
        Bs.append(B)
        specs.append((nx, ny, nz, 3 * m.N))
    return mats, Bs, specs


class LearnedNet:
    """The learned aggregation (hecmw33_net_*): a per-node MLP 16 -> 64 -> 64 -> 3
    (ReLU; BF16 MFMA kernels) whose first output is the logit of "the slow
    error moves this node rigidly with a stiff piece"; the flexible CG's
    two-level step (Iarray 41 = 2) splits the marked nodes of every cell into
    their own aggregates.  Trained on the GPU on labels built from the
    solver's own errors (add_sample)."""

    def __init__(self, seed=0):
        self._h = lib().hecmw33_net_create(seed)
        if not self._h:
            raise HecmwError("hecmw33_net_create failed: %s" % lib().hecmw33_last_error().decode())
        self.nw = lib().hecmw33_net_weight_count()

    def close(self):
        if getattr(self, "_h", None):
            lib().hecmw33_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self):
        """(W1 [16,64], W2 [64,64], W3 [64,16]) FP32 master weights."""
        w = np.zeros(self.nw, np.float32)
        _check(lib().hecmw33_net_get_weights(self._h, _ptr(w), self.nw))
        return w[:1024].reshape(16, 64), w[1024:5120].reshape(64, 64), w[5120:].reshape(64, 16)

    def set_weights(self, W1, W2, W3):
        w = np.concatenate([np.asarray(W1, np.float32).ravel(), np.asarray(W2, np.float32).ravel(),
                            np.asarray(W3, np.float32).ravel()])
        _check(lib().hecmw33_net_set_weights(self._h, _ptr(w), self.nw))

    def eval(self, X, dY=None):
        """Y = N(X) (n x 3); with dY also the weight gradients of sum(dY * Y)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        Y = np.zeros((n, 3), np.float32)
        dW = np.zeros(self.nw, np.float32) if dY is not None else None
        dYc = np.ascontiguousarray(dY, dtype=np.float32) if dY is not None else None
        _check(lib().hecmw33_net_eval(self._h, n, _ptr(X), _ptr(dYc), _ptr(Y), _ptr(dW)))
        if dY is None:
            return Y
        return Y, (dW[:1024].reshape(16, 64), dW[1024:5120].reshape(64, 64), dW[5120:].reshape(64, 16))

    def add_sample(self, ctx, e_ptr, tau=0.1):
        """An error sample e = A^-1 r (device FP64, 3*NP) on a loaded context
        with node coordinates; returns that context's sample count."""
        return _check(lib().hecmw33_net_add_sample(self._h, ctx._h, C.c_void_p(e_ptr), tau))

    def clear_samples(self):
        _check(lib().hecmw33_net_clear_samples(self._h))

    def train(self, epochs, lr=1e-3, seed=0):
        loss = np.zeros(epochs + 1)
        _check(lib().hecmw33_net_train(self._h, epochs, lr, seed, _ptr(loss)))
        return loss

    def score(self):
        """(mean logistic loss, fraction of nodes classified as their label) over the samples."""
        m = C.c_double(0.0)
        a = C.c_double(0.0)
        _check(lib().hecmw33_net_score(self._h, C.byref(m), C.byref(a)))
        return m.value, a.value


def set_net(net):
    """Attach a LearnedNet to the default context (hecmw_solve_iterative / _krylov)."""
    _check(lib().hecmw33_set_net(net._h if net is not None else None))


def harvest_errors(net, ctx, mat, b_dev, iters=(10, 40), tau=0.1):
    """Training samples from the solver itself: the FP64 multicolour-SSOR CG
    solution x* of the loaded system, then the errors e_k = x* - x_k = A^-1 r_k
    of its iterates at the cumulative iteration counts `iters` (the solve
    continues from the previous checkpoint's x), added to the network's samples
    (labels: hecmw33_rigid_labels).  Iterates that have converged (|e| below
    1e-10 |x*|) give no sample.  The context needs the node coordinates.
    b_dev: a torch float64 device tensor of 3*NP entries.  Returns the
    context's sample count (0 if none was added)."""
    import torch
    ia = mat.Iarray.copy()
    ia[mat.I_FLAG_NUMFACT - 1] = 0
    ia[mat.I_FLAG_SYMBFACT - 1] = 0
    ia[mat.I_SOLVER_OPT - 1] = 0
    ia[mat.I_PRECOND - 1] = 1
    xs = torch.zeros_like(b_dev)
    res = ctx.solve_device(b_dev.data_ptr(), xs.data_ptr(), ia, mat.Rarray)
    if res.error != 0:
        raise HecmwError("harvest_errors: the reference solve failed (%d)" % res.error)
    nx = float(torch.linalg.vector_norm(xs))
    x = torch.zeros_like(b_dev)
    count, done = 0, 0
    for k in sorted(iters):
        if k <= done:
            continue
        ia[mat.I_ITER - 1] = k - done
        torch.cuda.synchronize()
        res = ctx.solve_device(b_dev.data_ptr(), x.data_ptr(), ia, mat.Rarray)
        done = k
        if res.error not in (0, 3001):
            break
        e = xs - x
        if float(torch.linalg.vector_norm(e)) <= 1e-10 * nx:
            break
        torch.cuda.synchronize()
        count = net.add_sample(ctx, e.data_ptr(), tau)
    return count


def rigid_labels(mat, xyz, errors, tau=0.1):
    """hecmw33_rigid_labels (host): labels (N) of the nodes of every edge whose
    axial strain stays below tau times the sample median in every error
    sample; errors: (nsamp, 3N)."""
    e = np.ascontiguousarray(errors, dtype=np.float64).reshape(-1, 3 * mat.N)
    X = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
    t = np.zeros(mat.N, np.float32)
    _check(lib().hecmw33_rigid_labels(mat.N, _ptr(mat.indexL), _ptr(mat.itemL), _ptr(mat.indexU), _ptr(mat.itemU),
                                      _ptr(X), e.shape[0], _ptr(e), float(tau), _ptr(t)))
    return t
