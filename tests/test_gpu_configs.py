"""Parity at the full sizes of the five benchmark configurations.

* C1, 20^3 HEX8 cube (27,783 DOF): SpMV, SSOR and DIAG applications bitwise
  equal to the reference's own routines; SSOR-CG and DIAG-CG within +-2
  iterations of the reference's, x within 1e-6.
* C2, 100^3 HEX8 cube (3,090,903 DOF): the reference's full SSOR-CG and
  DIAG-CG solves on the host cores (16 threads) against the GPU's FP64 solves
  of the same arrays: x within 1e-6, iterations within +-2, the same colours.
* C3, 160^3 cube with per-element log-uniform E (12,519,843 DOF) and C5, 120^3
  two-phase inclusion cube (5,314,683 DOF): every solver mode reaches an FP64
  true residual (hecmw33_rel_resid_L2) of 1e-8, and its x is within 1e-6 of
  the GPU's FP64 multicolour SSOR-CG solution, which C1/C2 pin to the reference.
* C4, 48 TET10 boxes x 16 right-hand sides: every instance's FP64 true
  residual (the FP64 SpMM, bitwise hecmw_matvec_33) is at most 1e-8, and the
  smallest systems' solutions are within 1e-6 of the reference's DIAG-CG.
"""
import os

import numpy as np
import pytest

import frontistr_amd as fa
from oracle import ref

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]

TOL = 1e-8 * (1.0 + 1e-6)  # the solver's own FP64 residual against hecmw_rel_resid_L2's (another sum order)


def _host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 2)
    return max(2, min(16, n))


def _relerr(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


# ---- C1 ----------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c1():
    m = fa.gen_hex8(20)
    m.set_iter(20000)
    return m


def test_c1_kernels_bitwise(c1):
    rng = np.random.default_rng(11)
    x = rng.standard_normal(3 * c1.NP)
    assert np.array_equal(fa.hecmw_matvec(c1, x), ref.matvec(c1, x))
    for pc in (1, 3):
        m = c1.copy()
        m.set_precond(pc)
        assert np.array_equal(fa.hecmw_precond_apply(m, x), ref.precond(m, x)), pc


@pytest.mark.parametrize("pc", [1, 3])
def test_c1_cg(c1, pc):
    m = c1.copy()
    m.set_precond(pc)
    x_ref, info = ref.solve_krylov(m)
    r = fa.hecmw_solve_krylov(m)
    assert r.error == 0 and info["error"] == 0
    assert abs(r.iter - info["iter"]) <= 2, (r.iter, info["iter"])
    assert r.resid <= 1e-8 and ref.rel_resid(c1, m.X) <= TOL
    assert _relerr(m.X, x_ref) <= 1e-6


# ---- C2 ----------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2():
    m = fa.gen_hex8(100)
    m.set_iter(20000)
    return m


def test_c2_spmv_bitwise(c2):
    x = np.random.default_rng(12).standard_normal(3 * c2.NP)
    assert np.array_equal(fa.hecmw_matvec(c2, x), ref.matvec(c2, x, nthreads=_host_threads()))


@pytest.mark.parametrize("pc", [1, 3])
def test_c2_cg_against_full_reference_solve(c2, pc):
    m = c2.copy()
    m.set_precond(pc)
    x_ref, info = ref.solve_krylov(m, nthreads=_host_threads())
    assert info["error"] == 0 and info["resid"] <= 1e-8
    r = fa.hecmw_solve_krylov(m)
    print("C2 precond %d: GPU %d iterations, reference %d (%d threads, %.1f s)"
          % (pc, r.iter, info["iter"], _host_threads(), info["Tset"] + info["Tsol"]))
    assert r.error == 0
    assert abs(r.iter - info["iter"]) <= 2, (r.iter, info["iter"])
    assert _relerr(m.X, x_ref) <= 1e-6
    if pc == 1:
        _, _, _, _, ci = ref.ordering(c2)
        assert r.ncolor == len(ci) - 1 == 17


# ---- C3 and C5: every mode against the GPU's FP64 SSOR-CG ----------------------------------
# (name, Iarray 41, Iarray 43, Iarray 44): FP64 CG; flexible CG with FP32 operators; with the
# symmetric two-level step; with the additive one; with the deflation one; with the learned
# coarse modes (symmetric and deflation steps)
MODES = (("fp64_ssor_cg", 0, 0, 0), ("mixed_fcg", 1, 0, 0), ("mixed_coarse_fcg", 1, 1, 0),
         ("mixed_coarse_additive_fcg", 1, 1, 1), ("mixed_coarse_deflation_fcg", 1, 1, 2),
         ("learned_coarse_fcg", 2, 1, 0), ("learned_coarse_deflation_fcg", 2, 1, 2))


class _Big:
    def __init__(self, m, coords, iters):
        import torch
        self.m = m
        m.set_iter(iters)
        m.set_precond(1)
        m.set_coarse(1)
        self.ctx = fa.DeviceContext()
        self.ctx.set_coords(coords)
        self.ctx.load(m)
        self.b = torch.from_numpy(m.B).to("cuda:0")
        self.x = torch.zeros(3 * m.NP, dtype=torch.float64, device="cuda:0")
        self.x64 = None

    def solve(self, opt, coarse, additive):
        import torch
        ia = self.m.Iarray.copy()
        ia[96], ia[97] = 1, 0
        ia[40], ia[42], ia[43] = opt, coarse, additive
        self.x.zero_()
        torch.cuda.synchronize()
        r = self.ctx.solve_device(self.b.data_ptr(), self.x.data_ptr(), ia, self.m.Rarray)
        return r, self.x.cpu().numpy()

    def close(self):
        self.ctx.close()


def _check_mode(big, name, opt, coarse, additive):
    if big.x64 is None:
        r0, big.x64 = big.solve(0, 0, 0)
        assert r0.error == 0
    r, x = big.solve(opt, coarse, additive)
    big.m.X[:] = x
    resid = fa.hecmw_rel_resid_L2(big.m)
    err = _relerr(x, big.x64)
    print("%s: %d iterations, FP64 true residual %.3e, |x - x_fp64|/|x_fp64| = %.2e" % (name, r.iter, resid, err))
    assert r.error == 0
    assert resid <= TOL
    assert err <= 1e-6


@pytest.fixture(scope="module")
def c3():
    big = _Big(fa.gen_hex8(160, kind="lognormal", seed=0), fa.hex8_coords(160), 20000)
    yield big
    big.close()


@pytest.mark.parametrize("name,opt,coarse,additive", MODES, ids=[m[0] for m in MODES])
def test_c3_modes(c3, name, opt, coarse, additive):
    _check_mode(c3, name, opt, coarse, additive)


@pytest.fixture(scope="module")
def c5():
    m = fa.gen_hex8(120, kind="inclusions", seed=7, E=1000.0, nu=0.3, contrast=1e4, rmin=2, rmax=6, vfrac=0.15)
    big = _Big(m, fa.hex8_coords(120), 40000)
    yield big
    big.close()


@pytest.mark.parametrize("name,opt,coarse,additive", MODES, ids=[m[0] for m in MODES])
def test_c5_modes(c5, name, opt, coarse, additive):
    _check_mode(c5, name, opt, coarse, additive)


# ---- C4 ----------------------------------------------------------------------------------
def test_c4_full_batch():
    mats, Bs, specs = fa.gen_tet10_batch(nsys=48, nrhs=16, seed=0)
    m0 = fa.gen_hex8(1)
    m0.set_iter(20000)
    m0.set_precond(3)
    bs = fa.BatchSolver(mats, Bs)
    it, er, rs, _ = bs.solve(m0.Iarray, m0.Rarray)
    assert np.all(er == 0), np.argwhere(er != 0)
    Xs = [bs.x(s) for s in range(len(mats))]
    # FP64 true residual of every instance, from the FP64 SpMM (hecmw_matvec_33's sums)
    AX = bs.spmm(Xs, fp64=True)
    worst = 0.0
    for s, m in enumerate(mats):
        for k in range(Bs[s].shape[0]):
            b = Bs[s][k][:3 * m.N]
            r = np.linalg.norm(b - AX[s][k][:3 * m.N]) / np.linalg.norm(b)
            worst = max(worst, r)
            assert r <= TOL, (s, k, r)
    print("C4: %d instances, iterations %d..%d, worst FP64 true residual %.3e" % (it.size, it.min(), it.max(), worst))
    # the smallest systems against the reference's DIAG-CG, first and last right-hand side
    order = np.argsort([sp[3] for sp in specs])[:3]
    for s in order:
        for k in (0, Bs[s].shape[0] - 1):
            mm = mats[s].copy()
            mm.B[:] = Bs[s][k]
            mm.set_iter(20000)
            mm.set_precond(3)
            x_ref, info = ref.solve_krylov(mm)
            assert info["error"] == 0
            assert _relerr(Xs[s][k], x_ref) <= 1e-6, (s, k)
    bs.close()
