"""hecmw33_rel_resid_L2_device_noalloc: the relative residual of a system whose
arrays are in device memory, with no device allocation.

* against hecmw33_rel_resid_L2 (host arrays, the library's own device copy and
  partials) on the same system and x: the rows' residuals are the same
  statements in the same order, the sums run in another fixed order, so the two
  agree to 1e-12 relative;
* two calls in a row, the second on a smaller system (the scratch array keeps
  the first call's partials, which the second must not read), and the second
  leaves the device's free memory unchanged;
* b = 0 gives |r| (the reference takes |b|^2 = 0 as 1);
* a system above the scratch array's row limit is refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import frontistr_amd as fa
from frontistr_amd.hecmw import _Sys

pytestmark = [pytest.mark.gpu]


def _device_sys(m, torch):
    """device copies of the matrix arrays and a hecmw33_sys pointing at them"""
    keep = [torch.from_numpy(a).to("cuda:0") for a in (m.D, m.AL, m.AU, m.indexL, m.indexU, m.itemL, m.itemU)]
    ptr = [t.data_ptr() if t.numel() else None for t in keep]
    return _Sys(m.N, m.NP, m.NPL, m.NPU, *ptr), keep


def _noalloc(m, x, torch, b=None):
    s, keep = _device_sys(m, torch)
    bd = torch.from_numpy(m.B if b is None else b).to("cuda:0")
    xd = torch.from_numpy(x).to("cuda:0")
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = fa.lib().hecmw33_rel_resid_L2_device_noalloc(C.byref(s), C.c_void_p(bd.data_ptr()),
                                                      C.c_void_p(xd.data_ptr()), C.c_void_p(out.data_ptr()), None)
    assert rc == 0, fa.lib().hecmw33_last_error().decode()
    torch.cuda.synchronize()
    return float(out.item()), keep


def _reference_entry(m, x):
    mm = m.copy()
    mm.X[:] = x
    return fa.hecmw_rel_resid_L2(mm)


def test_matches_rel_resid_on_two_calls_the_second_smaller():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    big = fa.gen_hex8(40, kind="lognormal", seed=2)  # 68,921 rows: 270 partials
    x1 = rng.standard_normal(3 * big.NP)
    r1, keep1 = _noalloc(big, x1, torch)
    e1 = _reference_entry(big, x1)
    assert r1 == pytest.approx(e1, rel=1e-12, abs=0.0)
    small = fa.gen_hex8(7)  # 512 rows: 2 partials
    x2 = rng.standard_normal(3 * small.NP)
    s, keep2 = _device_sys(small, torch)
    bd = torch.from_numpy(small.B).to("cuda:0")
    xd = torch.from_numpy(x2).to("cuda:0")
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    rc = fa.lib().hecmw33_rel_resid_L2_device_noalloc(C.byref(s), C.c_void_p(bd.data_ptr()),
                                                      C.c_void_p(xd.data_ptr()), C.c_void_p(out.data_ptr()), None)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert rc == 0
    assert free1 == free0
    e2 = _reference_entry(small, x2)
    assert float(out.item()) == pytest.approx(e2, rel=1e-12, abs=0.0)
    # the solution of the system gives a small residual too
    res = fa.hecmw_solve_krylov(small)
    r3, _ = _noalloc(small, small.X.copy(), torch)
    assert r3 == pytest.approx(_reference_entry(small, small.X.copy()), rel=1e-9, abs=1e-15)
    assert r3 <= 1e-8 * 1.0001 and res.error == 0


def test_zero_rhs_and_external_nodes():
    torch = pytest.importorskip("torch")
    from tests.systems import random_system
    m = random_system(500, deg=9, seed=4, n_ext=40)  # NP > N: columns past the internal rows
    rng = np.random.default_rng(9)
    x = rng.standard_normal(3 * m.NP)
    r, _ = _noalloc(m, x, torch)
    assert r == pytest.approx(_reference_entry(m, x), rel=1e-12, abs=0.0)
    # b = 0: sqrt(|A x|^2 / 1)
    z = np.zeros(3 * m.NP)
    r0, _ = _noalloc(m, x, torch, b=z)
    y = fa.hecmw_matvec(m, x)[:3 * m.N]
    assert r0 == pytest.approx(np.linalg.norm(y), rel=1e-12)


def test_refuses_a_system_above_its_row_limit():
    pytest.importorskip("torch")
    limit = 256 * 65536
    s = _Sys(limit + 1, limit + 1, 0, 0, 1, None, None, 1, 1, None, None)
    rc = fa.lib().hecmw33_rel_resid_L2_device_noalloc(C.byref(s), C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), None)
    assert rc == -1  # HECMW33_ERROR_ARGS
    assert "exceeds" in fa.lib().hecmw33_last_error().decode()
