"""The measurement paths: TIMELOG's kernel times and the HBM read-ceiling probe.

* TIMELOG (Iarray 22) times the preconditioner and the SpMV of one iteration per
  enqueued chunk and scales to all iterations; the events must not change the
  solve (x, iterations and residual identical with TIMELOG on and off), and
  the times must be positive and within the solve time.
* hecmw33_stream_read reads a buffer larger than the 256 MB Infinity Cache;
  its figure must be physical: above a floor and below the published 8 TB/s.
"""
import numpy as np
import pytest

import frontistr_amd as fa

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,precond,opt", [(1, 1, 0), (2, 1, 0), (1, 3, 0), (1, 1, 1)],
                         ids=["cg_ssor", "bicgstab_ssor", "cg_diag", "fcg_ssor"])
def test_timelog_does_not_change_the_solve(method, precond, opt):
    base = fa.gen_hex8(12, kind="lognormal", seed=5)
    base.set_iter(3000)
    base.set_method(method)
    base.set_precond(precond)
    base.set_i(base.I_SOLVER_OPT, opt)
    out = {}
    for timelog in (0, 1):
        m = base.copy()
        m.set_i(m.I_TIMELOG, timelog)
        res = fa.hecmw_solve_iterative(m)
        out[timelog] = (res, m.X.copy())
    (r0, x0), (r1, x1) = out[0], out[1]
    assert r0.error == 0 and r1.error == 0
    assert r0.iter == r1.iter and r0.resid == r1.resid
    assert np.array_equal(x0, x1)
    assert r1.time_Ax > 0.0 and r1.time_precond > 0.0
    # scaled from one iteration in eight: within the solve's wall time, with room
    # for the sampling error of short solves
    assert r1.time_Ax + r1.time_precond <= 1.5 * r1.Tsol + 1e-3, (r1.time_Ax, r1.time_precond, r1.Tsol)


def test_stream_read_is_physical():
    torch = pytest.importorskip("torch")
    buf = torch.ones(1 << 26, dtype=torch.float64, device="cuda:0")  # 512 MiB
    part = torch.empty(65536, dtype=torch.float64, device="cuda:0")
    gbs = fa.stream_read(buf.data_ptr(), buf.numel(), part.data_ptr(), 3)
    print("stream read %.0f GB/s" % gbs)
    assert 500.0 < gbs < 8000.0
    # the last launch shape is 65536 workgroups: their partials of a read of
    # ones add up to the element count, so every element was read once
    torch.cuda.synchronize()
    assert torch.sum(part).item() == float(buf.numel())
