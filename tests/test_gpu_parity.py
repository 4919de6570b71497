"""GPU parity against the reference's own solver (oracle/_ref), through the C ABI.

Bars: SpMV, DIAG and SSOR applications bitwise equal to the reference CPU
routines (same operation order, no fused multiply-add); Krylov solves to
1e-8 relative residual within 1e-6 relative error of the reference's solution,
with the same breakdown codes and an iteration count within +-2 of the
reference's (the only difference is the summation order of the dot products).
"""
import numpy as np
import pytest

import frontistr_amd as fa
from oracle import ref
from tests.systems import random_system, shuffle_items, dense

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]


def _systems():
    yield "hex8_10", fa.gen_hex8(10)
    yield "hex8_7_shuffled", shuffle_items(fa.gen_hex8(7), seed=4)
    yield "random_700", random_system(700, deg=12, seed=11)
    yield "random_ext", random_system(300, deg=8, seed=12, n_ext=25)
    yield "hex8_1", fa.gen_hex8(1)


SYSTEMS = list(_systems())
IDS = [s[0] for s in SYSTEMS]


@pytest.mark.parametrize("name,m", SYSTEMS, ids=IDS)
def test_matvec_bitwise(name, m):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(3 * m.NP)
    y = fa.hecmw_matvec(m, x)
    yr = ref.matvec(m, x)
    assert np.array_equal(y, yr), np.max(np.abs(y - yr))
    # and the closed form
    yd = dense(m) @ x
    assert np.allclose(y, yd[: 3 * m.N], rtol=1e-12, atol=1e-12 * np.abs(yd).max())


@pytest.mark.parametrize("precond", [3, 1])
@pytest.mark.parametrize("sigma", [1.0, 1.35])
@pytest.mark.parametrize("name,m", SYSTEMS, ids=IDS)
def test_precond_bitwise(name, m, precond, sigma):
    m = m.copy()
    m.set_precond(precond)
    m.set_sigma_diag(sigma)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(3 * m.NP)
    z = fa.hecmw_precond_apply(m, r)
    zr = ref.precond(m, r, nthreads=2)
    assert np.array_equal(z, zr), np.max(np.abs(z - zr))


@pytest.mark.parametrize("precond", [1, 3, 0])
@pytest.mark.parametrize("name", ["hex8_10", "random_700", "random_ext"])
def test_cg_matches_reference(name, precond):
    m = dict(SYSTEMS)[name].copy()
    m.set_iter(5000)
    if precond == 0:
        m.set_iterpremax(0)
    else:
        m.set_precond(precond)
    x0 = np.zeros(3 * m.NP)
    if name == "random_ext":
        x0[3 * m.N:] = np.random.default_rng(5).standard_normal(3 * (m.NP - m.N))
    m.X[:] = x0
    res = fa.hecmw_solve_krylov(m)
    xr, info = ref.solve_krylov(m, x0=x0, nthreads=2)
    assert res.error == 0 and info["error"] == 0
    assert abs(res.iter - info["iter"]) <= 2, (res.iter, info["iter"])
    assert res.resid <= 1e-8
    n3 = 3 * m.N
    err = np.linalg.norm(m.X[:n3] - xr[:n3]) / np.linalg.norm(xr[:n3])
    assert err <= 1e-6, err
    # the solution also solves the system (independent of the oracle)
    assert ref.rel_resid(m, m.X) <= 1e-8 * 1.0001


def test_iteration_history_tracks_reference_closely():
    """Same ordering, same sweeps: the residual history agrees to rounding."""
    m = fa.gen_hex8(12)
    m.set_iter(2000)
    res = fa.hecmw_solve_krylov(m)
    _, info = ref.solve_krylov(m, nthreads=2)
    assert res.iter == info["iter"]
    assert res.resid == pytest.approx(info["resid"], rel=1e-6)


def test_zero_rhs():
    m = fa.gen_hex8(3)
    m.B[:] = 0.0
    m.X[:] = 1.0
    res = fa.hecmw_solve_krylov(m)
    _, info = ref.solve_krylov(m, x0=np.ones(3 * m.NP))
    assert res.iter == info["iter"] == 1
    assert res.resid == 0.0 and res.error == 0
    assert np.all(m.X == 0.0)


def test_noconv_maxit():
    m = fa.gen_hex8(8)
    m.set_iter(7)
    res = fa.hecmw_solve_krylov(m)
    _, info = ref.solve_krylov(m)
    assert res.error == info["error"] == 3001
    assert res.iter == info["iter"] == 8  # Fortran DO variable after the last trip
    assert res.resid == pytest.approx(info["resid"], rel=1e-9)


def test_diverge_indefinite():
    m = fa.gen_hex8(4)
    m.set_precond(3)
    m.D[:] = -m.D  # negative definite diagonal
    m.AL[:] = -m.AL
    m.AU[:] = -m.AU
    res = fa.hecmw_solve_krylov(m)
    _, info = ref.solve_krylov(m)
    assert res.error == info["error"]
    assert res.error in (3002, 3003)
    assert res.iter == info["iter"]


def test_diverge_nan():
    m = fa.gen_hex8(4)
    m.set_precond(3)
    m.B[5] = np.nan
    res = fa.hecmw_solve_krylov(m)
    _, info = ref.solve_krylov(m)
    assert res.error == info["error"] == 3004
    assert res.iter == info["iter"]


def test_solve_iterative_flags_and_recycle():
    m = fa.gen_hex8(6)
    m.set_iter(1000)
    mr = m.copy()
    res = fa.hecmw_solve_iterative(m)
    xr, iout, rout = ref.solve_iterative(mr)
    assert res.error == 0
    for k in (81, 82, 96, 97, 98):
        assert m.get_i(k) == iout[k - 1], k
    assert np.linalg.norm(m.X - xr) / np.linalg.norm(xr) < 1e-6
    assert res.resid_true < 1e-8


def test_zero_diag_reported():
    m = fa.gen_hex8(3)
    m.D[9 * 5 + 4] = 0.0
    res = fa.hecmw_solve_iterative(m)
    assert res.error == 2001


@pytest.mark.parametrize("precond", [1, 3, 0])
@pytest.mark.parametrize("name", ["hex8_10", "random_700", "random_ext"])
def test_bicgstab_matches_reference(name, precond):
    m = dict(SYSTEMS)[name].copy()
    m.set_iter(5000)
    m.set_method(2)
    if precond == 0:
        m.set_iterpremax(0)
    else:
        m.set_precond(precond)
    res = fa.hecmw_solve_krylov(m)
    xr, info = ref.solve_krylov(m, nthreads=2)
    assert res.error == info["error"] == 0
    # BiCGSTAB's convergence is irregular, so the summation order of the dot
    # products (the only difference) moves its iteration count by more than
    # CG's; the residual history test below pins the early iterations.
    assert abs(res.iter - info["iter"]) <= max(2, 0.1 * info["iter"]), (res.iter, info["iter"])
    n3 = 3 * m.N
    err = np.linalg.norm(m.X[:n3] - xr[:n3]) / np.linalg.norm(xr[:n3])
    assert err <= 1e-6, err
    assert ref.rel_resid(m, m.X) <= 1e-8 * 1.0001


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("precond", [1, 3])
def test_iterpremax2_matches_reference(method, precond):
    m = fa.gen_hex8(8)
    m.set_iter(3000)
    m.set_method(method)
    m.set_precond(precond)
    m.set_iterpremax(2)
    res = fa.hecmw_solve_krylov(m)
    xr, info = ref.solve_krylov(m, nthreads=2)
    # two DIAG sweeps make 2D^-1 - D^-1 A D^-1, which is slightly indefinite on
    # this mesh: both sides must stop with DIVERGE_PC at the same iteration
    assert res.error == info["error"]
    tol_it = 2 if method == 1 else max(2, 0.1 * info["iter"])
    assert abs(res.iter - info["iter"]) <= tol_it, (res.iter, info["iter"])
    if info["error"] == 0:
        assert np.linalg.norm(m.X - xr) / np.linalg.norm(xr) <= 1e-6


def test_bicgstab_noconv_iter_count():
    m = fa.gen_hex8(8)
    m.set_iter(5)
    m.set_method(2)
    res = fa.hecmw_solve_krylov(m)
    _, info = ref.solve_krylov(m)
    assert res.error == info["error"] == 3001
    assert res.iter == info["iter"] == 6


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("precond", [1, 3, 0])
def test_residual_history_matches_reference(method, precond):
    """The per-iteration relative residuals (ITERLOG) agree with the
    reference's printed ones (7 significant digits) over the first iterations."""
    m = fa.gen_hex8(9)
    m.set_iter(3000)
    m.set_method(method)
    if precond == 0:
        m.set_iterpremax(0)
    else:
        m.set_precond(precond)
    res = fa.hecmw_solve_krylov(m)
    h = fa.last_history()
    _, info = ref.solve_krylov(m, nthreads=2, history=True)
    hr = info["history"]
    k = min(15, len(h), len(hr))
    assert k >= 10
    np.testing.assert_allclose(h[:k], hr[:k], rtol=2e-6)
    if method == 1:
        # CG keeps tracking the reference over the first half of the solve
        # (late iterates drift by rounding: the dot products sum in another order)
        n = min(len(h), len(hr)) // 2
        np.testing.assert_allclose(h[:n], hr[:n], rtol=1e-3)


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("precond", [1, 3])
def test_scaling_matches_reference(method, precond):
    """SCALING=1 (Iarray 7): the reference scales A and b by 1/sqrt(|D_ii|) in
    place inside hecmw_solve_CG/BiCGSTAB and restores them afterwards.  The
    residual history, iteration count and solution follow the reference's;
    x0 enters the scaled system unscaled, as in the reference; the caller's
    arrays come back bitwise unchanged."""
    m = fa.gen_hex8(8, kind="lognormal", seed=5)
    m.set_iter(5000)
    m.set_method(method)
    m.set_precond(precond)
    m.set_i(m.I_SCALING, 1)
    x0 = np.random.default_rng(3).standard_normal(3 * m.NP) * 1e-3
    m.X[:] = x0
    keep = [a.copy() for a in (m.D, m.AL, m.AU, m.B)]
    res = fa.hecmw_solve_krylov(m)
    h = fa.last_history()
    for a, k in zip((m.D, m.AL, m.AU, m.B), keep):
        assert np.array_equal(a, k)
    xr, info = ref.solve_krylov(m.copy(), x0=x0, nthreads=2, history=True)
    assert res.error == info["error"] == 0
    hr = info["history"]
    k = min(15, len(h), len(hr))
    np.testing.assert_allclose(h[:k], hr[:k], rtol=2e-6)
    tol_it = 2 if method == 1 else max(2, 0.1 * info["iter"])
    assert abs(res.iter - info["iter"]) <= tol_it, (res.iter, info["iter"])
    err = np.linalg.norm(m.X - xr) / np.linalg.norm(xr)
    assert err <= 1e-6, err
    # the scaled system's 1e-8 is not the unscaled one's: both answers sit
    # within the same order of the unscaled residual
    assert ref.rel_resid(m, m.X) <= 10 * max(ref.rel_resid(m, xr), 1e-8)


def test_scaling_through_solve_iterative_and_mixed():
    m = fa.gen_hex8(7, kind="lognormal", seed=6)
    m.set_iter(5000)
    m.set_i(m.I_SCALING, 1)
    xr, ia, ra = ref.solve_iterative(m.copy())
    res = fa.hecmw_solve_iterative(m)
    assert res.error == 0
    assert np.linalg.norm(m.X - xr) / np.linalg.norm(xr) <= 1e-6
    m2 = m.copy()
    m2.X[:] = 0.0
    m2.set_mixed(True)
    res = fa.hecmw_solve_iterative(m2)
    assert res.error == 0
    assert np.linalg.norm(m2.X - xr) / np.linalg.norm(xr) <= 1e-6


@pytest.mark.parametrize("name", ["hex8_10", "random_700", "random_ext"])
def test_ssor_lane_per_row_layout_bitwise(name, monkeypatch):
    """The lane-per-row SSOR layout (SELL-64 slices, chosen for large colours)
    gives the same bits as the four-lane layout and the reference."""
    monkeypatch.setenv("HECMW33_SSOR_ROWS", "64")
    m = dict(SYSTEMS)[name].copy()
    m.set_precond(1)
    r = np.random.default_rng(2).standard_normal(3 * m.NP)
    z = fa.hecmw_precond_apply(m, r)
    assert np.array_equal(z, ref.precond(m, r, nthreads=2))
    m.set_iter(5000)
    res = fa.hecmw_solve_krylov(m)
    h = fa.last_history()
    _, info = ref.solve_krylov(m, nthreads=2, history=True)
    assert res.error == info["error"] == 0
    assert abs(res.iter - info["iter"]) <= 2
    k = min(15, len(h), len(info["history"]))
    np.testing.assert_allclose(h[:k], info["history"][:k], rtol=2e-6)
    # and through the mixed path
    m2 = m.copy()
    m2.X[:] = 0.0
    m2.set_mixed(True)
    res = fa.hecmw_solve_krylov(m2)
    assert res.error == 0 and ref.rel_resid(m2, m2.X) <= 1e-8 * 1.0001
