"""Pins the oracle restatement (oracle/restate.c) before anything is compared
with it:

* bitwise against the reference's own compiled solver (oracle/_ref) on
  orderings, SpMV, SSOR/DIAG applications and whole CG/BiCGSTAB solves,
  including the breakdown paths;
* against committed golden vectors written from the reference
  (tests/golden/make_golden.py), which hold without the reference checkout;
* against what the mathematics fixes: the dense product, the block-Jacobi
  inverse, (D+L) D^-1 (D+U) in the multicolour order for SSOR, a direct solve,
  and the race-freedom of the colouring;
* and the pins themselves are shown to catch planted mistakes (a transposed
  block, a reversed sweep order, a dropped colour boundary).
"""
import os

import numpy as np
import pytest

import frontistr_amd as fa
from oracle import ref, restate
from tests.systems import dense, random_system, shuffle_items

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hecmw33_golden.npz")


def systems():
    yield "hex8_5", fa.gen_hex8(5)
    yield "hex8_6_shuffled", shuffle_items(fa.gen_hex8(6), seed=2)
    yield "random_200", random_system(200, deg=10, seed=4)
    yield "random_ext", random_system(160, deg=8, seed=6, n_ext=15)
    yield "random_disconnected", random_system(180, deg=6, seed=8, components=3)


SYS = list(systems())
IDS = [s[0] for s in SYS]


# ---- restatement == reference build (bitwise) ------------------------------------
@needs_ref
@pytest.mark.parametrize("name,m", SYS, ids=IDS)
def test_restatement_ordering_equals_reference(name, m):
    for nc in (10, 3):
        pr, p, ip, ci = restate.ordering(m, nc)
        rpr, _, rp, rip, rci = ref.ordering(m, nc)
        assert np.array_equal(pr, rpr) and np.array_equal(p, rp) and np.array_equal(ci, rci)


@needs_ref
@pytest.mark.parametrize("name,m", SYS, ids=IDS)
def test_restatement_matvec_and_precond_equal_reference(name, m):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(3 * m.NP)
    assert np.array_equal(restate.matvec(m, x), ref.matvec(m, x))
    for pc in (1, 3):
        for sigma in (1.0, 1.7):
            for ipm in (1, 2):
                mm = m.copy()
                mm.set_precond(pc)
                mm.set_sigma_diag(sigma)
                mm.set_iterpremax(ipm)
                assert np.array_equal(restate.precond(mm, x), ref.precond(mm, x, nthreads=2)), (pc, sigma, ipm)


@needs_ref
@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("precond", [1, 3, 0])
@pytest.mark.parametrize("name,m", SYS[:4], ids=IDS[:4])
def test_restatement_solve_equals_reference(name, m, method, precond):
    mm = m.copy()
    mm.set_iter(3000)
    mm.set_method(method)
    if precond == 0:
        mm.set_iterpremax(0)
    else:
        mm.set_precond(precond)
    x1, i1 = restate.solve_krylov(mm)
    x2, i2 = ref.solve_krylov(mm, nthreads=2, history=True)
    assert (i1["iter"], i1["error"]) == (i2["iter"], i2["error"])
    assert i1["resid"] == i2["resid"]
    assert np.array_equal(x1, x2)
    np.testing.assert_allclose(i1["history"], i2["history"], rtol=6e-7)  # printed with 7 digits


@needs_ref
@pytest.mark.parametrize("case", ["noconv", "zero_rhs", "indefinite", "nan", "indefinite_pc"])
def test_restatement_error_paths_equal_reference(case):
    m = fa.gen_hex8(4)
    m.set_iter(2000)
    if case == "noconv":
        m.set_iter(6)
    elif case == "zero_rhs":
        m.B[:] = 0.0
    elif case == "indefinite":
        m.set_precond(3)
        m.D[:] = -m.D
        m.AL[:] = -m.AL
        m.AU[:] = -m.AU
    elif case == "nan":
        m.B[7] = np.nan
    elif case == "indefinite_pc":
        m = fa.gen_hex8(8)
        m.set_iter(2000)
        m.set_precond(3)
        m.set_iterpremax(2)
    for method in (1, 2):
        mm = m.copy()
        mm.set_method(method)
        x0 = np.ones(3 * m.NP) if case == "zero_rhs" else None
        x1, i1 = restate.solve_krylov(mm, x0=x0)
        x2, i2 = ref.solve_krylov(mm, x0=x0, nthreads=2)
        assert (i1["iter"], i1["error"]) == (i2["iter"], i2["error"]), (method, i1, i2)
        assert np.array_equal(x1, x2, equal_nan=True)


# ---- golden vectors (no reference checkout needed) ---------------------------------
def golden_systems():
    g = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in g.files})
    for n in names:
        N, NP = g[n + "/sizes"]
        m = fa.HecMatrix(int(N), g[n + "/D"], g[n + "/AL"], g[n + "/AU"], g[n + "/indexL"], g[n + "/itemL"],
                         g[n + "/indexU"], g[n + "/itemU"], B=g[n + "/B"], NP=int(NP))
        yield n, m, {k[len(n) + 1:]: g[k] for k in g.files if k.startswith(n + "/")}


@pytest.mark.parametrize("name,m,gold", list(golden_systems()), ids=[s[0] for s in golden_systems()])
def test_restatement_reproduces_golden(name, m, gold):
    pr, p, ip, ci = restate.ordering(m, 10)
    assert np.array_equal(pr, gold["perm_rcm"]) and np.array_equal(p, gold["perm"])
    assert np.array_equal(ci, gold["colorindex"])
    assert np.array_equal(restate.matvec(m, gold["x"]), gold["Ax"])
    for pc in (1, 3):
        mm = m.copy()
        mm.set_precond(pc)
        assert np.array_equal(restate.precond(mm, gold["r"]), gold["z_pc%d" % pc])
    for meth in (1, 2):
        for pc in (1, 3):
            mm = m.copy()
            mm.set_iter(2000)
            mm.set_method(meth)
            mm.set_precond(pc)
            x, info = restate.solve_krylov(mm)
            key = "solve_m%d_pc%d/" % (meth, pc)
            assert [info["iter"], info["error"]] == gold[key + "info"].tolist()
            assert np.array_equal(x, gold[key + "x"])
            np.testing.assert_allclose(info["history"], gold[key + "history"], rtol=6e-7)


# ---- mathematics --------------------------------------------------------------------
@pytest.mark.parametrize("name,m", SYS, ids=IDS)
def test_matvec_is_the_dense_product(name, m):
    x = np.random.default_rng(5).standard_normal(3 * m.NP)
    yd = dense(m) @ x
    np.testing.assert_allclose(restate.matvec(m, x), yd[: 3 * m.N], rtol=1e-13, atol=1e-13 * np.abs(yd).max())


@pytest.mark.parametrize("sigma", [1.0, 1.5])
@pytest.mark.parametrize("name,m", SYS, ids=IDS)
def test_diag_is_the_block_jacobi_inverse(name, m, sigma):
    mm = m.copy()
    mm.set_precond(3)
    mm.set_sigma_diag(sigma)
    r = np.random.default_rng(6).standard_normal(3 * m.NP)
    z = restate.precond(mm, r)
    for i in range(m.N):
        Db = m.D[9 * i:9 * i + 9].reshape(3, 3).copy()
        Db[np.diag_indices(3)] *= sigma
        np.testing.assert_allclose(z[3 * i:3 * i + 3], np.linalg.solve(Db, r[3 * i:3 * i + 3]), rtol=1e-11,
                                   atol=1e-14)


def ssor_dense(m, r, ncolor_in=10):
    """z = (D+U)^-1 D (D+L)^-1 r with L/U split by the multicolour order:
    symmetric block Gauss-Seidel, internal rows only."""
    _, p, _, _ = restate.ordering(m, ncolor_in)
    n3 = 3 * m.N
    A = dense(m)[:n3, :n3]
    idx = np.concatenate([[3 * (k - 1), 3 * (k - 1) + 1, 3 * (k - 1) + 2] for k in p])
    Ap = A[np.ix_(idx, idx)]
    Dp = np.zeros_like(Ap)
    for i in range(m.N):
        Dp[3 * i:3 * i + 3, 3 * i:3 * i + 3] = Ap[3 * i:3 * i + 3, 3 * i:3 * i + 3]
    lower = np.tril(Ap, -1) - np.tril(Dp, -1)
    upper = np.triu(Ap, 1) - np.triu(Dp, 1)
    rp = r[:n3][idx]
    # block-triangular (the 3x3 diagonal blocks are full): general solves
    y = np.linalg.solve(Dp + lower, rp)
    zp = np.linalg.solve(Dp + upper, Dp @ y)
    z = np.empty(n3)
    z[idx] = zp
    return z


@pytest.mark.parametrize("name,m", SYS, ids=IDS)
def test_ssor_is_symmetric_gauss_seidel_in_colour_order(name, m):
    mm = m.copy()
    mm.set_precond(1)
    r = np.random.default_rng(7).standard_normal(3 * m.NP)
    z = restate.precond(mm, r)
    np.testing.assert_allclose(z, ssor_dense(m, r), rtol=1e-10, atol=1e-12 * np.abs(z).max())


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("name,m", SYS[:4], ids=IDS[:4])
def test_solution_solves_the_system(name, m, method):
    mm = m.copy()
    mm.set_iter(5000)
    mm.set_method(method)
    x0 = np.zeros(3 * m.NP)
    x0[3 * m.N:] = 0.5  # external entries are inputs, kept as given
    x, info = restate.solve_krylov(mm, x0=x0)
    assert info["error"] == 0 and info["resid"] <= 1e-8
    n3 = 3 * m.N
    A = dense(m)
    rhs = m.B[:n3] - A[:n3, n3:] @ x0[n3:]
    xs = np.linalg.solve(A[:n3, :n3], rhs)
    assert np.linalg.norm(x[:n3] - xs) / np.linalg.norm(xs) < 1e-6


# ---- the pins catch planted mistakes -------------------------------------------------
def test_pins_catch_a_transposed_block():
    m = random_system(60, deg=6, seed=9)
    bad = m.copy()
    bad.AL = bad.AL.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1).copy()  # L blocks transposed
    x = np.random.default_rng(1).standard_normal(3 * m.NP)
    yd = (dense(m) @ x)[: 3 * m.N]
    assert not np.allclose(restate.matvec(bad, x), yd, rtol=1e-8)


def test_pins_catch_a_wrong_sweep_order():
    m = fa.gen_hex8(4)
    m.set_precond(1)
    r = np.random.default_rng(2).standard_normal(3 * m.NP)
    z = restate.precond(m, r)
    # natural order instead of the multicolour order is a different preconditioner
    _, p, _, _ = restate.ordering(m, 10)
    n3 = 3 * m.N
    A = dense(m)[:n3, :n3]
    Dm = np.zeros_like(A)
    for i in range(m.N):
        Dm[3 * i:3 * i + 3, 3 * i:3 * i + 3] = A[3 * i:3 * i + 3, 3 * i:3 * i + 3]
    L = np.tril(A, -1) - np.tril(Dm, -1)
    U = np.triu(A, 1) - np.triu(Dm, 1)
    y = np.linalg.solve(Dm + L, r[:n3])
    z_nat = np.linalg.solve(Dm + U, Dm @ y)
    assert not np.allclose(z, z_nat, rtol=1e-6)
    np.testing.assert_allclose(z, ssor_dense(m, r), rtol=1e-10, atol=1e-12 * np.abs(z).max())


def test_pins_catch_a_colour_that_is_not_independent():
    m = fa.gen_hex8(3)
    _, p, _, ci = restate.ordering(m, 10)
    merged = np.delete(ci, 1)  # first two colours merged: no longer independent
    color = np.zeros(m.N + 1, np.int64)
    for c in range(len(merged) - 1):
        color[p[merged[c]:merged[c + 1]]] = c
    clash = any(color[i] == color[j] for i in range(1, m.N + 1)
                for j in m.itemL[m.indexL[i - 1]:m.indexL[i]])
    assert clash
