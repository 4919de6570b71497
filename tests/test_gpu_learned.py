"""The learned parts on the GPU: the node classifier network and the learned coarse modes.

* forward and backward BF16 MFMA kernels against the CPU restatement
  (oracle/learned_ref.py: same BF16 rounding points, FP64 accumulation), on
  node counts that are not multiples of the 16/32-node tiles; the differences
  are those of the FP32 accumulation order (and the BF16 roundings it flips);
* training on labels built from the solver's own errors (SSOR-CG iterates on
  seeded two-phase cubes) lowers the loss, and on a held-out cube the network
  classifies the nodes as their labels;
* the learned coarse modes (per-aggregate linear autoencoders fitted to the
  solver's own slow errors of the system) against the rigid-body modes alone:
  far fewer iterations, harvest included, the learned step kept, an FP64 true
  residual of 1e-8 and the reference's FP64 SSOR-CG solution within 1e-6;
* the fitted modes are orthogonal to their aggregate's rigid-body modes and
  the twelve-mode E assembled on the GPU equals the dense Z^T A Z.
"""
import numpy as np
import pytest

import frontistr_amd as fa
from oracle import learned_ref as L
from oracle import ref

pytestmark = [pytest.mark.gpu]


def _weights(rng):
    W1 = rng.standard_normal((16, 64)).astype(np.float32) * 0.4
    W1[12:] = 0
    W2 = rng.standard_normal((64, 64)).astype(np.float32) * 0.2
    W3 = np.zeros((64, 16), np.float32)
    W3[:, :3] = rng.standard_normal((64, 3)) * 0.3
    return W1, W2, W3


@pytest.mark.parametrize("n", [1, 37, 300])
def test_net_forward_backward_match_oracle(n):
    rng = np.random.default_rng(n)
    net = fa.LearnedNet(seed=1)
    W = _weights(rng)
    net.set_weights(*W)
    Wr = net.weights()
    for a, b in zip(W, Wr):
        assert np.array_equal(a, b)
    X = rng.standard_normal((n, 16)).astype(np.float32)
    X[:, 12:] = 0
    dY = rng.standard_normal((n, 3)).astype(np.float32)
    Y, dW = net.eval(X, dY)
    Yo, cache = L.forward(X, *W)
    gW = L.backward(dY, cache)
    # FP32 accumulation in another order than the oracle's FP64: most entries
    # agree to ~1e-6; an intermediate whose BF16 rounding flips moves its
    # entries by up to 2^-8 relative, which bounds the norms
    rel = np.abs(Y - Yo) / (np.abs(Yo) + 1e-3 * np.abs(Yo).max())
    assert np.median(rel) <= 1e-5, np.median(rel)
    assert np.linalg.norm(Y - Yo) <= 2e-3 * np.linalg.norm(Yo), np.linalg.norm(Y - Yo) / np.linalg.norm(Yo)
    for name, g, go in zip(("W1", "W2", "W3"), dW, gW):
        err = np.linalg.norm(g - go) / (np.linalg.norm(go) + 1e-12)
        assert err <= 5e-3, (name, err)
    assert np.all(dW[2][:, 3:] == 0)
    net.close()


def _cube(n, seed, cells=8):
    """two-phase cube with the 120^3 workload's inclusion statistics (E contrast
    1e4, 15% volume, radii 2-6 elements); cells of 12 elements (Iarray 43)"""
    m = fa.gen_hex8(n, kind="inclusions", seed=seed, E=1000.0, contrast=1e4, rmin=2, rmax=6, vfrac=0.15)
    m.set_iter(20000)
    m.set_precond(1)
    m.set_coarse(cells)
    return m


def _context(m, n, torch):
    ctx = fa.DeviceContext()
    ctx.set_coords(fa.hex8_coords(n))
    ctx.load(m)
    return ctx, torch.tensor(m.B, device="cuda:0")


def test_training_learns_the_labels():
    torch = pytest.importorskip("torch")
    net = fa.LearnedNet(seed=3)
    keep = []
    for seed in (11, 12, 13):
        m = _cube(20, seed)
        ctx, b = _context(m, 20, torch)
        assert fa.harvest_errors(net, ctx, m, b) == 2
        keep.append((ctx, m, b))
    loss = net.train(epochs=60, lr=3e-3, seed=0)
    print("logistic loss: before %.4f, after %.4f" % (loss[0], loss[-1]))
    assert loss[-1] < 0.5 * loss[0]
    # held out
    net.clear_samples()
    m = _cube(20, 17)
    ctx, b = _context(m, 20, torch)
    fa.harvest_errors(net, ctx, m, b)
    l, acc = net.score()
    # the same labels on the host (hecmw33_rigid_labels over the same two errors),
    # for the share of the majority class
    errs = _errors(ctx, m, b, torch, (10, 40))
    t = fa.rigid_labels(m, fa.hex8_coords(20), errs)
    major = max(t.mean(), 1.0 - t.mean())
    print("held-out: loss %.4f, accuracy %.4f, labelled rigid %.4f" % (l, acc, t.mean()))
    assert 0.0 < t.mean() < 1.0
    # well above what one class alone scores
    assert acc >= 0.9 and (1.0 - acc) <= 0.6 * (1.0 - major)
    net.close()


def _errors(ctx, m, b, torch, iters):
    """e_k = x* - x_k of the FP64 SSOR-CG iterates at `iters` (as harvest_errors takes them), on the host"""
    ia = m.Iarray.copy()
    ia[m.I_FLAG_NUMFACT - 1] = ia[m.I_FLAG_SYMBFACT - 1] = ia[m.I_SOLVER_OPT - 1] = 0
    xs = torch.zeros_like(b)
    assert ctx.solve_device(b.data_ptr(), xs.data_ptr(), ia, m.Rarray).error == 0
    x = torch.zeros_like(b)
    out, done = [], 0
    for k in iters:
        ia[m.I_ITER - 1] = k - done
        ctx.solve_device(b.data_ptr(), x.data_ptr(), ia, m.Rarray)
        done = k
        out.append((xs - x).cpu().numpy()[:3 * m.N])
    return np.array(out)


def _solve(ctx, m, b, torch, opt):
    mm = m.copy()
    mm.set_i(mm.I_SOLVER_OPT, opt)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    r = ctx.solve_device(b.data_ptr(), x.data_ptr(), mm.Iarray, mm.Rarray)
    return r, x.cpu().numpy(), ctx.info()


@pytest.mark.parametrize("n,cells,seed", [(24, 64, 7), (32, 64, 3)])
def test_learned_modes_gain_over_the_rigid_body_modes(n, cells, seed):
    """The flexible CG with the learned coarse modes (fitted on the solver's
    own slow errors of this system) against the same two-level step with the
    rigid-body modes alone: far fewer iterations, the learned step kept, an
    FP64 true residual of 1e-8 and the reference's FP64 SSOR-CG solution
    within 1e-6."""
    torch = pytest.importorskip("torch")
    m = _cube(n, seed, cells=cells)
    ctx, b = _context(m, n, torch)
    x_ref, info = ref.solve_krylov(m, nthreads=2)
    res = {}
    for name, opt in (("rigid", 1), ("learned", 2)):
        r, xs, inf = _solve(ctx, m, b, torch, opt)
        res[name] = (r.iter, inf)
        assert r.error == 0, (name, r.error)
        assert ref.rel_resid(m, xs) <= 1e-8 * 1.0001, name
        assert np.linalg.norm(xs - x_ref) <= 1e-6 * np.linalg.norm(x_ref), name
    it_r, inf_r = res["rigid"]
    it_l, inf_l = res["learned"]
    print("%d^3 two-phase cube: rigid-body modes %d iterations (%d coarse unknowns), learned modes %d iterations "
          "(%d coarse unknowns, %d modes kept, %d samples x %d iterations, harvest %.3fs, fit %.3fs); reference "
          "SSOR-CG %d" % (n, it_r, inf_r["coarse_unknowns"], it_l, inf_l["coarse_unknowns"], inf_l["learned_modes"],
                          inf_l["mode_samples"], inf_l["mode_sample_iters"], inf_l["mode_harvest_s"],
                          inf_l["mode_fit_s"], info["iter"]))
    assert inf_r["modes_per_aggregate"] == 6 and inf_l["modes_per_aggregate"] == 12
    assert inf_l["learned_off"] == 0 and inf_l["fallback_iter"] == 0
    assert inf_l["learned_modes"] > 0
    assert it_l + inf_l["mode_samples"] * inf_l["mode_sample_iters"] < it_r
    assert it_l < 0.6 * it_r
    ctx.close()


def test_learned_modes_and_their_coarse_matrix():
    """The fitted modes lie outside the rigid-body modes of their aggregate
    (orthogonal to them, as the autoencoder is fitted to what the rigid-body
    modes leave), have the norm of a translation mode, and the twelve-mode
    E = Z^T A Z assembled on the GPU equals the dense product from the
    library's own aggregation, offsets and modes."""
    torch = pytest.importorskip("torch")
    import scipy.sparse as sp
    from tests.systems import to_csr
    n = 12
    m = _cube(n, 5, cells=8)
    ctx, b = _context(m, n, torch)
    r, xs, inf = _solve(ctx, m, b, torch, 2)
    assert r.error == 0 and inf["modes_per_aggregate"] == 12
    co = ctx.coarse(with_E=False)
    md = ctx.modes()
    agg, dv, xm = co["agg"], co["dv"].astype(np.float64), md["xm"].astype(np.float64)
    N = m.N
    # Z (3N x 12 n_agg): rigid-body modes from the offsets, then the learned ones
    rows, cols, vals = [], [], []
    i = np.arange(N)
    def put(comp, col, v):
        rows.append(3 * i + comp); cols.append(12 * agg + col); vals.append(v)
    for k in range(3):
        put(k, k, np.ones(N))
    put(1, 3, -dv[:, 2]); put(2, 3, dv[:, 1])
    put(0, 4, dv[:, 2]); put(2, 4, -dv[:, 0])
    put(0, 5, -dv[:, 1]); put(1, 5, dv[:, 0])
    for j in range(6):
        for k in range(3):
            put(k, 6 + j, xm[:, j, k])
    Z = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(3 * N, 12 * co["n_agg"]))
    for a in range(co["n_agg"]):
        nodes = np.nonzero(agg == a)[0]
        dof = (3 * nodes[:, None] + np.arange(3)).ravel()
        Za = Z[dof][:, 12 * a:12 * a + 12].toarray()
        R, X = Za[:, :6], Za[:, 6:]
        live = np.linalg.norm(X, axis=0) > 0
        if not live.any():
            continue
        X = X[:, live]
        # orthogonal to the rigid-body modes (the samples' rigid part removed in FP64, FP32 storage)
        assert np.abs(R.T @ X).max() <= 1e-3 * np.linalg.norm(R, axis=0).max() * np.linalg.norm(X, axis=0).max(), a
        # norm of a translation mode
        assert np.allclose(np.linalg.norm(X, axis=0), np.sqrt(nodes.size), rtol=2e-2), a
    # the GPU's fit against the oracle's on the same samples (R as the GPU kept it; the oracle
    # rounds R to BF16 for the Gram matrix as the MFMA pass does): the spans agree
    ns = inf["mode_samples"]
    R = md["R"].transpose(2, 0, 1)[:ns].astype(np.float64)  # (ns, N, 3)
    xo = L.fit_modes(R, agg, gram_bf16=True)
    cosines = []
    for a in range(co["n_agg"]):
        nodes = np.nonzero(agg == a)[0]
        Xg = xm[nodes].transpose(1, 0, 2).reshape(6, -1).T
        Xo = xo[nodes].transpose(1, 0, 2).reshape(6, -1).T
        lg, lo = np.linalg.norm(Xg, axis=0) > 0, np.linalg.norm(Xo, axis=0) > 0
        assert lg.sum() == lo.sum(), a
        if not lg.any():
            continue
        Qg, _ = np.linalg.qr(Xg[:, lg])
        Qo, _ = np.linalg.qr(Xo[:, lo])
        cosines.append(np.linalg.svd(Qg.T @ Qo, compute_uv=False).min())
    print("learned modes vs oracle fit: smallest principal cosine per aggregate: min %.4f median %.4f"
          % (min(cosines), np.median(cosines)))
    assert np.median(cosines) >= 0.99 and min(cosines) >= 0.9
    A = to_csr(m)
    E = (Z.T @ (A @ Z)).toarray()
    for p, (a, bb) in enumerate(zip(co["pair_a"], co["pair_b"])):
        Eab = E[12 * a:12 * a + 12, 12 * bb:12 * bb + 12]
        scale = np.abs(E[12 * a:12 * a + 12, 12 * a:12 * a + 12]).max()
        assert np.abs(md["Eblk"][p] - Eab).max() <= 1e-9 * scale, (p, a, bb)
    # the coarse solve operator: L^-1 of the GPU's own E = L L^T, mirrored (L^-T above the
    # diagonal); y = L^-T L^-1 c solves E y = c to a backward error of FP64 rounding
    W, f64 = ctx.coarse_inverse()
    ctx.close()
    nc = 12 * co["n_agg"]
    assert f64 and W.shape == (nc, nc)
    assert np.array_equal(np.triu(W, 1), np.tril(W, -1).T)
    Eg = np.zeros((nc, nc))
    for p, (a, bb) in enumerate(zip(co["pair_a"], co["pair_b"])):
        Eg[12 * a:12 * a + 12, 12 * bb:12 * bb + 12] = md["Eblk"][p]
    Li = np.tril(W)
    c = np.random.default_rng(3).standard_normal(nc)
    y = Li.T @ (Li @ c)
    berr = np.linalg.norm(Eg @ y - c) / (np.linalg.norm(Eg, 2) * np.linalg.norm(y))
    assert berr <= 1e-12, berr


def test_learned_solve_refuses_the_additive_step_and_missing_coordinates():
    torch = pytest.importorskip("torch")
    m = _cube(8, 1, cells=1)
    mm = m.copy()
    mm.set_i(mm.I_SOLVER_OPT, 2)
    mm.set_i(44, 1)  # the additive two-level step
    ctx, b = _context(m, 8, torch)
    x = torch.zeros_like(b)
    with pytest.raises(fa.HecmwError):
        ctx.solve_device(b.data_ptr(), x.data_ptr(), mm.Iarray, mm.Rarray)
    ctx.close()
    ctx = fa.DeviceContext()  # no coordinates
    ctx.load(m)
    mm.set_i(44, 0)
    with pytest.raises(fa.HecmwError):
        ctx.solve_device(b.data_ptr(), x.data_ptr(), mm.Iarray, mm.Rarray)
    ctx.close()


def test_score_is_the_loss_training_starts_from():
    """hecmw33_net_score and the first entry of hecmw33_net_train both evaluate
    the mean logistic loss of the current weights on the samples; with a zero
    learning rate the weights stay, so the two agree."""
    torch = pytest.importorskip("torch")
    net = fa.LearnedNet(seed=9)
    m = _cube(12, 3, cells=1)
    ctx, b = _context(m, 12, torch)
    fa.harvest_errors(net, ctx, m, b)
    l0, acc0 = net.score()
    assert np.isfinite(l0) and l0 > 0.0 and 0.0 <= acc0 <= 1.0
    loss = net.train(epochs=1, lr=0.0, seed=0)
    assert abs(loss[0] - l0) <= 1e-9 * l0
    assert abs(net.score()[0] - l0) <= 1e-9 * l0
    net.close()
