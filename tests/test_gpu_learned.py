"""The learned aggregation on the GPU.

* forward and backward BF16 MFMA kernels against the CPU restatement
  (oracle/learned_ref.py: same BF16 rounding points, FP64 accumulation), on
  node counts that are not multiples of the 16/32-node tiles; the differences
  are those of the FP32 accumulation order (and the BF16 roundings it flips);
* training on labels built from the solver's own errors (SSOR-CG iterates on
  seeded two-phase cubes) lowers the loss, and on a held-out cube the network
  classifies the nodes as their labels;
* on a held-out cube the learned two-level flexible CG takes fewer iterations
  than the same solve on the cells alone, keeps its learned step, reaches an
  FP64 true residual of 1e-8 and the reference's FP64 SSOR-CG solution within
  1e-6.
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


def test_learned_aggregation_gains_on_a_held_out_cube():
    torch = pytest.importorskip("torch")
    net = fa.LearnedNet(seed=5)
    keep = []
    for seed in (21, 22, 23):
        m = _cube(24, seed)
        ctx, b = _context(m, 24, torch)
        fa.harvest_errors(net, ctx, m, b)
        keep.append(ctx)
    net.train(epochs=60, lr=3e-3, seed=1)
    net.clear_samples()
    m = _cube(24, 7)
    ctx, b = _context(m, 24, torch)
    ctx.set_net(net)
    x_ref, info = ref.solve_krylov(m, nthreads=2)
    res = {}
    for name, opt in (("cells", 1), ("learned", 2)):
        mm = m.copy()
        mm.set_i(mm.I_SOLVER_OPT, opt)
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        r = ctx.solve_device(b.data_ptr(), x.data_ptr(), mm.Iarray, mm.Rarray)
        xs = x.cpu().numpy()
        inf = ctx.info()
        res[name] = (r.iter, inf)
        assert r.error == 0
        assert ref.rel_resid(m, xs) <= 1e-8 * 1.0001
        assert np.linalg.norm(xs - x_ref) <= 1e-6 * np.linalg.norm(x_ref)
    it_c, inf_c = res["cells"]
    it_l, inf_l = res["learned"]
    print("held-out 24^3: cells %d iterations (%d aggregates), learned %d iterations (%d aggregates, %d pieces); "
          "reference SSOR-CG %d" % (it_c, inf_c["coarse"], it_l, inf_l["coarse"], inf_l["pieces"], info["iter"]))
    assert inf_l["pieces"] > 0 and inf_l["learned_off"] == 0 and inf_l["fallback_iter"] == 0
    assert it_l < 0.8 * it_c
    ctx.set_net(None)
    net.close()


def test_learned_solve_needs_a_network():
    torch = pytest.importorskip("torch")
    m = _cube(8, 1, cells=1)
    ctx, b = _context(m, 8, torch)
    mm = m.copy()
    mm.set_i(mm.I_SOLVER_OPT, 2)
    x = torch.zeros_like(b)
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
