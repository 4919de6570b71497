"""The learned correction's CPU oracle pinned on its own: the BF16 rounding
against known bit patterns, the gradient formulas against central finite
differences (a dropped ReLU mask, a transposed weight or a missing layer
fails), and a planted mistake that the check detects."""
import numpy as np

from oracle import learned_ref as L


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.0e-39, np.inf], np.float32)
    b = L.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3*2^-8 ties up to 1 + 2^-6
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == np.float32(1.0 + 2 ** -6) and b[3] == np.float32(1.0 + 2 ** -7)
    assert b[4] == -2.5 and b[6] == np.inf
    assert np.all((L.bf16(x).view(np.uint32) & 0xFFFF) == 0)


def _params(rng):
    W1 = rng.standard_normal((16, 64)) * 0.4
    W1[12:] = 0
    W2 = rng.standard_normal((64, 64)) * 0.2
    W3 = np.zeros((64, 16))
    W3[:, :3] = rng.standard_normal((64, 3)) * 0.3
    return W1, W2, W3


def _fd_check(backward):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((7, 16))
    X[:, 12:] = 0
    W = list(_params(rng))
    dY = rng.standard_normal((7, 3))
    Y, cache = L.forward64(X, *W)
    grads = backward(dY, cache)
    h = 1e-6
    worst = 0.0
    for wi in range(3):
        for idx in [(0, 0), (3, 5), (11, 63), (63, 2), (10, 1)]:
            if idx[0] >= W[wi].shape[0] or idx[1] >= W[wi].shape[1]:
                continue
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[wi][idx] += h
            Wm[wi][idx] -= h
            fd = (np.sum(dY * L.forward64(X, *Wp)[0]) - np.sum(dY * L.forward64(X, *Wm)[0])) / (2 * h)
            worst = max(worst, abs(fd - grads[wi][idx]) / (1e-6 + abs(fd)))
    return worst


def test_gradients_match_finite_differences():
    assert _fd_check(L.backward64) < 1e-5


def test_planted_mistake_is_caught():
    def wrong(dY, cache):
        X, W1, W2, W3, H1p, H1, H2p, H2 = cache
        dYp = np.zeros((dY.shape[0], W3.shape[1]))
        dYp[:, :3] = dY
        dH2 = dYp @ W3.T  # ReLU mask dropped
        dH1 = (dH2 @ W2.T) * (H1p > 0)
        return X.T @ dH1, H1.T @ dH2, H2.T @ dYp
    assert _fd_check(wrong) > 1e-2


def test_rounded_version_tracks_exact():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((50, 16))
    W = _params(rng)
    Y, c = L.forward(X, *W)
    Y64, c64 = L.forward64(X, *W)
    assert np.linalg.norm(Y - Y64) <= 2e-2 * np.linalg.norm(Y64)
    dY = rng.standard_normal((50, 3))
    for g, g64 in zip(L.backward(dY, c), L.backward64(dY, c64)):
        # BF16 rounding of the pre-activations flips a few ReLU masks
        assert np.linalg.norm(g - g64) <= 0.1 * np.linalg.norm(g64)
