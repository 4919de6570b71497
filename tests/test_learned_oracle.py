"""The learned correction's CPU oracle pinned on its own: the BF16 rounding
against known bit patterns, the gradient formulas against central finite
differences (a dropped ReLU mask, a transposed weight or a missing layer
fails), and a planted mistake that the check detects."""
import numpy as np
import pytest

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


# ---- training labels and loss of the learned aggregation ------------------------------
def _region_sample(m, X, centre, radius, omega, stretch):
    """e = omega x (x - centre) inside the ball (a rigid rotation), a uniform
    stretch diag(stretch) x outside: edges with both ends inside have zero
    axial strain, the others a strain of the order of the stretch."""
    inside = np.linalg.norm(X - centre, axis=1) < radius
    e = X * np.asarray(stretch)
    e[inside] = np.cross(omega, X[inside] - centre)
    return e.ravel(), inside


def _expected(m, inside_sets):
    """(must, may): the nodes of the edges whose ends both lie inside the region
    of every sample, which must be labelled; the nodes with a neighbour (or
    themselves) inside, the only ones that may be (an edge that crosses the
    region's boundary has an arbitrary strain)."""
    must = np.zeros(m.N, bool)
    near = np.zeros(m.N, bool)
    both = np.logical_and.reduce(inside_sets)
    for idx, item in ((m.indexL, m.itemL), (m.indexU, m.itemU)):
        for i in range(m.N):
            for q in range(idx[i], idx[i + 1]):
                j = item[q] - 1
                if j >= m.N:
                    continue
                if both[i] and both[j]:
                    must[i] = must[j] = True
                if both[i] or both[j]:
                    near[i] = near[j] = True
    return must, near


def _check(lab, must, near):
    lab = lab > 0.5
    assert np.all(lab[must]), "an edge inside the rigid region is not labelled"
    assert not np.any(lab[~near]), "a node away from the rigid region is labelled"


def test_rigid_labels_mark_the_rigid_region():
    import frontistr_amd as fa
    m = fa.gen_hex8_structure(8)
    X = fa.hex8_coords(8).reshape(-1, 3)
    e, inside = _region_sample(m, X, np.array([0.45, 0.5, 0.55]), 0.3, [0.3, -0.2, 0.5], [1.0, 2.0, 3.0])
    must, near = _expected(m, [inside])
    assert must.sum() > 20 and (~near).sum() > 300
    got_ref = L.rigid_labels(m.N, m.indexL, m.itemL, m.indexU, m.itemU, X, e, 0.1)
    got_lib = fa.rigid_labels(m, X, e[None, :], 0.1)
    assert np.array_equal(got_lib, got_ref)
    _check(got_lib, must, near)


def test_rigid_labels_need_every_sample_and_skip_fixed_nodes():
    import frontistr_amd as fa
    m = fa.gen_hex8_structure(8)
    X = fa.hex8_coords(8).reshape(-1, 3)
    e1, in1 = _region_sample(m, X, np.array([0.4, 0.5, 0.5]), 0.32, [0.3, -0.2, 0.5], [1.0, 2.0, 3.0])
    e2, in2 = _region_sample(m, X, np.array([0.6, 0.5, 0.5]), 0.32, [-0.1, 0.4, 0.2], [3.0, 1.0, 2.0])
    # a fixed bottom face: e = 0 there in both samples, so its edges say nothing
    bottom = X[:, 2] == 0.0
    for e in (e1, e2):
        e.reshape(-1, 3)[bottom] = 0.0
    must, near = _expected(m, [in1, in2])
    got_lib = fa.rigid_labels(m, X, np.stack([e1, e2]), 0.1)
    got_ref = L.rigid_labels(m.N, m.indexL, m.itemL, m.indexU, m.itemU, X, np.stack([e1, e2]), 0.1)
    assert np.array_equal(got_lib, got_ref)
    _check(got_lib, must, near)
    # the second sample removes what only the first region holds
    only1 = _expected(m, [in1])[0] & ~_expected(m, [in1, in2])[1]
    assert only1.sum() > 0 and not np.any(got_lib[only1] > 0.5)
    assert not np.any(got_lib[bottom & ~near] > 0.5)


def test_rigid_labels_planted_mistake_is_caught():
    """Taking the whole relative motion |e_j - e_i| instead of its component
    along the edge marks nothing inside a rotating piece: the region check fails."""
    import frontistr_amd as fa
    m = fa.gen_hex8_structure(8)
    X = fa.hex8_coords(8).reshape(-1, 3)
    e, inside = _region_sample(m, X, np.array([0.45, 0.5, 0.55]), 0.3, [0.3, -0.2, 0.5], [1.0, 2.0, 3.0])
    must, near = _expected(m, [inside])
    E = e.reshape(-1, 3)
    rows = np.repeat(np.arange(m.N), np.diff(m.indexU))
    cols = m.itemU - 1
    d = X[cols] - X[rows]
    wrong = np.linalg.norm(E[cols] - E[rows], axis=1) / np.linalg.norm(d, axis=1)
    lab = np.zeros(m.N, np.float32)
    small = wrong / np.median(wrong) < 0.1
    lab[rows[small]] = 1
    lab[cols[small]] = 1
    with pytest.raises(AssertionError):
        _check(lab, must, near)


def test_logistic_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    Y = rng.standard_normal((9, 3))
    t = (rng.random(9) > 0.5).astype(np.float64)
    loss, dY = L.logistic(Y, t)
    h = 1e-6
    for i in range(9):
        Yp, Ym = Y.copy(), Y.copy()
        Yp[i, 0] += h
        Ym[i, 0] -= h
        fd = (L.logistic(Yp, t)[0] - L.logistic(Ym, t)[0]) / (2 * h)
        assert abs(fd - dY[i, 0]) <= 1e-7
    assert np.all(dY[:, 1:] == 0)
    # closed form at y = 0: log 2 whatever the labels
    assert abs(L.logistic(np.zeros((4, 3)), np.array([0, 1, 0, 1]))[0] - np.log(2.0)) <= 1e-15


# ---- learned coarse modes ---------------------------------------------------------------------
def _aggregates(rng, n=400, na=5):
    """na aggregates of random points with their scaled offsets from the centroid"""
    X = rng.standard_normal((n, 3))
    agg = rng.integers(0, na, n)
    dv = np.zeros_like(X)
    for a in range(na):
        m = agg == a
        dv[m] = X[m] - X[m].mean(0)
    return agg, dv / np.abs(dv).max()


def test_mode_fit_recovers_planted_fields_beyond_the_rigid_modes():
    """Samples = rigid-body motions + combinations of three fields per
    aggregate: the fit removes the rigid part exactly, and its modes span the
    planted fields' part beyond the rigid-body modes, with the norm of a
    translation mode; modes past the planted rank stay zero."""
    rng = np.random.default_rng(3)
    agg, dv = _aggregates(rng)
    N, ns = agg.size, 10
    Z = L.rigid_modes(dv)
    fields = rng.standard_normal((3, N, 3))
    S = np.zeros((ns, N, 3))
    for s in range(ns):
        for a in np.unique(agg):
            m = agg == a
            S[s, m] = (Z[m] @ rng.standard_normal(6)) + np.einsum("f,fnd->nd", rng.standard_normal(3), fields[:, m])
    R = L.remove_rigid(S, agg, dv)
    xm = L.fit_modes(R, agg)
    for a in np.unique(agg):
        m = np.nonzero(agg == a)[0]
        Za = Z[m].reshape(-1, 6)
        # the rigid part is gone from the samples and absent from the modes
        assert np.abs(Za.T @ R[:, m].reshape(ns, -1).T).max() <= 1e-9 * np.abs(R).max() * m.size
        X = xm[m].transpose(1, 0, 2).reshape(6, -1).T  # (3 n_a, 6)
        live = np.linalg.norm(X, axis=0) > 0
        assert live.sum() == 3, live
        assert np.allclose(np.linalg.norm(X[:, live], axis=0), np.sqrt(m.size))
        # the planted fields beyond the rigid modes lie in the modes' span
        F = fields[:, m].reshape(3, -1).T
        F = F - Za @ np.linalg.lstsq(Za, F, rcond=None)[0]
        Q, _ = np.linalg.qr(X[:, live])
        assert np.linalg.norm(F - Q @ (Q.T @ F)) <= 1e-9 * np.linalg.norm(F)


def test_mode_fit_catches_a_kept_rigid_part():
    """A fit that skips the rigid-body removal gives modes that are not
    orthogonal to the rigid-body modes (what the GPU test checks)."""
    rng = np.random.default_rng(4)
    agg, dv = _aggregates(rng, na=3)
    Z = L.rigid_modes(dv)
    S = np.stack([Z @ rng.standard_normal(6) + 0.1 * rng.standard_normal((agg.size, 3)) for _ in range(8)])
    bad = L.fit_modes(S, agg)
    good = L.fit_modes(L.remove_rigid(S, agg, dv), agg)
    for xm, ok in ((good, True), (bad, False)):
        worst = 0.0
        for a in np.unique(agg):
            m = np.nonzero(agg == a)[0]
            X = xm[m].transpose(1, 0, 2).reshape(6, -1).T
            worst = max(worst, np.abs(Z[m].reshape(-1, 6).T @ X).max() / np.sqrt(m.size))
        assert (worst <= 1e-8) == ok, worst


def test_mode_fit_bf16_gram_keeps_the_span():
    """The GPU forms the Gram matrix from BF16-rounded samples: the fitted
    span moves by about the BF16 rounding, not more."""
    rng = np.random.default_rng(5)
    agg, dv = _aggregates(rng, na=2)
    S = rng.standard_normal((12, agg.size, 3)) * np.linspace(1.0, 0.05, 12)[:, None, None]
    R = L.remove_rigid(S, agg, dv)
    x64 = L.fit_modes(R, agg)
    x16 = L.fit_modes(R, agg, gram_bf16=True)
    for a in np.unique(agg):
        m = np.nonzero(agg == a)[0]
        A = x64[m].transpose(1, 0, 2).reshape(6, -1).T
        B = x16[m].transpose(1, 0, 2).reshape(6, -1).T
        Qa, _ = np.linalg.qr(A)
        Qb, _ = np.linalg.qr(B)
        # principal angles between the two 6-dimensional spans
        s = np.linalg.svd(Qa.T @ Qb, compute_uv=False)
        assert s.min() >= 0.98, s
