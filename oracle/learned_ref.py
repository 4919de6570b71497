"""CPU restatement of the learned aggregation's network and its training
labels (test infrastructure only: tests and the benchmark's checks import it,
the product never does).

The network of frontistr_amd/csrc/learned.hip: per node, features x (16) ->
H1 = relu(x W1) (64) -> H2 = relu(H1 W2) (64) -> y = H2 W3[:, :3]; no biases.
The GPU kernels round their MFMA operands to BF16 (round to nearest even):
the features, the weights, the post-ReLU activations, the upstream gradient
and the back-propagated activation gradients; products accumulate in FP32.
`forward`/`backward` here apply the same roundings and accumulate in FP64,
so they differ from the GPU only by the FP32 accumulation order (and the rare
BF16 rounding flip that order can cause).

`forward64`/`backward64` are the same formulas without rounding, pinned by
finite differences in tests/test_learned_oracle.py.  `logistic` is the
training loss of output 0 against labels, `rigid_labels` the labels
(frontistr_amd/csrc/solver.hip, rigid_edge_ratios / rigid_labels_from_ratios).
There is no reference counterpart: FrontISTR has no learned preconditioner.
"""
import numpy as np


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def relu(a):
    return np.maximum(a, 0.0)


def forward(X, W1, W2, W3):
    """Returns (Y [n,3], cache) with the GPU's BF16 rounding points."""
    Xb, W1b, W2b, W3b = (bf16(X).astype(np.float64), bf16(W1).astype(np.float64), bf16(W2).astype(np.float64),
                         bf16(W3).astype(np.float64))
    H1 = bf16(relu(Xb @ W1b)).astype(np.float64)
    H2 = bf16(relu(H1 @ W2b)).astype(np.float64)
    Y = H2 @ W3b
    return Y[:, :3], (Xb, W1b, W2b, W3b, H1, H2)


def backward(dY, cache):
    """Weight gradients of sum(dY * Y) with the GPU's rounding points."""
    Xb, W1b, W2b, W3b, H1, H2 = cache
    n = dY.shape[0]
    dYp = np.zeros((n, W3b.shape[1]))
    dYp[:, :3] = bf16(dY).astype(np.float64)
    dH2 = bf16((dYp @ W3b.T) * (H2 > 0)).astype(np.float64)
    dH1 = bf16((dH2 @ W2b.T) * (H1 > 0)).astype(np.float64)
    return Xb.T @ dH1, H1.T @ dH2, H2.T @ dYp


def forward64(X, W1, W2, W3):
    H1p = X @ W1
    H1 = relu(H1p)
    H2p = H1 @ W2
    H2 = relu(H2p)
    return (H2 @ W3)[:, :3], (X, W1, W2, W3, H1p, H1, H2p, H2)


def backward64(dY, cache):
    X, W1, W2, W3, H1p, H1, H2p, H2 = cache
    dYp = np.zeros((dY.shape[0], W3.shape[1]))
    dYp[:, :3] = dY
    dH2 = (dYp @ W3.T) * (H2p > 0)
    dH1 = (dH2 @ W2.T) * (H1p > 0)
    return X.T @ dH1, H1.T @ dH2, H2.T @ dYp


def logistic(Y, t):
    """Mean logistic loss of Y[:, 0] against labels t and its gradient dL/dY (n x 3)."""
    y = np.asarray(Y, np.float64)[:, 0]
    t = np.asarray(t, np.float64)
    n = y.size
    loss = np.mean(np.logaddexp(0.0, y) - t * y)
    dY = np.zeros((n, 3))
    dY[:, 0] = (1.0 / (1.0 + np.exp(-y)) - t) / n
    return loss, dY


def rigid_labels(N, indexL, itemL, indexU, itemU, X, errors, tau):
    """Labels of the nodes of every stored edge (i, j), both internal, whose
    axial strain |(e_j - e_i).(x_j - x_i)| / |x_j - x_i|^2, divided by the
    median over the sample's edges, stays below tau in every error sample;
    edges whose e is zero at both ends are left out (median included)."""
    X = np.asarray(X, np.float64).reshape(-1, 3)
    rows, cols = [], []
    for idx, item in ((indexL, itemL), (indexU, itemU)):
        cnt = np.diff(np.asarray(idx[:N + 1]))
        rows.append(np.repeat(np.arange(N), cnt))
        cols.append(np.asarray(item[:idx[N]]) - 1)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    internal = cols < N
    worst = np.zeros(rows.size)
    for e in np.asarray(errors, np.float64).reshape(-1, 3 * N):
        e = e.reshape(-1, 3)
        ratio = np.full(rows.size, np.inf)
        ii, jj = rows[internal], cols[internal]
        d = X[jj] - X[ii]
        de = e[jj] - e[ii]
        mag = np.abs(e[jj]).sum(1) + np.abs(e[ii]).sum(1)
        den = (d * d).sum(1)
        live = (mag > 0) & (den > 0)
        eps = np.abs((de * d).sum(1)) / np.where(den > 0, den, 1.0)
        med = np.sort(eps[live])[live.sum() // 2] if live.any() else 0.0
        r = np.where(eps > 0, np.inf, 0.0) if med == 0 else eps / med
        sub = np.where(live, r, np.inf)
        ratio[np.nonzero(internal)[0]] = sub
        worst = np.maximum(worst, ratio)
    lab = np.zeros(N, np.float32)
    small = worst < tau
    lab[rows[small]] = 1
    lab[cols[small]] = 1
    return lab


# ---- learned coarse modes -------------------------------------------------------------------
# The fit of frontistr_amd/csrc/solver.hip (learned_modes) and learned.hip
# (k_modes_rigid / k_modes_resid / k_modes_gram / k_modes_make), in FP64:
# per aggregate, the samples' least-squares rigid-body part (through the
# eigenpairs of Z^T Z, directions below 1e-12 of its largest eigenvalue
# dropped) is removed, the Gram matrix of the rest R is formed (the GPU rounds
# R to BF16 for it), and its eigenvectors of the six largest eigenvalues above
# 1e-6 of the largest give the modes R v / sqrt(lambda) * sqrt(n_a).

def rigid_modes(dv):
    """Z_i (3 x 6) per node from the scaled offsets: translations, then
    rotations about x, y, z, as the coarse kernels write them."""
    n = dv.shape[0]
    Z = np.zeros((n, 3, 6))
    for k in range(3):
        Z[:, k, k] = 1.0
    dx, dy, dz = dv[:, 0], dv[:, 1], dv[:, 2]
    Z[:, 1, 3], Z[:, 2, 3] = -dz, dy
    Z[:, 0, 4], Z[:, 2, 4] = dz, -dx
    Z[:, 0, 5], Z[:, 1, 5] = -dy, dx
    return Z


def remove_rigid(S, agg, dv):
    """S (ns, N, 3) -> R: each sample minus its least-squares rigid-body part
    on every aggregate."""
    Z = rigid_modes(np.asarray(dv, np.float64))
    R = np.array(S, dtype=np.float64, copy=True)
    for a in np.unique(agg):
        nodes = np.nonzero(agg == a)[0]
        Za = Z[nodes].reshape(-1, 6)
        lam, V = np.linalg.eigh(Za.T @ Za)
        keep = lam > 1e-12 * lam.max()
        P = V[:, keep] @ np.diag(1.0 / lam[keep]) @ V[:, keep].T
        Sa = R[:, nodes, :].reshape(R.shape[0], -1).T  # (3 n_a, ns)
        c = P @ (Za.T @ Sa)
        R[:, nodes, :] = (Sa - Za @ c).T.reshape(R.shape[0], nodes.size, 3)
    return R


def fit_modes(R, agg, k=6, gram_bf16=False):
    """R (ns, N, 3) (the samples beyond the rigid-body modes) -> xm (N, k, 3)."""
    ns, N, _ = R.shape
    xm = np.zeros((N, k, 3))
    for a in np.unique(agg):
        nodes = np.nonzero(agg == a)[0]
        Ra = R[:, nodes, :].reshape(ns, -1).T  # (3 n_a, ns)
        Rg = bf16(Ra).astype(np.float64) if gram_bf16 else Ra
        lam, V = np.linalg.eigh(Rg.T @ Rg)
        order = np.argsort(lam)[::-1]
        lmax = lam[order[0]]
        for j in range(min(k, ns)):
            q = order[j]
            if not lmax > 0.0 or lam[q] <= 1e-6 * lmax:
                continue
            xm[nodes, j, :] = (Ra @ V[:, q] * (np.sqrt(nodes.size) / np.sqrt(lam[q]))).reshape(nodes.size, 3)
    return xm
