"""The TET10 box generator (fistr1 element 342, Kuhn split) against what the
mechanics fixes: a symmetric positive definite operator, zero internal force
for rigid motions and for any linear (constant-strain) displacement field at
nodes off the boundary (the patch test, exact for quadratic tetrahedra with
the 4-point rule), and the hecmw_mat_con profile rules."""
import numpy as np
import pytest

import frontistr_amd as fa
from tests.systems import dense


def _interior_mask(nx, ny, nz):
    X, Y, Z = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    idx = np.arange(X * Y * Z)
    x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
    # away from every face by more than one cell (rows coupled to fixed or loaded nodes excluded)
    return (x > 2) & (x < X - 3) & (y > 2) & (y < Y - 3) & (z > 2) & (z < Z - 3), (x, y, z)


@pytest.mark.parametrize("shape,box", [((4, 4, 4), (1.0, 1.0, 1.0)), ((5, 4, 6), (2.0, 0.7, 1.3))])
def test_patch_and_rigid_modes(shape, box):
    nx, ny, nz = shape
    m, _ = fa.gen_tet10(nx, ny, nz, box=box, nrhs=1)
    A = dense(m)
    inner, (x, y, z) = _interior_mask(nx, ny, nz)
    assert inner.sum() > 0
    h = np.array(box) / np.array(shape) / 2.0
    P = np.stack([x * h[0], y * h[1], z * h[2]], 1)  # node coordinates
    scale = np.abs(A).max() * np.abs(P).max()
    rng = np.random.default_rng(0)
    G = rng.standard_normal((3, 3))
    fields = [np.tile([1.0, 0, 0], len(x)), np.tile([0, 0, 1.0], len(x)),
              np.cross([0.3, -0.2, 0.5], P).ravel(),       # rotation
              (P @ G.T).ravel()]                           # arbitrary linear field
    for u in fields:
        f = (A @ u).reshape(-1, 3)
        assert np.abs(f[inner]).max() <= 1e-9 * scale, np.abs(f[inner]).max() / scale
    # a quadratic field is not in the patch: interior forces appear (the test can fail)
    uq = np.stack([P[:, 0] ** 2, 0 * P[:, 0], 0 * P[:, 0]], 1).ravel()
    assert np.abs((A @ uq).reshape(-1, 3)[inner]).max() > 1e-6 * scale


def test_spd_symmetric_and_profile():
    m, B = fa.gen_tet10(3, 2, 2, box=(1.0, 0.7, 0.6), nrhs=4, seed=1)
    assert m.N == 7 * 5 * 5
    A = dense(m)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
    for v in range(m.N):
        L = m.itemL[m.indexL[v]:m.indexL[v + 1]]
        U = m.itemU[m.indexU[v]:m.indexU[v + 1]]
        assert np.all(L < v + 1) and np.all(U > v + 1)
        assert len(set(L)) == len(L) and len(set(U)) == len(U)
    # symmetric profile: w in row v iff v in row w
    pairs = set()
    for v in range(m.N):
        for w in m.itemL[m.indexL[v]:m.indexL[v + 1]]:
            pairs.add((v + 1, int(w)))
    for v in range(m.N):
        for w in m.itemU[m.indexU[v]:m.indexU[v + 1]]:
            assert (int(w), v + 1) in pairs
    # loads only on the top face, N(0,1), columns differ; fixed face rows zero
    X, Y = 7, 5
    top = np.arange(m.N) // (X * Y) == 4
    bot = np.arange(m.N) // (X * Y) == 0
    Bn = B.reshape(4, -1, 3)
    assert np.all(Bn[:, ~top] == 0.0)
    assert np.all(Bn[:, bot] == 0.0)
    assert not np.allclose(Bn[0], Bn[1])
    assert 0.5 < Bn[:, top].std() < 1.5
