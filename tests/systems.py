"""Small test systems in hecmw1's 3x3 block layout."""
import numpy as np

import frontistr_amd as fa


def random_system(N, deg=8, seed=0, shuffle=True, n_ext=0, components=1):
    """Random symmetric block pattern with an SPD, block diagonally dominant
    matrix.  Items are left in random order within each row (as
    hecmw_mat_con leaves them in element order).  n_ext extra external nodes
    (index > N) are coupled from the U part of internal rows, as in a
    partitioned domain.  components > 1 splits the nodes into disconnected
    groups."""
    rng = np.random.default_rng(seed)
    NP = N + n_ext
    edges = set()
    comp = np.arange(N) % components
    for i in range(N):
        for _ in range(deg // 2):
            j = int(rng.integers(0, N))
            if j != i and comp[j] == comp[i]:
                edges.add((min(i, j), max(i, j)))
    ext_edges = set()
    for e in range(n_ext):
        for _ in range(3):
            i = int(rng.integers(0, N))
            ext_edges.add((i, N + e))
    blocks = {}
    for (i, j) in list(edges) + list(ext_edges):
        blocks[(i, j)] = rng.standard_normal((3, 3)) * 0.5
    rows_L = [[] for _ in range(NP)]
    rows_U = [[] for _ in range(NP)]
    for (i, j), v in blocks.items():
        rows_U[i].append((j, v))
        rows_L[j].append((i, v.T.copy()))
    diag = np.zeros((NP, 3, 3))
    for i in range(NP):
        s = sum(np.abs(v).sum() for _, v in rows_L[i]) + sum(np.abs(v).sum() for _, v in rows_U[i])
        m = rng.standard_normal((3, 3)) * 0.1
        diag[i] = m @ m.T + (s + 1.0) * np.eye(3)
    iL = np.zeros(NP + 1, np.int32)
    iU = np.zeros(NP + 1, np.int32)
    tL, tU, AL, AU = [], [], [], []
    for i in range(NP):
        l = list(rows_L[i])
        u = list(rows_U[i])
        if shuffle:
            rng.shuffle(l)
            rng.shuffle(u)
        else:
            l.sort(key=lambda t: t[0])
            u.sort(key=lambda t: t[0])
        for j, v in l:
            tL.append(j + 1)
            AL.append(v.reshape(9))
        for j, v in u:
            tU.append(j + 1)
            AU.append(v.reshape(9))
        iL[i + 1] = len(tL)
        iU[i + 1] = len(tU)
    AL = np.array(AL).reshape(-1) if AL else np.zeros(0)
    AU = np.array(AU).reshape(-1) if AU else np.zeros(0)
    B = rng.standard_normal(3 * NP)
    m = fa.HecMatrix(N, diag.reshape(-1), AL, AU, iL, np.array(tL, np.int32), iU, np.array(tU, np.int32), B=B,
                     NP=NP)
    return m


def dense(m):
    """Dense 3NP x 3NP matrix of the rows 1..N (external rows left zero)."""
    A = np.zeros((3 * m.NP, 3 * m.NP))
    for i in range(m.N):
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] = m.D[9 * i:9 * i + 9].reshape(3, 3)
        for q in range(m.indexL[i], m.indexL[i + 1]):
            j = m.itemL[q] - 1
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = m.AL[9 * q:9 * q + 9].reshape(3, 3)
        for q in range(m.indexU[i], m.indexU[i + 1]):
            j = m.itemU[q] - 1
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = m.AU[9 * q:9 * q + 9].reshape(3, 3)
    return A


def shuffle_items(m, seed=0):
    """Same matrix, random item order inside every row."""
    rng = np.random.default_rng(seed)
    m2 = m.copy()
    for i in range(m.NP):
        for idx, item, A in ((m2.indexL, m2.itemL, m2.AL), (m2.indexU, m2.itemU, m2.AU)):
            a, b = idx[i], idx[i + 1]
            if b - a < 2:
                continue
            p = rng.permutation(b - a)
            item[a:b] = item[a:b][p]
            blk = A[9 * a:9 * b].reshape(-1, 9)[p].reshape(-1)
            A[9 * a:9 * b] = blk
    return m2


def to_csr(m):
    """A (3N x 3NP) as scipy CSR, the rows of the internal nodes."""
    import scipy.sparse as sp
    D, AL, AU = m.D.reshape(-1, 3, 3), m.AL.reshape(-1, 3, 3), m.AU.reshape(-1, 3, 3)
    idx, blk, ptr = [], [], [0]
    for i in range(m.N):
        l0, l1, u0, u1 = m.indexL[i], m.indexL[i + 1], m.indexU[i], m.indexU[i + 1]
        idx += list(m.itemL[l0:l1] - 1) + [i] + list(m.itemU[u0:u1] - 1)
        blk += [AL[l0:l1], D[i:i + 1], AU[u0:u1]]
        ptr.append(len(idx))
    return sp.bsr_matrix((np.concatenate(blk), np.array(idx), np.array(ptr)), shape=(3 * m.N, 3 * m.NP)).tocsr()
