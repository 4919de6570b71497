"""Multipoint-constraint fixtures for the MPC-CG tests (hecMESH%mpc layout)."""
import numpy as np


def tie_constraints(n):
    """For the gen_hex8(n) cube ((n+1)^3 nodes, node i + (n+1)(j + (n+1)k)):
    every top-face node's z displacement tied to the top-face centre node's
    (u_z(s) - u_z(c) = 0), one node's x displacement tied to the mean of two
    others (u_x(s) - u_x(a)/2 - u_x(b)/2 = 0), and one constraint on a rotation
    dof (4), which the reference's MPC-CG skips.  Returns (index, item, dof, val)."""
    nn = n + 1
    nid = lambda i, j, k: i + nn * (j + nn * k) + 1
    top, c = n, n // 2
    centre = nid(c, c, top)
    index, item, dof, val = [0], [], [], []
    for j in range(nn):
        for i in range(nn):
            s = nid(i, j, top)
            if s == centre or (i, j) == (0, 0):
                continue
            item += [s, centre]
            dof += [3, 3]
            val += [1.0, -1.0]
            index.append(len(item))
    mid = nid(c, c, c)
    item += [mid, nid(c - 1, c, c), nid(c + 1, c, c)]
    dof += [1, 1, 1]
    val += [1.0, -0.5, -0.5]
    index.append(len(item))
    item += [nid(0, 0, top), nid(1, 0, top)]
    dof += [4, 4]
    val += [1.0, -1.0]
    index.append(len(item))
    return (np.array(index, np.int32), np.array(item, np.int32), np.array(dof, np.int32), np.array(val))


def dense_T(N, mpc):
    """T (3N x 3N) of hecmw_Tvec_33: identity except each slave dof's row, which is
    minus the master coefficients; constraints with a dof above 3 are skipped."""
    index, item, dof, val = mpc
    T = np.eye(3 * N)
    for i in range(len(index) - 1):
        terms = range(index[i], index[i + 1])
        if any(dof[j] > 3 for j in terms):
            continue
        ks = 3 * (item[index[i]] - 1) + dof[index[i]] - 1
        T[ks, :] = 0.0
        for j in list(terms)[1:]:
            T[ks, 3 * (item[j] - 1) + dof[j] - 1] -= val[j]
    return T
