"""hecmw_mat_dump's on-disk formats (hecmw1/src/solver/matrix/hecmw_matrix_dump.f90).

fistr1 dumps the assembled system with ``!SOLVER`` ``DUMPTYPE=MM|CSR|BSR``
(Iarray 31, ``hecmw_matrix_misc.F90:123``): ``hecmw_mat_dump`` writes
``dump_matrix_<call>_<rank>.{mm,csr,bsr}`` plus ``.rhs``, and
``hecmw_mat_dump_solution`` writes ``.sol``.  This module reads such files into
a :class:`HecMatrix` (so a system dumped by a real fistr1 run can be solved on
the GPU as it is) and writes a :class:`HecMatrix` back in the same formats,
byte for byte as the reference writes them:

* values as Fortran ``e20.12e3`` (12 significant digits: a value read back is
  the 12-digit rounding of the one written), integers as ``I0``, one item per
  line;
* rows are all NP nodes (the reference dumps ``n = NP``), each row's entries
  in the order L (storage order), D, U (storage order)
  (``hecmw_matrix_dump.f90:85-119,181-220,252-297``);
* BSR keeps the 3x3 blocks row-major as ``hecmwST_matrix`` stores them; CSR and
  MM expand every block row into its 3 scalar rows (1-based scalar indices).

Reading reverses that exactly: the diagonal entry of each block row splits its
items into the ``itemL`` part (before) and the ``itemU`` part (after), so the
profile, the item order within each row and every block come back as dumped.
A dump does not record N separately from NP, so the matrix read has N = NP.
"""
import math
import os

import numpy as np

from .hecmw import HecMatrix

NDOF = 3
KINDS = {"mm": 1, "csr": 2, "bsr": 3}


def _fe(v):
    """Fortran e20.12e3: sign column, 0.<12 digits>E<sign><3-digit exponent>."""
    v = float(v)
    if math.isnan(v):
        return "%20s" % "NaN"
    if math.isinf(v):
        return "%20s" % ("-Infinity" if v < 0 else "Infinity")
    sign = "-" if math.copysign(1.0, v) < 0 else " "
    if v == 0.0:
        return sign + "0.000000000000E+000"
    mant, exp = ("%.11e" % abs(v)).split("e")
    e = int(exp) + 1
    return "%s0.%sE%s%03d" % (sign, mant.replace(".", ""), "+" if e >= 0 else "-", abs(e))


def _lines_f(a):
    return "".join(_fe(v) + "\n" for v in np.asarray(a, dtype=np.float64).ravel())


def _lines_i(a):
    return "".join("%d\n" % v for v in np.asarray(a).ravel())


def _rows(m):
    """Per block row i (0-based, all NP rows): node ids (1-based) and blocks (k, 3, 3)
    in the dump's L, D, U order."""
    NP = m.NP
    D = m.D.reshape(NP, 3, 3)
    AL = m.AL.reshape(-1, 3, 3)
    AU = m.AU.reshape(-1, 3, 3)
    for i in range(NP):
        l0, l1 = m.indexL[i], m.indexL[i + 1]
        u0, u1 = m.indexU[i], m.indexU[i + 1]
        items = np.concatenate([m.itemL[l0:l1], [i + 1], m.itemU[u0:u1]]).astype(np.int64)
        blocks = np.concatenate([AL[l0:l1], D[i:i + 1], AU[u0:u1]])
        yield items, blocks


def write_dump(m, kind, path):
    """Write m as hecmw_mat_dump would with DUMPTYPE=kind ("mm", "csr" or "bsr") to
    `path` (the matrix file); B goes to the same stem + ".rhs" and X to ".sol"."""
    if kind not in KINDS:
        raise ValueError("kind must be one of %s" % sorted(KINDS))
    NP = m.NP
    nnzb = NP + int(m.indexL[NP]) + int(m.indexU[NP])
    out = []
    if kind == "bsr":
        out.append("%%Block-CSR matrix real general\n% nrow ncol nnonzero ndof\n")
        out.append("%d %d %d %d\n" % (NP, NP, nnzb, NDOF))
        out.append("% index(0:nrow)\n")
        cnt = np.diff(m.indexL[:NP + 1]) + 1 + np.diff(m.indexU[:NP + 1])
        out.append(_lines_i(np.concatenate([[0], np.cumsum(cnt)])))
        rows = list(_rows(m))
        out.append("% item(1:nnonzero)\n")
        out.append(_lines_i(np.concatenate([r[0] for r in rows])))
        out.append("% value(1:nnonzero*ndof*ndof)\n")
        out.append(_lines_f(np.concatenate([r[1].ravel() for r in rows])))
    else:
        rows = list(_rows(m))
        nnz = NDOF * NDOF * nnzb
        if kind == "csr":
            out.append("%%CSR matrix real general\n% nrow ncol nnonzero\n")
            out.append("%d %d %d\n" % (NP * NDOF, NP * NDOF, nnz))
            out.append("% index(0:nrow)\n")
            per = np.repeat([NDOF * len(r[0]) for r in rows], NDOF)
            out.append(_lines_i(np.concatenate([[0], np.cumsum(per)])))
            out.append("% item(1:nnonzero)\n")
            cols, vals = [], []
            for items, blocks in rows:
                c = ((items[:, None] - 1) * NDOF + np.arange(1, NDOF + 1)[None, :]).ravel()
                for d in range(NDOF):
                    cols.append(c)
                    vals.append(blocks[:, d, :].ravel())
            out.append(_lines_i(np.concatenate(cols)))
            out.append("% value(1:nnonzero)\n")
            out.append(_lines_f(np.concatenate(vals)))
        else:
            out.append("%%MatrixMarket matrix coordinate real general\n")
            out.append("%d %d %d\n" % (NP * NDOF, NP * NDOF, nnz))
            for i, (items, blocks) in enumerate(rows):
                c = ((items[:, None] - 1) * NDOF + np.arange(1, NDOF + 1)[None, :]).ravel()
                for d in range(NDOF):
                    ii = NDOF * i + d + 1
                    v = blocks[:, d, :].ravel()
                    out.append("".join("%d %d %s\n" % (ii, jj, _fe(x)) for jj, x in zip(c, v)))
    with open(path, "w") as f:
        f.write("".join(out))
    stem = os.path.splitext(path)[0]
    with open(stem + ".rhs", "w") as f:
        f.write(_lines_f(m.B[:NDOF * NP]))
    with open(stem + ".sol", "w") as f:
        f.write(_lines_f(m.X[:NDOF * NP]))


def _split_rows(NP, ptr, items, blocks):
    """Block rows (items 1-based, blocks (k,3,3)) -> HecMatrix arrays."""
    iL = np.zeros(NP + 1, np.int32)
    iU = np.zeros(NP + 1, np.int32)
    tL, tU, AL, AU = [], [], [], []
    D = np.zeros((NP, 3, 3))
    for i in range(NP):
        it = items[ptr[i]:ptr[i + 1]]
        bl = blocks[ptr[i]:ptr[i + 1]]
        k = np.nonzero(it == i + 1)[0]
        if k.size != 1:
            raise ValueError("block row %d has %d diagonal entries" % (i + 1, k.size))
        k = int(k[0])
        D[i] = bl[k]
        tL.append(it[:k])
        AL.append(bl[:k])
        tU.append(it[k + 1:])
        AU.append(bl[k + 1:])
        iL[i + 1] = iL[i] + k
        iU[i + 1] = iU[i] + (len(it) - k - 1)
    cat = lambda a, shape: np.concatenate(a).reshape(shape) if a else np.zeros(shape)
    return (D.ravel(), cat(AL, (-1,)), cat(AU, (-1,)), iL, cat(tL, (-1,)).astype(np.int32), iU,
            cat(tU, (-1,)).astype(np.int32))


def _numbers(lines, conv):
    return np.array([conv(t) for t in lines], dtype=np.float64 if conv is float else np.int64)


def read_dump(path):
    """Read a hecmw_mat_dump matrix file (.mm, .csr or .bsr, recognised by its
    header) and, when present beside it, the .rhs (B) and .sol (X) files."""
    with open(path) as f:
        text = f.read().splitlines()
    head = text[0]
    body = [l for l in text[1:] if not l.startswith("%")]
    if head.startswith("%%Block-CSR"):
        nrow, ncol, nnz, ndof = (int(t) for t in body[0].split())
        if ndof != NDOF:
            raise ValueError("ndof %d: only 3x3 blocks are supported" % ndof)
        NP = nrow
        ptr = _numbers(body[1:NP + 2], int)
        items = _numbers(body[NP + 2:NP + 2 + nnz], int)
        vals = _numbers(body[NP + 2 + nnz:NP + 2 + nnz + 9 * nnz], float).reshape(nnz, 3, 3)
    elif head.startswith("%%CSR") or head.startswith("%%MatrixMarket"):
        nrow, ncol, nnz = (int(t) for t in body[0].split())
        if nrow % NDOF:
            raise ValueError("%d scalar rows: not a 3-DOF system" % nrow)
        NP = nrow // NDOF
        if head.startswith("%%CSR"):
            sptr = _numbers(body[1:nrow + 2], int)
            scol = _numbers(body[nrow + 2:nrow + 2 + nnz], int)
            sval = _numbers(body[nrow + 2 + nnz:nrow + 2 + 2 * nnz], float)
        else:
            trip = [l.split() for l in body[1:1 + nnz]]
            srow = np.array([int(t[0]) for t in trip])
            scol = np.array([int(t[1]) for t in trip])
            sval = np.array([float(t[2]) for t in trip])
            sptr = np.concatenate([[0], np.cumsum(np.bincount(srow - 1, minlength=nrow))])
        # scalar rows 3i+d hold the same node sequence (3 columns per block)
        ptr = np.zeros(NP + 1, np.int64)
        items_l, blocks_l = [], []
        for i in range(NP):
            r0 = [scol[sptr[NDOF * i + d]:sptr[NDOF * i + d + 1]] for d in range(NDOF)]
            v0 = [sval[sptr[NDOF * i + d]:sptr[NDOF * i + d + 1]] for d in range(NDOF)]
            nodes = (r0[0][::NDOF] - 1) // NDOF + 1
            for d in range(NDOF):
                if not np.array_equal(r0[d], r0[0]):
                    raise ValueError("scalar rows of node %d differ in their columns" % (i + 1))
            items_l.append(nodes)
            blocks_l.append(np.stack([v0[d].reshape(-1, NDOF) for d in range(NDOF)], axis=1))
            ptr[i + 1] = ptr[i] + len(nodes)
        items = np.concatenate(items_l)
        vals = np.concatenate(blocks_l)
    else:
        raise ValueError("%s: not a hecmw_mat_dump matrix file" % path)
    D, AL, AU, iL, tL, iU, tU = _split_rows(NP, ptr, items, vals)
    stem = os.path.splitext(path)[0]
    B = X = None
    if os.path.exists(stem + ".rhs"):
        with open(stem + ".rhs") as f:
            B = _numbers(f.read().split(), float)
    if os.path.exists(stem + ".sol"):
        with open(stem + ".sol") as f:
            X = _numbers(f.read().split(), float)
    return HecMatrix(NP, D, AL, AU, iL, tL, iU, tU, B=B, X=X)
