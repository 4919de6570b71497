"""fistr1 input models in hecmw's text formats, for the linear static systems
the solver receives.

The subset read is what fistr1's own solver tests use
(tests/solver/CG/*.msh, *.cnt in the reference):

* mesh (``.msh``, hecmw's native format): ``!NODE``, ``!ELEMENT, TYPE=341``,
  ``342`` or ``361`` (4- and 10-node tetrahedra, 8-node hexahedron),
  ``!MATERIAL`` with ``!ITEM=1, SUBITEM=2`` (Young's
  modulus, Poisson's ratio) and ``!ITEM=2`` (density), ``!NGROUP`` (list or
  ``GENERATE`` start, end, step); other blocks are skipped;
* control (``.cnt``): ``!BOUNDARY`` (node or group, first dof, last dof,
  value), ``!CLOAD`` (node or group, dof, value), ``!DLOAD`` body forces on
  all elements (``BX``/``BY``/``BZ`` force per volume, ``GRAV`` acceleration
  times the density along the normalised direction, ``CENT`` angular speed,
  axis point and axis direction) and pressure ``P1``..``P6`` on a face of the
  elements of an ``!EGROUP`` (types 341 and 361), as ``DL_C3``,
  fistr1/src/lib/static_LIB_3d.f90:213-382, applies them) and ``!SOLVER``
  (METHOD, PRECOND and the
  two data lines: ITER, ITERPREMAX / RESID, SIGMA_DIAG, SIGMA);
* result (``.res``, ``*fstrresult 2.0``): the nodal DISPLACEMENT of every node.

:func:`assemble` builds the static system fistr1 hands to ``hecmw_solve``:

* element stiffness as ``STF_C3`` computes it
  (fistr1/src/lib/static_LIB_3d.f90:50-208): small strain, isotropic
  elasticity, K_e = sum over quadrature points of B^T D B det(J) w.  Type
  341 has one point, so K_e = V B^T D B with the constant B of the linear
  tetrahedron; type 342 uses the quadratic shape functions of
  fistr1/src/lib/element/tet10n.f90 (vertices 1-4, then the midpoints of
  edges 1-2, 2-3, 3-1, 1-4, 2-4, 3-4) and the 4-point rule (a = (5 - sqrt 5)
  / 20, b = (5 + 3 sqrt 5) / 20, w = 1/24).  A ``.msh`` file lists a 342's
  midpoints in hecmw's order, edges 2-3, 3-1, 1-2, 1-4, 2-4, 3-4 (the order
  hecmw1/src/common/hecmw_conn_conv.c:18 converts Abaqus' C3D10 into), so
  nodes 5-7 are rotated onto the shape functions' edges.  Type 361 is the
  formulation fistr1's static analysis picks by default (incompatible modes,
  fistr1/src/common/fstr_setup.f90:415-424; ``STF_C3D8IC``,
  fistr1/src/lib/static_LIB_3dIC.f90:21-215): trilinear shape functions
  (fistr1/src/lib/element/hex8n.f90), 2x2x2 Gauss points, three bubble modes
  1 - xi_i^2 whose gradients -2 xi_i grad(xi_i) use the Jacobian at the
  element centre scaled by det(J_centre) / det(J), and the 9 bubble dofs
  condensed out: K = K_uu - K_ua K_aa^-1 K_au;
* the nodes are those the elements use, in the order of ``!NODE`` (the mesh
  reader drops nodes no element references), the block profile is every
  node pair sharing an element with items ascending per row
  (``hecmw_mat_con``, hecmw1/src/solver/matrix/hecmw_mat_con.f90:21-107),
  assembled sparse (element matrices summed in a scipy COO, then 3x3 blocks)
  so models of any size load;
* Dirichlet values by row and column elimination with a unit pivot
  (``hecmw_mat_ass_bc``, hecmw1/src/solver/matrix/hecmw_mat_ass.f90:328-463),
  the right-hand side moved by the prescribed values;
* concentrated loads added to the right-hand side, and body forces
  integrated against each shape function with the element's own quadrature
  as ``DL_C3`` does (1 point for type 341, 4 for 342, 8 for 361; exact for a
  constant force, so V f / 4 per node of a linear tetrahedron).
"""
import re

import numpy as np
import scipy.sparse as sp

from .hecmw import HecMatrix


def _blocks(path):
    """(header keyword line, [data lines]) in file order; '!!' and '#' lines are comments."""
    out = []
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("!!") or line.startswith("#"):
                continue
            if line.startswith("!"):
                out.append((line, []))
            elif out:
                out[-1][1].append(line)
    return out


def _opts(header):
    """'!KEY, A=1, B' -> ('KEY', {'A': '1', 'B': ''})"""
    parts = [p.strip() for p in header[1:].split(",")]
    d = {}
    for p in parts[1:]:
        k, _, v = p.partition("=")
        d[k.strip().upper()] = v.strip()
    return parts[0].upper(), d


def _fields(line):
    return [t for t in re.split(r"[,\s]+", line.strip()) if t]


def read_msh(path):
    """Nodes {id: (x, y, z)} in file order, elements [(id, type, node ids)],
    material (E, nu, density or None), node groups and element groups
    ({name: [ids]} each)."""
    nodes, elems, groups, egroups = {}, [], {}, {}
    material = None
    density = None
    n_material = n_section = 0
    for header, lines in _blocks(path):
        key, opt = _opts(header)
        if key == "END":
            break
        # one material on every element: a second !MATERIAL or !SECTION would
        # give some elements other properties, which this reader does not map
        if key == "MATERIAL":
            n_material += 1
            if n_material > 1:
                raise ValueError("%s: more than one !MATERIAL; only single-material models are read" % path)
        elif key == "SECTION":
            n_section += 1
            if n_section > 1:
                raise ValueError("%s: more than one !SECTION; only single-section models are read" % path)
            if opt.get("TYPE", "SOLID").upper() != "SOLID":
                raise ValueError("%s: !SECTION TYPE=%s is not read" % (path, opt.get("TYPE")))
        if key == "NODE":
            for ln in lines:
                f = _fields(ln)
                nodes[int(f[0])] = tuple(float(v) for v in f[1:4])
        elif key == "ELEMENT":
            etype = int(opt.get("TYPE", "0"))
            for ln in lines:
                f = [int(v) for v in _fields(ln)]
                elems.append((f[0], etype, f[1:]))
        elif key in ("NGROUP", "EGROUP"):
            ids = []
            for ln in lines:
                f = [int(v) for v in _fields(ln)]
                ids += list(range(f[0], f[1] + 1, f[2] if len(f) > 2 else 1)) if "GENERATE" in opt else f
            name = opt["NGRP" if key == "NGROUP" else "EGRP"].upper()
            (groups if key == "NGROUP" else egroups).setdefault(name, []).extend(ids)
        elif key.replace(" ", "") == "ITEM=1":
            # '!ITEM=1, SUBITEM=2': Young's modulus and Poisson's ratio
            if opt.get("SUBITEM", "2") != "2":
                raise ValueError("%s: only isotropic elasticity (ITEM=1, SUBITEM=2) is read" % path)
            f = [float(v) for v in _fields(lines[0])]
            material = (f[0], f[1])
        elif key.replace(" ", "") == "ITEM=2":
            density = float(_fields(lines[0])[0])
    if material is not None:
        material = (material[0], material[1], density)
    return nodes, elems, material, groups, egroups


def read_cnt(path):
    """Boundary conditions [(node or group, first dof, last dof, value)], loads
    [(node or group, dof, value)], body loads [(element group, kind, values)]
    and solver settings."""
    bcs, loads, dloads, solver = [], [], [], {}
    for header, lines in _blocks(path):
        key, opt = _opts(header)
        if key == "END":
            break
        if key == "BOUNDARY":
            for ln in lines:
                f = _fields(ln)
                bcs.append((f[0], int(f[1]), int(f[2]), float(f[3]) if len(f) > 3 else 0.0))
        elif key == "CLOAD":
            for ln in lines:
                f = _fields(ln)
                loads.append((f[0], int(f[1]), float(f[2])))
        elif key == "DLOAD":
            for ln in lines:
                f = _fields(ln)
                kind = f[1].upper()
                if kind not in ("BX", "BY", "BZ", "GRAV", "CENT") and kind not in ("P1", "P2", "P3", "P4", "P5", "P6"):
                    raise ValueError("%s: !DLOAD %s is not read" % (path, kind))
                dloads.append((f[0].upper(), kind, [float(v) for v in f[2:]]))
        elif key == "SOLVER":
            # header parameters and data lines as fstr_ctrl_get_SOLVER reads them
            # (fistr1/src/common/fstr_ctrl_common.f90:138-201)
            solver["METHOD"] = opt.get("METHOD", "CG").upper()
            if "METHOD2" in opt:
                solver["METHOD2"] = opt["METHOD2"].upper()
            solver["PRECOND"] = int(opt.get("PRECOND", "1"))
            for name in ("SCALING", "ITERLOG", "TIMELOG"):
                if name in opt:
                    words = ("NO", "YES", "VERBOSE") if name == "TIMELOG" else ("NO", "YES")
                    v = opt[name].upper()
                    if v not in words:
                        raise ValueError("%s: !SOLVER %s=%s" % (path, name, opt[name]))
                    solver[name] = words.index(v)
            if len(lines) > 0:
                # nier, iterpremax, nrest, ncolor_in, recyclepre (the 'iiiiii' line)
                f = [int(v) for v in _fields(lines[0])]
                for name, v in zip(("ITER", "ITERPREMAX", "NREST", "NCOLOR_IN", "RECYCLEPRE"), f):
                    solver[name] = v
            if len(lines) > 1:
                f = [float(v) for v in _fields(lines[1])]
                solver["RESID"] = f[0]
                if len(f) > 1:
                    solver["SIGMA_DIAG"] = f[1]
                if len(f) > 2:
                    solver["SIGMA"] = f[2]
    return bcs, loads, dloads, solver


def read_res_displacement(path):
    """{node id: (ux, uy, uz)} from an ``*fstrresult 2.0`` file whose first
    nodal field is DISPLACEMENT (3 components)."""
    with open(path) as f:
        tok = f.read().split()
    k = tok.index("*data")
    nn, _ne = int(tok[k + 1]), int(tok[k + 2])
    nfield = int(tok[k + 3])
    p = k + 5
    sizes = [int(tok[p + i]) for i in range(nfield)]
    p += nfield
    names = tok[p:p + nfield]
    if names[0] != "DISPLACEMENT" or sizes[0] != 3:
        raise ValueError("%s: the first nodal field is not a 3-component DISPLACEMENT" % path)
    p += nfield
    width = sum(sizes)
    out = {}
    for _ in range(nn):
        nid = int(tok[p])
        vals = [float(v) for v in tok[p + 1:p + 1 + width]]
        out[nid] = tuple(vals[:3])
        p += 1 + width
    return out


def _D_iso(E, nu):
    """Isotropic elasticity in (xx, yy, zz, xy, yz, zx) with engineering shears."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2.0 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return D


def _B(dN):
    """6 x 3n strain-displacement matrix from the shape-function gradients (n x 3)."""
    n = dN.shape[0]
    B = np.zeros((6, 3 * n))
    for a in range(n):
        gx, gy, gz = dN[a]
        c = 3 * a
        B[0, c] = gx
        B[1, c + 1] = gy
        B[2, c + 2] = gz
        B[3, c], B[3, c + 1] = gy, gx
        B[4, c + 1], B[4, c + 2] = gz, gy
        B[5, c], B[5, c + 2] = gz, gx
    return B


_HEC342_TO_SHAPE = (0, 1, 2, 3, 6, 4, 5, 7, 8, 9)
_QA, _QB = (5.0 - np.sqrt(5.0)) / 20.0, (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
_TET10_POINTS = ((_QA, _QA, _QA), (_QB, _QA, _QA), (_QA, _QB, _QA), (_QA, _QA, _QB))


def _tet10_shape(xi, et, ze):
    a = 1.0 - xi - et - ze
    N = np.array([a * (2 * a - 1), xi * (2 * xi - 1), et * (2 * et - 1), ze * (2 * ze - 1),
                  4 * a * xi, 4 * xi * et, 4 * et * a, 4 * a * ze, 4 * xi * ze, 4 * et * ze])
    g = np.array([[1 - 4 * a, 1 - 4 * a, 1 - 4 * a],
                  [4 * xi - 1, 0, 0],
                  [0, 4 * et - 1, 0],
                  [0, 0, 4 * ze - 1],
                  [4 * (a - xi), -4 * xi, -4 * xi],
                  [4 * et, 4 * xi, 0],
                  [-4 * et, 4 * (a - et), -4 * et],
                  [-4 * ze, -4 * ze, 4 * (a - ze)],
                  [4 * ze, 0, 4 * xi],
                  [0, 4 * ze, 4 * et]])
    return N, g


def _tet10_points(X):
    """(N, dN/dx, det J * w) at the 4 quadrature points of a 10-node tetrahedron."""
    out = []
    for xi, et, ze in _TET10_POINTS:
        N, g = _tet10_shape(xi, et, ze)
        J = g.T @ X  # J[r, d] = d x_d / d xi_r
        det = np.linalg.det(J)
        if det <= 0.0:
            raise ValueError("tetrahedron with non-positive Jacobian")
        out.append((N, g @ np.linalg.inv(J).T, det / 24.0))
    return out


def _tet10_stiffness(X, D):
    """30x30 stiffness of a 10-node tetrahedron with node coordinates X (10x3)."""
    K = np.zeros((30, 30))
    for _, dN, w in _tet10_points(X):
        B = _B(dN)
        K += w * (B.T @ D @ B)
    return K


_HEX8_SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                        [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
_G2 = 1.0 / np.sqrt(3.0)
_HEX8_POINTS = tuple((a, b, c) for c in (-_G2, _G2) for b in (-_G2, _G2) for a in (-_G2, _G2))


def _hex8_shape(xi):
    f = 1.0 + _HEX8_SIGNS * np.asarray(xi)[None, :]  # (8, 3)
    N = 0.125 * f.prod(axis=1)
    g = np.empty((8, 3))
    for r in range(3):
        o = [d for d in range(3) if d != r]
        g[:, r] = 0.125 * _HEX8_SIGNS[:, r] * f[:, o[0]] * f[:, o[1]]
    return N, g


def _hex8_points(X):
    """(N, dN/dx, det J * w, det J, xi) at the 8 Gauss points of a hexahedron."""
    out = []
    for xi in _HEX8_POINTS:
        N, g = _hex8_shape(xi)
        J = g.T @ X
        det = np.linalg.det(J)
        if det <= 0.0:
            raise ValueError("hexahedron with non-positive Jacobian")
        out.append((N, g @ np.linalg.inv(J).T, det, det, xi))
    return out


def _hex8ic_stiffness(X, D):
    """24x24 stiffness of an incompatible-mode hexahedron (9 bubble dofs condensed)."""
    _, g0 = _hex8_shape((0.0, 0.0, 0.0))
    J0 = g0.T @ X
    det0 = np.linalg.det(J0)
    grad_xi = np.linalg.inv(J0)  # column i: the gradient of xi_i in x
    K = np.zeros((33, 33))
    for _, dN, w, det, xi in _hex8_points(X):
        bub = np.array([-2.0 * xi[i] * grad_xi[:, i] * det0 / det for i in range(3)])
        B = _B(np.vstack([dN, bub]))
        K += w * (B.T @ D @ B)
    Kaa = K[24:, 24:]
    return K[:24, :24] - K[:24, 24:] @ np.linalg.solve(Kaa, K[24:, :24])


# face k of an element: its nodes in the order whose right-hand rule gives the
# normal (fistr1/src/lib/element/element.f90 getSubFace)
_FACES = {
    341: ((1, 2, 3), (4, 2, 1), (4, 3, 2), (4, 1, 3)),
    361: ((1, 2, 3, 4), (8, 7, 6, 5), (5, 6, 2, 1), (6, 7, 3, 2), (7, 8, 4, 3), (8, 5, 1, 4)),
}


def _face_points(X):
    """(shape values, dX/dxi x dX/deta * weight) over a 3- or 4-node face (SurfaceNormal
    in element.f90: the cross product of the two tangents, not normalised)."""
    if X.shape[0] == 3:
        # linear triangle: constant tangents, one point, weight 1/2
        nrm = np.cross(X[1] - X[0], X[2] - X[0])
        return [(np.full(3, 1.0 / 3.0), 0.5 * nrm)]
    out = []
    s = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=np.float64)
    for b in (-_G2, _G2):
        for a in (-_G2, _G2):
            shp = 0.25 * (1 + s[:, 0] * a) * (1 + s[:, 1] * b)
            dxi = 0.25 * s[:, 0] * (1 + s[:, 1] * b)
            deta = 0.25 * s[:, 1] * (1 + s[:, 0] * a)
            out.append((shp, np.cross(dxi @ X, deta @ X)))
    return out


def _quadrature(X):
    """(shape values, det J * weight) at the quadrature points fistr1 uses for the
    element with nodes X: 1 (the centroid) for 341, 4 for 342, 8 for 361."""
    n = X.shape[0]
    if n == 4:
        V = np.linalg.det((X[1:] - X[0]).T) / 6.0
        return [(np.full(4, 0.25), V)]
    if n == 10:
        return [(shp, w) for shp, _, w in _tet10_points(X)]
    return [(shp, w) for shp, _, w, _, _ in _hex8_points(X)]


def _tet4_stiffness(X, D):
    """12x12 stiffness of a linear tetrahedron with corner coordinates X (4x3)."""
    J = (X[1:] - X[0]).T  # columns: edges from node 1
    det = np.linalg.det(J)
    if det <= 0.0:
        raise ValueError("tetrahedron with non-positive volume")
    G = np.linalg.inv(J)  # rows: gradients of the barycentric coordinates of nodes 2-4
    B = _B(np.vstack([-G.sum(axis=0), G]))
    return (det / 6.0) * (B.T @ D @ B)


# !SOLVER METHOD names and numbers (fstr_ctrl_common.f90:119-171: names after
# the five numbers, CR and the GMRES variants counted among the iterative ones)
_METHODS = {"1": 1, "2": 2, "3": 3, "4": 4, "5": 5, "CG": 1, "BICGSTAB": 2, "GMRES": 3, "GPBICG": 4,
            "GMRESR": 5, "GMRESREN": 6, "CR": 7}


def _method_code(name, path, library_only=True):
    code = _METHODS.get(name.upper())
    if code is None:
        raise ValueError("%s: !SOLVER METHOD=%s is not an iterative method this reader knows" % (path, name))
    if library_only and code not in (1, 2):
        raise ValueError("%s: !SOLVER METHOD=%s: the GPU library provides CG (1) and BiCGSTAB (2) only"
                         % (path, name))
    return code


def assemble(msh_path, cnt_path):
    """The static system of a fistr1 model as a :class:`HecMatrix` (solver
    settings of the ``!SOLVER`` block applied) and the node ids of its rows."""
    nodes, elems, material, groups, egroups = read_msh(msh_path)
    bcs, loads, dloads, solver = read_cnt(cnt_path)
    if material is None:
        raise ValueError("%s: no !MATERIAL elasticity" % msh_path)
    if any(t not in (341, 342, 361) for _, t, _ in elems):
        raise ValueError("only element types 341, 342 and 361 are assembled")
    # 342: hecmw's midpoint order (2-3, 3-1, 1-2, ...) -> the shape functions' (1-2, 2-3, 3-1, ...)
    elems = [(e, t, [ns[i] for i in _HEC342_TO_SHAPE] if t == 342 else ns) for e, t, ns in elems]
    used = set(n for _, _, ns in elems for n in ns)
    ids = [n for n in nodes if n in used]
    idx = {n: i for i, n in enumerate(ids)}
    N = len(ids)
    X = np.array([nodes[n] for n in ids])
    D = _D_iso(material[0], material[1])
    # element matrices summed into a sparse K; every entry of an element's
    # node pairs stays stored (the profile of hecmw_mat_con), zero or not
    rows, cols, vals = [], [], []
    for _, _, ns in elems:
        loc = [idx[n] for n in ns]
        ke = {4: _tet4_stiffness, 10: _tet10_stiffness, 8: _hex8ic_stiffness}[len(loc)](X[loc], D)
        dofs = np.array([[3 * i, 3 * i + 1, 3 * i + 2] for i in loc]).ravel()
        rows.append(np.repeat(dofs, dofs.size))
        cols.append(np.tile(dofs, dofs.size))
        vals.append(ke.ravel())
    K = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(3 * N, 3 * N)).tocsr()
    K.sum_duplicates()
    K.sort_indices()
    B = np.zeros(3 * N)

    def members(target):
        t = target.upper()
        return [idx[n] for n in groups[t] if n in idx] if t in groups else [idx[int(target)]]

    for target, dof, val in loads:
        for i in members(target):
            B[3 * i + dof - 1] += val
    for egrp, kind, vals in dloads:
        if kind.startswith("P"):
            face = int(kind[1:])
            members_e = set(egroups[egrp]) if egrp in egroups else None
            if members_e is None and egrp != "ALL":
                raise ValueError("!DLOAD on unknown element group %s" % egrp)
            for eid, t, ns in elems:
                if members_e is not None and eid not in members_e:
                    continue
                if t not in _FACES:
                    raise ValueError("pressure on element type %d is not read" % t)
                loc = [idx[ns[k - 1]] for k in _FACES[t][face - 1]]
                for shp, nrm in _face_points(X[loc]):
                    for a, i in enumerate(loc):
                        B[3 * i:3 * i + 3] += vals[0] * shp[a] * nrm
            continue
        if egrp != "ALL":
            raise ValueError("!DLOAD on element group %s: only ALL is read" % egrp)
        if kind in ("GRAV", "CENT") and material[2] is None:
            raise ValueError("%s: %s needs the density (!ITEM=2)" % (msh_path, kind))
        if kind == "GRAV":
            d = np.array(vals[1:4])
            f = material[2] * vals[0] * d / np.linalg.norm(d)  # the direction is normalised (DL_C3)
        elif kind == "CENT":
            # rho w|w| (x - its projection on the axis through A along R), at each quadrature point
            omega, A_, R_ = vals[0], np.array(vals[1:4]), np.array(vals[4:7])
            coef = material[2] * omega * abs(omega)
        else:
            f = np.zeros(3)
            f["XYZ".index(kind[1])] = vals[0]
        for _, _, ns in elems:
            loc = [idx[n] for n in ns]
            for shp, w in _quadrature(X[loc]):
                if kind == "CENT":
                    xg = shp @ X[loc]
                    f = coef * (xg - (A_ + ((xg - A_) @ R_) / (R_ @ R_) * R_))
                for a, i in enumerate(loc):
                    B[3 * i:3 * i + 3] += (w * shp[a]) * f
    fixed = np.zeros(3 * N, dtype=bool)
    xfix = np.zeros(3 * N)
    for target, d0, d1, val in bcs:
        for i in members(target):
            for d in range(d0, d1 + 1):
                fixed[3 * i + d - 1] = True
                xfix[3 * i + d - 1] = val
    # elimination with a unit pivot: move the prescribed values to the right-hand
    # side, then zero the rows and columns in place (the entries stay stored)
    B -= K @ xfix
    r_of = np.repeat(np.arange(3 * N), np.diff(K.indptr))
    K.data[fixed[r_of] | fixed[K.indices]] = 0.0
    K.data[fixed[r_of] & (r_of == K.indices)] = 1.0
    B[fixed] = xfix[fixed]
    # 3x3 blocks: row i's items ascending, split at the diagonal
    Kb = K.tobsr(blocksize=(3, 3))
    Kb.sort_indices()
    indexL, indexU, itemL, itemU, AL, AU = [0], [0], [], [], [], []
    Dd = np.zeros((N, 3, 3))
    for i in range(N):
        for q in range(Kb.indptr[i], Kb.indptr[i + 1]):
            j = Kb.indices[q]
            if j < i:
                itemL.append(j + 1)
                AL.append(Kb.data[q])
            elif j > i:
                itemU.append(j + 1)
                AU.append(Kb.data[q])
            else:
                Dd[i] = Kb.data[q]
        indexL.append(len(itemL))
        indexU.append(len(itemU))
    m = HecMatrix(N, Dd.ravel(), np.array(AL).ravel() if AL else np.zeros(0),
                  np.array(AU).ravel() if AU else np.zeros(0), np.array(indexL, np.int32),
                  np.array(itemL, np.int32), np.array(indexU, np.int32), np.array(itemU, np.int32), B=B)
    m.set_method(_method_code(solver.get("METHOD", "CG"), cnt_path))
    if "METHOD2" in solver:
        m.set_method2(_method_code(solver["METHOD2"], cnt_path, library_only=False))
    m.set_precond(solver.get("PRECOND", 1))
    if "ITER" in solver:
        m.set_iter(solver["ITER"])
    if "ITERPREMAX" in solver:
        m.set_iterpremax(solver["ITERPREMAX"])
    if "NREST" in solver:
        m.set_i(6, solver["NREST"])
    if "NCOLOR_IN" in solver:
        m.set_ncolor_in(solver["NCOLOR_IN"])
    if "RECYCLEPRE" in solver:
        m.set_i(35, solver["RECYCLEPRE"])
    if "SCALING" in solver:
        m.set_i(7, solver["SCALING"])
    if "ITERLOG" in solver:
        m.set_i(21, solver["ITERLOG"])
    if "TIMELOG" in solver:
        m.set_i(22, solver["TIMELOG"])
    if "RESID" in solver:
        m.set_resid(solver["RESID"])
    if "SIGMA_DIAG" in solver:
        m.Rarray[m.R_SIGMA_DIAG - 1] = solver["SIGMA_DIAG"]
    if "SIGMA" in solver:
        m.Rarray[m.R_SIGMA - 1] = solver["SIGMA"]
    return m, ids
