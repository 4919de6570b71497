"""fistr1's own test models assembled by frontistr_amd.mesh, pinned on the
CPU: the CG solver test (tests/golden/fistr_cg, from the reference's
tests/solver/CG) and the static examples A-E with TET4, TET10 and HEX8
elements (tests/golden/fistr_static, from tests/analysis/static: a point
load, a face pressure (B, TET4 and HEX8), a body force per volume, gravity
times the density, a centrifugal load):

* the displacement fistr1 reported (.res, SSOR and DIAG runs) solves the
  assembled system to fistr1's RESID (1e-8): the element stiffness, material,
  profile, Dirichlet elimination and load are fistr1's;
* a dense direct solve of the assembled system equals that displacement to
  the CG's accuracy;
* the reference's own CG (oracle/_ref) on the assembled system reproduces it;
* planted mistakes are caught by the residual check: the tensor instead of
  the engineering shear modulus, a TET10's midpoints taken in the file's
  order instead of rotated onto the shape functions' edges, and a HEX8 with
  plain integration instead of fistr1's incompatible modes.
"""
import os

import numpy as np
import pytest

from frontistr_amd import mesh
from oracle import ref
from tests.systems import dense

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fistr_cg")
STATIC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fistr_static")
# name -> (mesh, control, result)
MODELS = {
    "cg_ssor": (os.path.join(HERE, "SSOR.msh"), os.path.join(HERE, "SSOR.cnt"), os.path.join(HERE, "SSOR.res.0.1")),
    "cg_diag": (os.path.join(HERE, "SSOR.msh"), os.path.join(HERE, "DIAG.cnt"), os.path.join(HERE, "DIAG.res.0.1")),
    "static_A341": (os.path.join(HERE, "SSOR.msh"), os.path.join(STATIC, "A341.cnt"),
                    os.path.join(STATIC, "A341.res.0.1")),
    "static_C341": (os.path.join(STATIC, "C341.msh"), os.path.join(STATIC, "C341.cnt"),
                    os.path.join(STATIC, "C341.res.0.1")),
    "static_D341": (os.path.join(STATIC, "D341.msh"), os.path.join(STATIC, "D341.cnt"),
                    os.path.join(STATIC, "D341.res.0.1")),
}
for _t in ("341", "361"):
    MODELS["static_B%s" % _t] = tuple(os.path.join(STATIC, "B%s.%s" % (_t, ext)) for ext in ("msh", "cnt", "res.0.1"))
for _t in ("341", "342", "361"):
    MODELS["static_E%s" % _t] = tuple(os.path.join(STATIC, "E%s.%s" % (_t, ext)) for ext in ("msh", "cnt", "res.0.1"))
for _x in "ACD":
    for _t in ("342", "361"):
        MODELS["static_%s%s" % (_x, _t)] = tuple(os.path.join(STATIC, "%s%s.%s" % (_x, _t, ext))
                                                  for ext in ("msh", "cnt", "res.0.1"))


def model(name):
    msh, cnt, res = MODELS[name]
    m, ids = mesh.assemble(msh, cnt)
    u = mesh.read_res_displacement(res)
    return m, ids, np.array([u[n] for n in ids]).ravel()


def _model(prec):
    return model("cg_" + prec.lower())


def test_reader_sizes_and_solver_settings():
    nodes, elems, material, groups, egroups = mesh.read_msh(os.path.join(HERE, "SSOR.msh"))
    assert egroups == {}
    assert mesh.read_msh(MODELS["static_B341"][0])[4]["B340P3"] == list(range(125, 240, 6)) + list(range(126, 241, 6))
    assert len(elems) == 240 and all(t == 341 and len(ns) == 4 for _, t, ns in elems)
    assert material == (4000.0, 0.3, None)
    assert mesh.read_msh(MODELS["static_D341"][0])[2] == (4000.0, 0.3, 8.0102e-10)
    assert groups["FIX"] == [1001, 1101, 1201, 3001, 3101, 3201, 5001, 5101, 5201]
    for prec, code in (("SSOR", 1), ("DIAG", 3)):
        m, ids, u = _model(prec)
        assert m.N == len(ids) == 99 and u.size == 297
        assert m.get_i(m.I_METHOD) == 1 and m.get_i(m.I_PRECOND) == code
        assert m.get_i(m.I_ITER) == 10000 and m.Rarray[0] == 1e-8


@pytest.mark.parametrize("name", sorted(MODELS))
def test_fistr1_displacement_solves_the_assembled_system(name):
    m, ids, u = model(name)
    A = dense(m)
    # symmetric to rounding: B^T D B summed over quadrature points
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
    rel = np.linalg.norm(A @ u - m.B) / np.linalg.norm(m.B)
    assert rel <= 2e-8, rel
    x = np.linalg.solve(A, m.B)
    assert np.abs(x - u).max() <= 1e-8 * np.abs(u).max()


def test_planted_tet10_midpoint_order_mistake_is_caught(monkeypatch):
    """The .msh lists a 342's midpoints as edges 2-3, 3-1, 1-2: taken in the
    shape functions' order instead, the mid-edge nodes land on the wrong edges."""
    m, ids, u = model("static_A342")
    monkeypatch.setattr(mesh, "_HEC342_TO_SHAPE", tuple(range(10)))
    monkeypatch.setattr(mesh, "_tet10_points", _tet10_points_any_sign)
    msh, cnt, _ = MODELS["static_A342"]
    mw, _ = mesh.assemble(msh, cnt)
    rel = np.linalg.norm(dense(mw) @ u - mw.B) / np.linalg.norm(mw.B)
    assert rel > 1e-2, rel


def _tet10_points_any_sign(X):
    out = []
    for xi, et, ze in mesh._TET10_POINTS:
        shp, g = mesh._tet10_shape(xi, et, ze)
        J = g.T @ X
        out.append((shp, g @ np.linalg.inv(J).T, abs(np.linalg.det(J)) / 24.0))
    return out


def test_planted_hex8_full_integration_is_caught(monkeypatch):
    """fistr1's static 361 is the incompatible-mode element: plain 2x2x2
    integration (no bubble modes) is stiffer in bending and is caught."""
    m, ids, u = model("static_A361")

    def full(X, D):
        K = np.zeros((24, 24))
        for _, dN, w, _, _ in mesh._hex8_points(X):
            B = mesh._B(dN)
            K += w * (B.T @ D @ B)
        return K
    monkeypatch.setattr(mesh, "_hex8ic_stiffness", full)
    msh, cnt, _ = MODELS["static_A361"]
    mw, _ = mesh.assemble(msh, cnt)
    rel = np.linalg.norm(dense(mw) @ u - mw.B) / np.linalg.norm(mw.B)
    assert rel > 1e-2, rel


def test_planted_pressure_sign_mistake_is_caught(monkeypatch):
    """The pressure's normal follows the face's node order (right-hand rule);
    the opposite orientation is caught."""
    m, ids, u = model("static_B361")
    orig = mesh._face_points
    monkeypatch.setattr(mesh, "_face_points", lambda X: [(s_, -n_) for s_, n_ in orig(X)])
    msh, cnt, _ = MODELS["static_B361"]
    mw, _ = mesh.assemble(msh, cnt)
    rel = np.linalg.norm(dense(mw) @ u - mw.B) / np.linalg.norm(mw.B)
    assert rel > 1e-2, rel


def test_planted_shear_mistake_is_caught():
    m, ids, u = _model("SSOR")
    orig = mesh._D_iso
    try:
        def wrong(E, nu):
            D = orig(E, nu)
            D[3:, 3:] *= 2.0  # tensor instead of engineering shear strains
            return D
        mesh._D_iso = wrong
        mw, _ = mesh.assemble(os.path.join(HERE, "SSOR.msh"), os.path.join(HERE, "SSOR.cnt"))
    finally:
        mesh._D_iso = orig
    rel = np.linalg.norm(dense(mw) @ u - mw.B) / np.linalg.norm(mw.B)
    assert rel > 1e-2, rel


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", sorted(MODELS))
def test_reference_cg_reproduces_fistr1(name):
    m, ids, u = model(name)
    x, info = ref.solve_krylov(m)
    assert info["error"] == 0
    assert np.abs(x - u).max() <= 1e-6 * np.abs(u).max()
