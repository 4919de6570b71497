"""hecmw_mat_dump's on-disk formats (frontistr_amd/dump.py) against the
reference's own writer (hecmw1/src/solver/matrix/hecmw_matrix_dump.f90, built
in oracle/_ref and called through ref_mat_dump):

* the files frontistr_amd.dump.write_dump produces are byte-identical to the
  reference's for MM, CSR and BSR, with the .rhs and .sol files;
* reading the reference's files gives back the profile, the item order in
  every row and every block exactly, with values equal to their 12-digit
  (e20.12e3) rounding;
* the formatter agrees with known Fortran outputs (zero, signed zero,
  subnormal, a rounding tie).
"""
import filecmp
import os

import numpy as np
import pytest

import frontistr_amd as fa
from frontistr_amd import dump
from oracle import ref
from tests.systems import random_system

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def _systems():
    m = fa.gen_hex8(3, kind="lognormal", seed=4)
    yield "hex8_3", m
    yield "random_60", random_system(60, deg=6, seed=2)
    yield "random_ext", random_system(40, deg=6, seed=5, n_ext=6)  # NP > N: the dump holds all NP rows
    m, _ = fa.gen_tet10(1, 2, 1, box=(1.0, 2.0, 0.5), seed=3)
    yield "tet10", m


SYSTEMS = dict(_systems())


def _with_x(m, seed):
    m = m.copy()
    m.X[:] = np.random.default_rng(seed).standard_normal(m.X.size) * 1e-3
    return m


def test_fortran_e_format_known_values():
    assert dump._fe(0.0) == " 0.000000000000E+000"
    assert dump._fe(-0.0) == "-0.000000000000E+000"
    assert dump._fe(210000.0) == " 0.210000000000E+006"
    assert dump._fe(-2.5) == "-0.250000000000E+001"
    assert dump._fe(9.9999999999995) == " 0.100000000000E+002"
    assert dump._fe(123456789012.5) == " 0.123456789012E+012"  # tie to even
    assert dump._fe(4.9e-324) == " 0.494065645841E-323"
    assert dump._fe(1e300) == " 0.100000000000E+301"


@pytest.mark.parametrize("kind", ["mm", "csr", "bsr"])
@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_writer_matches_reference_bytes(tmp_path, name, kind):
    m = _with_x(SYSTEMS[name], 1)
    refpath = ref.mat_dump(m, kind, str(tmp_path))
    mine = os.path.join(str(tmp_path), "mine." + kind)
    dump.write_dump(m, kind, mine)
    stem_r, stem_m = os.path.splitext(refpath)[0], os.path.splitext(mine)[0]
    for ext in (kind, "rhs", "sol"):
        assert filecmp.cmp(stem_r + "." + ext, stem_m + "." + ext, shallow=False), ext


@pytest.mark.parametrize("kind", ["mm", "csr", "bsr"])
@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_reader_restores_reference_dump(tmp_path, name, kind):
    m = _with_x(SYSTEMS[name], 2)
    refpath = ref.mat_dump(m, kind, str(tmp_path))
    r = dump.read_dump(refpath)
    assert r.N == m.NP and r.NP == m.NP
    for a in ("indexL", "indexU", "itemL", "itemU"):
        assert np.array_equal(getattr(r, a), getattr(m, a)), a
    r12 = lambda a: np.array([float(dump._fe(v)) for v in a])
    for a in ("D", "AL", "AU", "B", "X"):
        assert np.array_equal(getattr(r, a), r12(getattr(m, a))), a
        assert np.allclose(getattr(r, a), getattr(m, a), rtol=1e-11, atol=0.0), a


def test_planted_mistake_is_caught(tmp_path):
    """A transposed block in the writer would change the bytes."""
    m = _with_x(SYSTEMS["hex8_3"], 3)
    refpath = ref.mat_dump(m, "bsr", str(tmp_path))
    bad = m.copy()
    bad.AL[:] = bad.AL.reshape(-1, 3, 3).transpose(0, 2, 1).ravel()  # off-diagonal blocks are not symmetric
    mine = os.path.join(str(tmp_path), "bad.bsr")
    dump.write_dump(bad, "bsr", mine)
    assert not filecmp.cmp(refpath, mine, shallow=False)
