"""The MPC-CG product of the reference build (hecmw_TtmatTvec_33 through
oracle/_ref) against the closed form T^T A T x with T written out densely
(tests/mpc_systems.py): pins what the GPU's MPC product is compared with."""
import numpy as np
import pytest

import frontistr_amd as fa
from oracle import ref
from tests.mpc_systems import dense_T, tie_constraints
from tests.systems import dense

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def test_reference_mpc_product_is_TtAT():
    m = fa.gen_hex8(4, kind="lognormal", seed=1)
    mpc = tie_constraints(4)
    x = np.random.default_rng(0).standard_normal(3 * m.NP)
    ref.set_mpc(mpc)
    try:
        y = ref.matvec_mpc(m, x, nthreads=1)
    finally:
        ref.set_mpc(None)
    T = dense_T(m.N, mpc)
    want = T.T @ dense(m) @ T @ x
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # a planted mistake: T instead of T^T on the left
    assert not np.allclose(y, T @ dense(m) @ T @ x, rtol=1e-6)
