"""A system the reference wrote with hecmw_mat_dump (BSR and CSR, with its
right-hand side) becomes a GPU fixture as it is: read back by
frontistr_amd.dump and solved through the C ABI, it gives the reference's
SSOR-CG solution of the same arrays within 1e-6, with the same iteration
count to within 2."""
import numpy as np
import pytest

import frontistr_amd as fa
from frontistr_amd import dump
from oracle import ref

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")]


@pytest.mark.parametrize("kind", ["bsr", "csr"])
def test_dumped_system_solves_on_gpu(tmp_path, kind):
    m = fa.gen_hex8(6, kind="lognormal", seed=9)
    path = ref.mat_dump(m, kind, str(tmp_path))
    r = dump.read_dump(path)
    r.set_iter(5000)
    r.set_precond(1)
    xr, info = ref.solve_krylov(r, nthreads=2)
    assert info["error"] == 0
    g = r.copy()
    res = fa.hecmw_solve_iterative(g)
    assert res.error == 0
    assert abs(res.iter - info["iter"]) <= 2
    assert np.linalg.norm(g.X - xr) <= 1e-6 * np.linalg.norm(xr)
    assert ref.rel_resid(r, g.X) <= 1e-8 * 1.0001
