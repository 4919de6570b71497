"""bench.py on the host: the one-GPU rule and the algorithmic-byte formulas
its roofline fields are computed from (no GPU needed)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import frontistr_amd as fa  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr and p.stdout == ""


def test_spmv_bytes_is_the_reference_formula_on_the_100_cube():
    # hecmw_solver_Iterative.f90:212-217 with merged indices, on the 100^3 cube's own sizes
    m = fa.gen_hex8_structure(100)
    nnzb = m.N + int(m.indexL[m.N]) + int(m.indexU[m.N])
    assert (m.N, m.NP, nnzb) == (1030301, 1030301, 27270901)
    b = bench.spmv_bytes(m.N, m.NP, nnzb, 8)
    # 72 B of values and 4 B of column per block, row pointers, x read and y written once
    assert b == 72 * nnzb + 4 * nnzb + 4 * (m.N + 1) + 24 * m.NP + 24 * m.N == 2126164132
    assert bench.spmv_bytes(m.N, m.NP, nnzb, 4) < b


def test_ssor_bytes_counts_every_off_diagonal_block_once():
    m = fa.gen_hex8_structure(160)
    nnz_lu = int(m.indexL[m.N]) + int(m.indexU[m.N])
    assert (m.N, nnz_lu) == (4173281, 107111360)
    assert bench.ssor_bytes(m.N, nnz_lu, 4) == 4735168748
    # each block of L' and U' once (values + column), plus per-row terms that do not depend on the blocks
    assert bench.ssor_bytes(m.N, nnz_lu + 1, 4) - bench.ssor_bytes(m.N, nnz_lu, 4) == 9 * 4 + 4
    # BF16 L'/U' values: 22 B per block, the per-row terms unchanged
    assert bench.ssor_bytes(m.N, nnz_lu, 4, 2) == 4735168748 - 18 * nnz_lu == 2807164268
    assert bench.ssor_bytes(m.N, nnz_lu + 1, 4, 2) - bench.ssor_bytes(m.N, nnz_lu, 4, 2) == 9 * 2 + 4
