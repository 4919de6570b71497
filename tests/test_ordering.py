"""RCM + multicolour ordering: bit-exact with the reference's own routines
(hecmw_matrix_ordering_CM.f90, hecmw_matrix_ordering_MC.f90, run from
oracle/_ref) and with the race-freedom invariant of check_ordering
(hecmw_precond_SSOR_33.F90:476-513)."""
import numpy as np
import pytest

import frontistr_amd as fa
from oracle import ref
from tests.systems import random_system, shuffle_items

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def cases():
    yield "hex8_3", fa.gen_hex8(3), 10
    yield "hex8_11", fa.gen_hex8(11), 10
    yield "hex8_20", fa.gen_hex8(20), 10
    yield "hex8_9_shuffled", shuffle_items(fa.gen_hex8(9), seed=3), 10
    yield "hex8_6_ncolor1", fa.gen_hex8(6), 1
    yield "hex8_6_ncolor37", fa.gen_hex8(6), 37
    yield "hex8_2_ncolor_gt_N", fa.gen_hex8(2), 100
    yield "random_300", random_system(300, deg=10, seed=1), 10
    yield "random_500_ext", random_system(500, deg=6, seed=2, n_ext=40), 10
    yield "random_disconnected", random_system(400, deg=6, seed=5, components=3), 7
    yield "random_sparse", random_system(257, deg=2, seed=9), 10


@needs_ref
@pytest.mark.parametrize("name,m,nc", list(cases()), ids=[c[0] for c in cases()])
def test_ordering_bit_exact_with_reference(name, m, nc):
    pr, p, ip, ci = fa.hecmw_ordering(m, nc)
    rpr, _, rp, rip, rci = ref.ordering(m, nc)
    assert np.array_equal(pr, rpr), "RCM permutation differs"
    assert np.array_equal(p, rp), "multicolour permutation differs"
    assert np.array_equal(ip, rip)
    assert np.array_equal(ci, rci), "COLORindex differs"


@pytest.mark.parametrize("name,m,nc", list(cases()), ids=[c[0] for c in cases()])
def test_coloring_is_race_free(name, m, nc):
    """No L or U entry joins two internal nodes of one colour, so the rows of a
    colour can be swept in parallel; perm/iperm are inverse permutations and
    colours respect the cap N/ncolor_in."""
    pr, p, ip, ci = fa.hecmw_ordering(m, nc)
    N = m.N
    assert sorted(p.tolist()) == list(range(1, N + 1))
    assert np.array_equal(ip[p - 1], np.arange(1, N + 1))
    assert ci[0] == 0 and ci[-1] == N and np.all(np.diff(ci) > 0)
    cap = N // nc
    if cap > 0:
        assert np.all(np.diff(ci) <= cap)
    color = np.zeros(N + 1, np.int64)
    for c in range(len(ci) - 1):
        color[p[ci[c]:ci[c + 1]]] = c
    for i in range(1, N + 1):
        nb = list(m.itemL[m.indexL[i - 1]:m.indexL[i]]) + [j for j in m.itemU[m.indexU[i - 1]:m.indexU[i]] if j <= N]
        for j in nb:
            assert color[i] != color[j], (i, j)


@needs_ref
@pytest.mark.parametrize("n", [100, 160])
def test_ordering_bit_exact_at_benchmark_sizes(n):
    """The 100^3 and 160^3 cubes (1,030,301 and 4,173,281 rows, 17 colours):
    long BFS levels where the unstable middle-pivot sort's tie-breaks and the
    capped greedy colouring decide most positions
    (hecmw_matrix_ordering_CM.f90:207-233, hecmw_matrix_ordering_MC.f90:15-72)."""
    m = fa.gen_hex8_structure(n)
    pr, p, ip, ci = fa.hecmw_ordering(m, 10)
    rpr, _, rp, rip, rci = ref.ordering(m, 10)
    assert np.array_equal(pr, rpr), "RCM permutation differs"
    assert np.array_equal(p, rp), "multicolour permutation differs"
    assert np.array_equal(ip, rip)
    assert np.array_equal(ci, rci), "COLORindex differs"
    assert len(ci) - 1 == 17
