#!/usr/bin/env python3
"""Writes tests/golden/hecmw33_golden.npz: inputs and outputs of the
reference's own solver routines (oracle/_ref, compiled from the FrontISTR
sources) on small systems, so the oracle restatement and the GPU library can be
pinned without the reference checkout.

Systems: a 4^3 HEX8 cube as fistr1 assembles it (frontistr_amd.gen_hex8), the
same cube with items shuffled inside every row, and a random SPD block system
with external nodes (tests/systems.py).  Outputs per system: RCM and
multicolour permutations, COLORindex, y = A x, z = M^-1 r for SSOR and DIAG,
CG and BiCGSTAB solutions with iteration counts, residuals and ITERLOG histories.

    python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import frontistr_amd as fa  # noqa: E402
from oracle import ref  # noqa: E402
from tests.systems import random_system, shuffle_items  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hecmw33_golden.npz")


def systems():
    yield "hex8_4", fa.gen_hex8(4)
    yield "hex8_4_shuffled", shuffle_items(fa.gen_hex8(4), seed=7)
    yield "random_150_ext", random_system(150, deg=8, seed=21, n_ext=12)


def main():
    data = {}
    rng = np.random.default_rng(1234)
    for name, m in systems():
        for k in ("D", "AL", "AU", "indexL", "itemL", "indexU", "itemU", "B"):
            data["%s/%s" % (name, k)] = getattr(m, k)
        data["%s/sizes" % name] = np.array([m.N, m.NP], np.int32)
        pr, _, p, ip, ci = ref.ordering(m, 10)
        data["%s/perm_rcm" % name] = pr
        data["%s/perm" % name] = p
        data["%s/colorindex" % name] = ci
        x = rng.standard_normal(3 * m.NP)
        data["%s/x" % name] = x
        data["%s/Ax" % name] = ref.matvec(m, x)
        r = rng.standard_normal(3 * m.NP)
        data["%s/r" % name] = r
        for pc in (1, 3):
            mm = m.copy()
            mm.set_precond(pc)
            data["%s/z_pc%d" % (name, pc)] = ref.precond(mm, r, nthreads=2)
        for meth in (1, 2):
            for pc in (1, 3):
                mm = m.copy()
                mm.set_iter(2000)
                mm.set_method(meth)
                mm.set_precond(pc)
                xs, info = ref.solve_krylov(mm, nthreads=2, history=True)
                key = "%s/solve_m%d_pc%d" % (name, meth, pc)
                data[key + "/x"] = xs
                data[key + "/info"] = np.array([info["iter"], info["error"]], np.int32)
                data[key + "/resid"] = np.array([info["resid"]])
                data[key + "/history"] = info["history"]
    np.savez_compressed(OUT, **data)
    print("wrote %s (%d arrays, %.0f KB)" % (OUT, len(data), os.path.getsize(OUT) / 1024))


if __name__ == "__main__":
    main()
