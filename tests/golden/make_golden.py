"""Writes the committed fixtures of this directory (run from the repo root):

  small8_jk.npz  - small8 workload: Cartesian density, oracle J and K (reference
                   per-component screening over this repo's lists) and the
                   unscreened 4-index brute-force J and K, energies.
  md_blocks.npz  - McMurchie-Davidson (mpmath Boys) blocks of one random
                   primitive quartet per class (s,p,d)^4, with their inputs.

The reference repository holds no numeric fixtures; these are produced by the
independent evaluators of this repository (the MD evaluator does not use the oracle).
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "vlx-twoe-ints_amd"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import _util as U  # noqa: E402
import md_eri  # noqa: E402
import oracle  # noqa: E402
from vlxmi import configs  # noqa: E402
from vlxmi import pairs as P  # noqa: E402


def small8():
    mol, D = configs.make_small8()
    fl = P.build_lists(mol, U.oracle_schwarz)
    Dc = U.cart_density(mol, D)
    J, K = U.oracle_fock(fl, Dc, "component")
    eri = oracle.ao_eri_tensor(mol)
    Jb, Kb = oracle.jk_from_tensor(eri, Dc)
    np.savez_compressed(os.path.join(HERE, "small8_jk.npz"), D=D, D_cart=Dc, J=J, K=K, J_bf=Jb, K_bf=Kb,
                        E_J=0.5 * np.sum(D * J), E_K=0.5 * np.sum(D * K))


def md_blocks():
    rng = np.random.default_rng(2024)
    out = {}
    for la in range(3):
        for lb in range(3):
            for lc in range(3):
                for ld in range(3):
                    pts = rng.uniform(-2, 2, (4, 3))
                    ex = np.exp(rng.uniform(math.log(0.1), math.log(20), 4))
                    blk = np.array(md_eri.eri_block_md(la, lb, lc, ld, ex[0], pts[0], ex[1], pts[1],
                                                       ex[2], pts[2], ex[3], pts[3]))
                    key = "spd"[la] + "spd"[lb] + "spd"[lc] + "spd"[ld]
                    out[key + "_exps"] = ex
                    out[key + "_centres"] = pts
                    out[key + "_block"] = blk
    np.savez_compressed(os.path.join(HERE, "md_blocks.npz"), **out)


if __name__ == "__main__":
    small8()
    md_blocks()
    print("fixtures written")
