"""The list-based oracle J/K (reference screening over this repository's
reference-layout lists) against the unscreened 4-index brute force, the
committed fixtures, symmetry/positivity, and sensitivity to planted mistakes."""

import os

import numpy as np
import pytest

import _util as U
import oracle
from vlxmi import configs
from vlxmi import pairs as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def small8():
    mol, D = configs.make_small8()
    fl = P.build_lists(mol, U.oracle_schwarz)
    Dc = U.cart_density(mol, D)
    return mol, D, Dc, fl


def test_md_fixture_blocks():
    g = np.load(os.path.join(GOLD, "md_blocks.npz"))
    for la in range(3):
        for lb in range(3):
            for lc in range(3):
                for ld in range(3):
                    key = "spd"[la] + "spd"[lb] + "spd"[lc] + "spd"[ld]
                    ex, pts, blk = g[key + "_exps"], g[key + "_centres"], g[key + "_block"]
                    v = oracle.eri_block(la, lb, lc, ld, ex[0], pts[0], ex[1], pts[1], ex[2], pts[2], ex[3], pts[3])
                    assert np.max(np.abs(v - blk)) <= 1e-12 * np.max(np.abs(blk)), key


def test_small8_vs_bruteforce_and_fixture(small8):
    mol, D, Dc, fl = small8
    J, K = U.oracle_fock(fl, Dc, "component")
    g = np.load(os.path.join(GOLD, "small8_jk.npz"))
    assert np.max(np.abs(J - g["J"])) <= 1e-12 and np.max(np.abs(K - g["K"])) <= 1e-12
    # unscreened brute force: only screened terms (each <= 2 thr) differ
    assert np.max(np.abs(J - g["J_bf"])) < 1e-9
    assert np.max(np.abs(K - g["K_bf"])) < 1e-9
    assert np.max(np.abs(J - J.T)) < 1e-12 and np.max(np.abs(K - K.T)) < 1e-12
    EJ, EK = 0.5 * np.sum(D * J), 0.5 * np.sum(D * K)
    assert EJ > 0 and EK > 0 and EK < EJ


def test_planted_weight_mistake_is_caught(small8):
    """Dropping the k<->l doubling of off-diagonal ket entries (gen_coulomb.py:276-281)
    must be detected by the brute-force comparison."""
    mol, D, Dc, fl = small8
    g = np.load(os.path.join(GOLD, "small8_jk.npz"))
    mJ = U.oracle_matJ(fl, Dc, "component")
    # planted mistake: count every ket entry once (halve the off-diagonal contributions)
    Dhalf = Dc.copy()
    mJ_bad = U.oracle_matJ(fl, 0.5 * Dhalf, "component")
    Jbad = P.assemble_J(fl, {k: 0.5 * (mJ[k] + mJ_bad[k]) for k in mJ})
    assert np.max(np.abs(Jbad - g["J_bf"])) > 1e-3


def test_block_rule_differs_from_reference_rule(small8):
    """Screening per primitive-pair block instead of per Cartesian component pair
    changes J/K by more than the 1e-11 parity tolerance: the HIP path must apply
    the per-component rule."""
    mol, D, Dc, fl = small8
    Jc, Kc = U.oracle_fock(fl, Dc, "component")
    Jb, Kb = U.oracle_fock(fl, Dc, "block")
    assert max(np.max(np.abs(Jc - Jb)), np.max(np.abs(Kc - Kb))) > 1e-11
