"""Host-side logic: basis conventions, workload sizes, the reference data layout
and the ordering contract the kernels rely on, AO assembly."""

import math

import numpy as np
import pytest

import _util as U
from vlxmi import basis, configs
from vlxmi import pairs as P
from vlxmi.basis import NCART


def test_workload_sizes():
    mol, D = configs.make_small8()
    assert (mol.natoms, mol.nbf, mol.naos_cart) == (8, 40, 40) and D.shape == (40, 40)
    mol, D = configs.make_lattice64()
    assert (mol.natoms, mol.nbf, mol.naos_cart) == (64, 896, 960) and D.shape == (896, 896)
    rng = np.random.default_rng(1003)
    mol = configs.lattice_molecule(rng, (8, 8, 4))
    assert (mol.natoms, mol.nbf) == (256, 3584)
    coords = configs.make_small8()[0].coords
    dmin = min(np.linalg.norm(coords[i] - coords[j]) for i in range(8) for j in range(i))
    assert dmin >= 1.5 and np.all(np.abs(coords) <= 3.0)


def test_normalisation_and_spherical_d():
    mol, _ = configs.make_small8()
    S = basis.overlap_matrix_cart(mol)
    assert np.max(np.abs(np.diag(S) - 1)) < 1e-13
    rng = np.random.default_rng(1)
    m = basis.Molecule(np.zeros((1, 3)), [basis.Shell(0, 2, np.zeros(3), [0.7, 2.0], [0.4, 0.6], spherical=True)])
    Sc = basis.overlap_matrix_cart(m)
    T = m.cart2sph()
    Ss = T @ Sc @ T.T
    assert np.max(np.abs(Ss - np.eye(5))) < 1e-13


@pytest.fixture(scope="module")
def lists():
    mol, D = configs.make_small8()
    fl = P.build_lists(mol, U.oracle_schwarz)
    return mol, D, fl


def test_jlist_block_contract(lists):
    mol, D, fl = lists
    Dc = U.cart_density(mol, D)
    for pc, jl in fl.jlists.items():
        na, nb = NCART[jl.la], NCART[jl.lb]
        f = jl.first.reshape(-1, na, nb)
        s = jl.second.reshape(-1, na, nb)
        # complete blocks: one primitive pair per block, a-major component order
        assert np.all(f // na == f[:, :1, :1] // na) and np.all(s // nb == s[:, :1, :1] // nb)
        assert np.all(f % na == np.arange(na)[None, :, None]) and np.all(s % nb == np.arange(nb)[None, None, :])
        if jl.la == jl.lb:
            assert np.all(f[:, 0, 0] // na <= s[:, 0, 0] // nb)
        mD = jl.mat_D(Dc)
        assert np.all(mD[jl.lower_diag] == 0.0)
        order = jl.ket_order(mD)
        beta = np.abs(jl.mat_Q * mD).reshape(jl.nblk, -1).max(axis=1)[order]
        assert np.all(np.diff(beta) <= 0)
        # bra order: non-increasing Qmax
        assert np.all(np.diff(jl.qmax_blk) <= 0)


def test_kcsr_contract(lists):
    mol, D, fl = lists
    for pc, cs in fl.kcsr.items():
        na, nb = NCART[cs.la], NCART[cs.lb]
        cnt_a = fl.prims[cs.la].count
        for I in range(cnt_a):
            cnt = cs.counts[I * na]
            assert np.all(cs.counts[I * na:(I + 1) * na] == cnt) and cnt % nb == 0
            segs = [cs.D_inds[cs.displs[I * na + t]: cs.displs[I * na + t] + cnt] for t in range(na)]
            for sgm in segs[1:]:
                assert np.array_equal(sgm, segs[0])            # same partner order for every component
            q = np.max([np.abs(cs.Q[cs.displs[I * na + t]: cs.displs[I * na + t] + cnt]).reshape(-1, nb).max(axis=1)
                        for t in range(na)], axis=0) if cnt else np.zeros(0)
            assert np.all(np.diff(q) <= 0)                      # partner blocks by non-increasing QKmax
            if cnt:
                assert np.array_equal(segs[0].reshape(-1, nb) % nb, np.tile(np.arange(nb), (cnt // nb, 1)))


def test_assembly_is_symmetric(lists):
    mol, D, fl = lists
    Dc = U.cart_density(mol, D)
    mJ = U.oracle_matJ(fl, Dc)
    J = P.assemble_J(fl, mJ)
    assert np.max(np.abs(J - J.T)) < 1e-12
    counts = {"J": len(P.j_classes()), "K": len(P.k_classes())}
    assert counts == {"J": 36, "K": 54}
