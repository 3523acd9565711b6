"""GPU parity of the Coulomb and exchange energy-gradient kernels.

* small8 and lattice4 (every class through (dd|dd)): the per-block rows of
  every J and K gradient class against the oracle's (orc_coulomb_grad_class /
  orc_exchange_grad_class, pinned by finite differences in
  test_oracle_grad.py), 1e-11 relative to the largest row element, and the
  per-atom gradients; omega = 0.4 exchange gradient;
* lattice64: every J class on sampled bra blocks and every K class on sampled
  (i,k) blocks (full ket lists / CSR), against the oracle on the same blocks;
  bitwise reproducibility of the whole gradient; the sum over atoms (zero for
  a translation-invariant energy, up to the screened quartets);
* small8: the per-atom gradient against central finite differences of the
  GPU energies E_J = 1/2 tr(D J), E_K = 1/2 tr(D K) at displaced geometries
  (1e-6 relative: the step's truncation error and the screening change).
"""

import numpy as np
import pytest

import _util as U
from vlxmi import configs
from vlxmi import grad as G
from vlxmi import pairs as P
from vlxmi.basis import Molecule, Shell

TOL = 1e-11
pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def dev():
    from vlxmi.runtime import Device
    return Device()


def _engines(dev, mol, omega=0.0):
    from vlxmi.fock import FockEngine
    eng = FockEngine(mol, dev=dev, omega=omega)
    return eng, G.GradientEngine(eng)


def _rows_close(got, ref, what):
    scale = max(1.0, max(np.abs(v).max() for v in ref.values() if v.size))
    for k in ref:
        err = np.abs(got[k] - ref[k]).max() if ref[k].size else 0.0
        assert err <= TOL * scale, (what, k, err, scale)


@pytest.mark.parametrize("wl", ["small8", "lattice4"])
def test_gradient_rows_and_atoms(dev, wl):
    mol, D = configs.make_workload(wl)
    _say("case: %s J/K gradient, every class against the oracle" % wl)
    eng, ge = _engines(dev, mol)
    Dc = eng.cart_density(D)
    Dc_np = U.cart_density(mol, D)
    rJ = {k: v.cpu().numpy() for k, v in ge.run_J(Dc).items()}
    rK = {k: v.cpu().numpy() for k, v in ge.run_K(Dc).items()}
    oJ, oK = U.oracle_gradJ(eng.fl, Dc_np), U.oracle_gradK(eng.fl, Dc_np)
    _rows_close(rJ, oJ, "J")
    _rows_close(rK, oK, "K")
    gJ, gK = ge.gradient(D)
    gJ, gK = gJ.cpu().numpy(), gK.cpu().numpy()
    aJ, aK = U.oracle_gradient(eng.fl, Dc_np)
    _say("  max|gJ| %.3e max|gK| %.3e, errors vs oracle %.2e %.2e"
         % (np.abs(aJ).max(), np.abs(aK).max(), np.abs(gJ - aJ).max(), np.abs(gK - aK).max()))
    assert np.abs(gJ - aJ).max() <= TOL * max(1.0, np.abs(aJ).max())
    assert np.abs(gK - aK).max() <= TOL * max(1.0, np.abs(aK).max())


def test_lattice4_range_separated_exchange_gradient(dev):
    mol, D = configs.make_lattice4()
    _say("case: lattice4 exchange gradient, omega = 0.4")
    eng, ge = _engines(dev, mol, omega=0.4)
    Dc = eng.cart_density(D)
    rK = {k: v.cpu().numpy() for k, v in ge.run_K(Dc).items()}
    _rows_close(rK, U.oracle_gradK(eng.fl, U.cart_density(mol, D), omega=0.4), "K omega")


@pytest.fixture(scope="module")
def lat64(dev):
    mol, D = configs.make_lattice64()
    eng, ge = _engines(dev, mol)
    return mol, D, eng, ge


def test_lattice64_j_gradient_sampled(lat64):
    mol, D, eng, ge = lat64
    Dc = eng.cart_density(D)
    Dc_np = U.cart_density(mol, D)
    kets = eng.prepare_kets(Dc)
    rng = np.random.default_rng(15)
    sel = {}
    for ab in eng.jl:
        n = eng.fl.jlists[ab].nblk
        sel[ab] = np.unique(np.concatenate([[0, n - 1], rng.choice(n, size=min(n, 10), replace=False)])).astype(np.uint32)
    for ab, cd in P.j_classes():
        _say("case: lattice64 J gradient class %s%s (sampled bra blocks)" % (ab, cd))
        g = ge.run_J(Dc, kets=kets, classes=[(ab, cd)], bra_blocks=sel)[ab].cpu().numpy()
        r = U.oracle_gradJ(eng.fl, Dc_np, bra_blocks=sel, classes=[(ab, cd)])[ab]
        assert np.abs(g - r).max() <= TOL * max(1.0, np.abs(r).max()), (ab, cd)


@pytest.mark.parametrize("cls", P.k_classes())
def test_lattice64_k_gradient_sampled(lat64, cls):
    mol, D, eng, ge = lat64
    Dc = eng.cart_density(D)
    ac = cls[0] + cls[2]
    ik = eng.fl.iklists[ac]
    rng = np.random.default_rng(700 + P.k_classes().index(cls))
    sel = np.unique(np.concatenate([[0, ik.nblk - 1], rng.choice(ik.nblk, size=min(ik.nblk, 4), replace=False)]))
    sel = sel.astype(np.uint32)
    _say("case: lattice64 K gradient class %s (%d sampled (i,k) blocks)" % (cls, sel.size))
    g = ge.run_K(Dc, classes=[cls], ik_blocks={ac: sel})[ac].cpu().numpy()
    r = U.oracle_gradK(eng.fl, U.cart_density(mol, D), ik_blocks={ac: sel}, classes=[cls])[ac]
    assert np.abs(g - r).max() <= TOL * max(1.0, np.abs(r).max())


def test_lattice64_gradient_properties(lat64):
    mol, D, eng, ge = lat64
    _say("case: lattice64 full gradient twice: bitwise, sum over atoms")
    gJ1, gK1 = ge.gradient(D)
    gJ2, gK2 = ge.gradient(D)
    assert bool((gJ1 == gJ2).all()) and bool((gK1 == gK2).all())
    gJ, gK = gJ1.cpu().numpy(), gK1.cpu().numpy()
    sJ, sK = np.abs(gJ.sum(axis=0)).max(), np.abs(gK.sum(axis=0)).max()
    _say("  max|gJ| %.3e |sum gJ| %.2e, max|gK| %.3e |sum gK| %.2e" % (np.abs(gJ).max(), sJ, np.abs(gK).max(), sK))
    assert sJ <= 1e-8 * np.abs(gJ).max() and sK <= 1e-8 * np.abs(gK).max()


def _displaced(mol, atom, axis, h):
    coords = mol.coords.copy()
    coords[atom, axis] += h
    shells = [Shell(s.atom, s.l, coords[s.atom], s.exps, s.coefs, s.spherical) for s in mol.shells]
    return Molecule(coords, shells, name=mol.name)


def test_small8_gradient_vs_finite_difference(dev):
    from vlxmi.fock import FockEngine
    mol, D = configs.make_small8()
    _say("case: small8 gradient vs central differences of the GPU energies")
    eng, ge = _engines(dev, mol)
    gJ, gK = ge.gradient(D)
    gJ, gK = gJ.cpu().numpy(), gK.cpu().numpy()
    h = 1e-4
    fJ, fK = np.zeros_like(gJ), np.zeros_like(gK)
    for at in range(mol.natoms):
        for ax in range(3):
            e = []
            for s in (h, -h):
                m2 = _displaced(mol, at, ax, s)
                e2 = FockEngine(m2, dev=dev)
                J, K = e2.build(D)
                e.append(e2.energies(D, J, K))
            fJ[at, ax] = (e[0][0] - e[1][0]) / (2 * h)
            fK[at, ax] = (e[0][1] - e[1][1]) / (2 * h)
    _say("  max|gJ - fd| %.2e (max|gJ| %.2e), max|gK - fd| %.2e (max|gK| %.2e)"
         % (np.abs(gJ - fJ).max(), np.abs(gJ).max(), np.abs(gK - fK).max(), np.abs(gK).max()))
    assert np.abs(gJ - fJ).max() <= 1e-6 * np.abs(fJ).max()
    assert np.abs(gK - fK).max() <= 1e-6 * np.abs(fK).max()
