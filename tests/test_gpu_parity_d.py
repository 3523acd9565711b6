"""GPU parity for every class through (dd|dd) and at the lattice64 size.

lattice4 (2x2x1 lattice, s3p2d1 basis, nbf 56): every J and K class against
the full CPU oracle, 1e-11 max-abs (reference per-component screening).
lattice64 (nbf 896): every J class on a sample of bra blocks, every K class on
a slice of (i,k) blocks, against the oracle on the same inputs; and
size-independent properties of the full build (J, K symmetric, E_J, E_K > 0).
omega != 0: range-separated K against the oracle.
"""

import numpy as np
import pytest

import _util as U
from vlxmi import configs
from vlxmi import pairs as P

TOL = 1e-11
pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def dev():
    from vlxmi.runtime import Device
    return Device()


@pytest.fixture(scope="module")
def lat4(dev):
    from vlxmi.fock import FockEngine
    mol, D = configs.make_lattice4()
    return mol, D, FockEngine(mol, dev=dev)


@pytest.fixture(scope="module")
def lat64(dev):
    from vlxmi.fock import FockEngine
    mol, D = configs.make_lattice64()
    return mol, D, FockEngine(mol, dev=dev)


def test_lattice4_all_classes(lat4):
    mol, D, eng = lat4
    Dc_np = U.cart_density(mol, D)
    Dc = eng.cart_density(D)
    _say("case: lattice4 all 36 J classes")
    gJ = {k: v.cpu().numpy() for k, v in eng.run_J(Dc).items()}
    rJ = U.oracle_matJ(eng.fl, Dc_np, "component")
    for k in rJ:
        assert np.max(np.abs(gJ[k] - rJ[k])) <= TOL, k
    _say("case: lattice4 all 54 K classes")
    gK = {k: v.cpu().numpy() for k, v in eng.run_K(Dc).items()}
    rK = U.oracle_matK(eng.fl, Dc_np, "component")
    for k in rK:
        assert np.max(np.abs(gK[k] - rK[k])) <= TOL, k
    J, K = eng.build(D)
    Jr, Kr = P.assemble_J(eng.fl, rJ), P.assemble_K(eng.fl, rK)
    T = mol.cart2sph()
    assert np.max(np.abs(J.cpu().numpy() - T @ Jr @ T.T)) <= TOL
    assert np.max(np.abs(K.cpu().numpy() - T @ Kr @ T.T)) <= TOL


def test_lattice4_range_separated_k(lat4, dev):
    from vlxmi.fock import FockEngine
    mol, D, _ = lat4
    _say("case: lattice4 K with omega = 0.4")
    eng = FockEngine(mol, dev=dev, omega=0.4)
    Dc_np = U.cart_density(mol, D)
    gK = {k: v.cpu().numpy() for k, v in eng.run_K(eng.cart_density(D)).items()}
    rK = U.oracle_matK(eng.fl, Dc_np, "component", omega=0.4)
    for k in rK:
        assert np.max(np.abs(gK[k] - rK[k])) <= TOL, k


def test_lattice64_j_sampled(lat64):
    mol, D, eng = lat64
    Dc_np = U.cart_density(mol, D)
    Dc = eng.cart_density(D)
    kets = eng.prepare_kets(Dc)
    rng = np.random.default_rng(5)
    for ab, cd in P.j_classes():
        _say("case: lattice64 J class %s%s (sampled bra blocks)" % (ab, cd))
        g = eng.run_J(Dc, kets=kets, classes=[(ab, cd)])[ab].cpu().numpy()
        n = eng.fl.jlists[ab].nblk
        sel = np.sort(np.concatenate([[0, n - 1], rng.choice(n, size=min(n, 14), replace=False)])).astype(np.uint32)
        sel = np.unique(sel)
        r = U.oracle_matJ(eng.fl, Dc_np, "component", bra_blocks=sel, classes=[(ab, cd)])[ab]
        nab = eng.jl[ab]["nab"]
        idx = (sel[:, None] * nab + np.arange(nab)[None, :]).ravel()
        assert np.max(np.abs(g[idx] - r[idx])) <= TOL, (ab, cd)


@pytest.fixture(scope="module")
def lat64_dens(lat64):
    mol, D, eng = lat64
    Dc = eng.cart_density(D)
    return U.cart_density(mol, D), Dc, eng.max_D(Dc)


@pytest.mark.parametrize("cls", P.k_classes())
def test_lattice64_k_sampled(lat64, lat64_dens, cls):
    """Every exchange class on a slice of (i,k) blocks (the reference's _local
    share) against the full CSR lists, compared with the oracle."""
    mol, D, eng = lat64
    Dc_np, Dc, maxD = lat64_dens
    ac = cls[0] + cls[2]
    ik = eng.fl.iklists[ac]
    rng = np.random.default_rng(600 + P.k_classes().index(cls))
    sel = np.unique(np.concatenate([[0, ik.nblk - 1], rng.choice(ik.nblk, size=min(ik.nblk, 6), replace=False)]))
    sel = sel.astype(np.uint32)
    _say("case: lattice64 K class %s (%d sampled (i,k) blocks)" % (cls, sel.size))
    g = eng.run_K(Dc, classes=[cls], maxD=maxD, ik_blocks={ac: sel})[ac].cpu().numpy()
    r = U.oracle_matK(eng.fl, Dc_np, "component", ik_blocks=sel, classes=[cls])[ac]
    nac = P.NCART[P.LVAL[cls[0]]] * P.NCART[P.LVAL[cls[2]]]
    idx = (sel[:, None].astype(np.int64) * nac + np.arange(nac)[None, :]).ravel()
    assert np.max(np.abs(g - r[idx])) <= TOL


def test_lattice64_full_build_properties(lat64):
    mol, D, eng = lat64
    _say("case: lattice64 full J+K build: symmetry, positivity")
    J, K = eng.build(D)
    J, K = J.cpu().numpy(), K.cpu().numpy()
    assert np.max(np.abs(J - J.T)) <= 1e-10 and np.max(np.abs(K - K.T)) <= 1e-10
    EJ, EK = 0.5 * np.sum(D * J), 0.5 * np.sum(D * K)
    assert EJ > 0 and EK > 0 and EK < EJ


def test_lattice64_fp64_bit_reproducible(lat64):
    """Coulomb launches split their ket lists into chunks (grid.y) and combine
    the chunks' partial sums in a fixed order, so two FP64 builds of the same
    density give the same bits."""
    mol, D, eng = lat64
    _say("case: lattice64 FP64 build twice, bitwise")
    J1, K1 = eng.build(D)
    J2, K2 = eng.build(D)
    assert eng.torch.equal(J1, J2) and eng.torch.equal(K1, K2)
