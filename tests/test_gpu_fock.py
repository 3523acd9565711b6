"""The Fock matrix on the device: F = H + c_J J + c_K K with the Coulomb and
exchange per-pair outputs assembled in one deterministic pass.

small8: against the oracle's 2J - K (+ H) to 1e-11; lattice64 (all classes
through (dd|dd)): against c_J J + c_K K from the separate J and K assemblies
(same sums, other association) to 1e-10, bitwise reproducible."""

import numpy as np
import pytest

import _util as U
from vlxmi import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from vlxmi.runtime import Device
    return Device()


def test_fock_small8_vs_oracle(dev):
    from vlxmi.fock import FockEngine
    print("case: Fock matrix small8 vs oracle", flush=True)
    mol, D = configs.make_small8()
    eng = FockEngine(mol, dev=dev)
    rng = np.random.default_rng(4)
    H = rng.standard_normal(D.shape)
    H = H + H.T
    F = eng.fock(D, H=H).cpu().numpy()
    Dc = U.cart_density(mol, D)
    Jr, Kr = U.oracle_fock(eng.fl, Dc, "component")
    T = mol.cart2sph()
    Fr = H + T @ (2.0 * Jr - Kr) @ T.T
    assert np.abs(F - Fr).max() <= 1e-11
    G = eng.fock(D, cJ=1.0, cK=0.0).cpu().numpy()       # c_K = 0: J alone
    assert np.abs(G - T @ Jr @ T.T).max() <= 1e-11


def test_fock_lattice64_vs_separate_assembly(dev):
    from vlxmi.fock import FockEngine
    print("case: Fock matrix lattice64 vs J, K assembled separately", flush=True)
    mol, D = configs.make_lattice64()
    eng = FockEngine(mol, dev=dev)
    F1 = eng.fock(D)
    F2 = eng.fock(D)
    J, K = eng.build(D)
    assert eng.torch.equal(F1, F2)
    ref = 2.0 * J - K
    assert float((F1 - ref).abs().max()) <= 1e-10 * max(1.0, float(ref.abs().max()))
