"""Mixed precision (FP32 VRR/HRR for quartets whose Schwarz x |D| bound is
below 1e-6; Boys, prefactors and accumulation FP64) against the FP64 oracle.

Bars: |dE_J|, |dE_K| <= 1e-8 Eh (the task's target) and elementwise
max|dJ|, max|dK| <= 1e-8; the plain FP64 path stays within 1e-11."""

import numpy as np
import pytest

import _util as U
from vlxmi import configs
from vlxmi import pairs as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from vlxmi.runtime import Device
    return Device()


@pytest.mark.parametrize("policy", ["all", "measured"])
@pytest.mark.parametrize("wl", ["small8", "lattice4"])
def test_mixed_vs_oracle(dev, wl, policy):
    """policy "all": every class runs its _mixed kernel; "measured": only the
    classes of vlxmi/gen/mixed_classes.json do (the engines' default)."""
    from vlxmi.fock import FockEngine
    print("case: mixed precision %s, policy %s" % (wl, policy), flush=True)
    mol, D = configs.make_workload(wl)
    eng = FockEngine(mol, dev=dev, precision="mixed", mixed_policy=policy)
    J, K = eng.build(D)
    J, K = J.cpu().numpy(), K.cpu().numpy()
    Dc = U.cart_density(mol, D)
    Jr, Kr = U.oracle_fock(eng.fl, Dc, "component")
    T = mol.cart2sph()
    Jr, Kr = T @ Jr @ T.T, T @ Kr @ T.T
    dEJ = 0.5 * np.sum(D * (J - Jr))
    dEK = 0.5 * np.sum(D * (K - Kr))
    print("  max|dJ| %.2e max|dK| %.2e dE_J %.2e dE_K %.2e" % (np.abs(J - Jr).max(), np.abs(K - Kr).max(), dEJ, dEK))
    assert abs(dEJ) <= 1e-8 and abs(dEK) <= 1e-8
    assert np.abs(J - Jr).max() <= 1e-8 and np.abs(K - Kr).max() <= 1e-8
    eng.precision = "fp64"
    J64, K64 = eng.build(D)
    assert np.abs(J64.cpu().numpy() - Jr).max() <= 1e-11 and np.abs(K64.cpu().numpy() - Kr).max() <= 1e-11


def test_mixed_lattice64_energy_bar(dev):
    """lattice64 (nbf 896) full J+K build: the engines' default mixed path
    (FP32 below 1e-7, measured classes) against the plain FP64 path, which the
    parity suite pins to the oracle at 1e-11.  |dE_J|, |dE_K| <= 1e-8 Eh, and
    the quartet counters show that FP32 quartets were evaluated."""
    from vlxmi.fock import FockEngine
    print("case: lattice64 mixed (default threshold) vs FP64, energies", flush=True)
    mol, D = configs.make_lattice64()
    eng = FockEngine(mol, dev=dev)
    Dd = dev.tensor(D)
    J64, K64 = eng.build(Dd)
    eng.precision = "mixed"
    assert eng.fp32_threshold == 1e-7
    J, K = eng.build(Dd)
    n64, n32 = eng.quartet_counts()
    EJ64, EK64 = eng.energies(Dd, J64, K64)
    EJ, EK = eng.energies(Dd, J, K)
    dJ, dK = float((J - J64).abs().max()), float((K - K64).abs().max())
    print("  FP32 quartets %d of %d in mixed-kernel classes; dE_J %.2e dE_K %.2e max|dJ| %.2e max|dK| %.2e"
          % (n32, n32 + n64, EJ - EJ64, EK - EK64, dJ, dK), flush=True)
    assert n32 > 0 and n64 > 0
    assert abs(EJ - EJ64) <= 1e-8 and abs(EK - EK64) <= 1e-8
    assert dJ <= 1e-8 and dK <= 1e-8


def test_build_with_report(dev):
    """A mixed build reports its quartet counts and its errors against the
    plain FP64 build; the reported errors are the ones measured here against
    the oracle-pinned FP64 path."""
    from vlxmi.fock import FockEngine
    print("case: lattice4 mixed build with report", flush=True)
    mol, D = configs.make_lattice4()
    eng = FockEngine(mol, dev=dev, precision="mixed", mixed_policy="all", fp32_threshold=1e-6)
    J, K, rep = eng.build_with_report(D)
    print("  report %s" % rep, flush=True)
    assert rep["precision"] == "mixed" and rep["reference"] == "fp64"
    assert rep["mixed_kernel_quartets"] > 0 and 0 <= rep["fp32_quartets"] <= rep["mixed_kernel_quartets"]
    eng.precision = "fp64"
    J64, K64 = eng.build(D)
    assert rep["max_abs_dJ"] == float((J - J64).abs().max()) and rep["max_abs_dK"] == float((K - K64).abs().max())
    assert abs(rep["dE_J"]) <= 1e-8 and abs(rep["dE_K"]) <= 1e-8
