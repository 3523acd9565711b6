"""The mixed-precision configuration at full size: lattice256 (256 atoms on an
8x8x4 lattice, 2.5 bohr, per atom 3 s + 2 p + 1 spherical d, nbf 3584, 3840
Cartesian AOs), FP32 VRR/HRR for quartets whose Schwarz x |D| bound is below
1e-6.

Every Coulomb class runs on a sample of bra blocks against the full ket list,
and every exchange class on a sample of (i,k) blocks against the full CSR
lists: the reference's ``_local`` convention (gen_coulomb.py:81-86), so each
launch is small while the lists are the full-size ones.  The oracle evaluates
the same blocks.  Bars:

* plain FP64 kernels: 1e-11 max-abs per element against the oracle;
* mixed kernels (every class, FP32 below 1e-6): the per-element error against
  the oracle is printed per class and must stay below the FP32 threshold
  itself; the quartet counters must show FP32 quartets in every class, and
  the mixed result must differ from the FP64 one (the FP32 loop ran).
"""

import numpy as np
import pytest

import _util as U
from vlxmi import configs
from vlxmi import pairs as P

TOL = 1e-11
FP32_THRESHOLD = 1e-6
pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _sample(n, k, rng):
    return np.unique(np.concatenate([[0, n - 1], rng.choice(n, size=min(n, k), replace=False)])).astype(np.uint32)


@pytest.fixture(scope="module")
def lat256():
    from vlxmi.fock import FockEngine
    from vlxmi.runtime import Device
    _say("case: lattice256 set-up (lists, Schwarz factors, density)")
    mol, D = configs.make_lattice256()
    assert (mol.natoms, mol.nbf, mol.naos_cart) == (256, 3584, 3840)
    eng = FockEngine(mol, dev=Device(), fp32_threshold=FP32_THRESHOLD, mixed_policy="all")
    Dc_np = U.cart_density(mol, D)
    Dc = eng.cart_density(D)
    kets = eng.prepare_kets(Dc)
    maxD = eng.max_D(Dc)
    _say("case: lattice256 ready (%.1f s of set-up)" % eng.setup_s)
    return eng, Dc_np, Dc, kets, maxD


def _both_precisions(eng, run):
    eng.precision = "fp64"
    g64 = run()
    eng.precision = "mixed"
    eng.reset_counts()
    gm = run()
    n64, n32 = eng.quartet_counts()
    eng.precision = "fp64"
    return g64, gm, n64, n32


def test_lattice256_coulomb_sampled(lat256):
    eng, Dc_np, Dc, kets, _ = lat256
    rng = np.random.default_rng(11)
    worst = 0.0
    for ab, cd in P.j_classes():
        sel = _sample(eng.fl.jlists[ab].nblk, 6, rng)
        run = lambda: eng.run_J(Dc, kets=kets, classes=[(ab, cd)], bra_blocks={ab: sel})[ab].cpu().numpy()
        g64, gm, n64, n32 = _both_precisions(eng, run)
        r = U.oracle_matJ(eng.fl, Dc_np, "component", bra_blocks=sel, classes=[(ab, cd)])[ab]
        nab = eng.jl[ab]["nab"]
        r = r[(sel[:, None].astype(np.int64) * nab + np.arange(nab)[None, :]).ravel()]
        e64, em = float(np.abs(g64 - r).max()), float(np.abs(gm - r).max())
        worst = max(worst, em)
        _say("case: lattice256 J %s%s: %d bra blocks, FP64 max|dJ| %.2e, mixed max|dJ| %.2e, "
             "quartets FP64 %d FP32 %d" % (ab, cd, sel.size, e64, em, n64, n32))
        assert e64 <= TOL, (ab, cd, e64)
        assert em <= FP32_THRESHOLD, (ab, cd, em)
        assert n32 > 0, (ab, cd)
        assert not np.array_equal(g64, gm), (ab, cd)
    _say("lattice256 J: largest mixed error over all classes %.2e" % worst)


def test_lattice256_exchange_sampled(lat256):
    eng, Dc_np, Dc, _, maxD = lat256
    rng = np.random.default_rng(12)
    worst = 0.0
    for cls in P.k_classes():
        ac = cls[0] + cls[2]
        sel = _sample(eng.fl.iklists[ac].nblk, 3, rng)
        run = lambda: eng.run_K(Dc, classes=[cls], maxD=maxD, ik_blocks={ac: sel})[ac].cpu().numpy()
        g64, gm, n64, n32 = _both_precisions(eng, run)
        r = U.oracle_matK(eng.fl, Dc_np, "component", ik_blocks=sel, classes=[cls])[ac]
        nac = P.NCART[P.LVAL[cls[0]]] * P.NCART[P.LVAL[cls[2]]]
        r = r[(sel[:, None].astype(np.int64) * nac + np.arange(nac)[None, :]).ravel()]
        e64, em = float(np.abs(g64 - r).max()), float(np.abs(gm - r).max())
        worst = max(worst, em)
        _say("case: lattice256 K %s: %d (i,k) blocks, FP64 max|dK| %.2e, mixed max|dK| %.2e, "
             "quartets FP64 %d FP32 %d" % (cls, sel.size, e64, em, n64, n32))
        assert e64 <= TOL, (cls, e64)
        assert em <= FP32_THRESHOLD, (cls, em)
        assert n32 > 0, cls
        assert not np.array_equal(g64, gm), cls
    _say("lattice256 K: largest mixed error over all classes %.2e" % worst)
