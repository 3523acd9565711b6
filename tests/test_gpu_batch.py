"""Batched persistent launches: many molecules per class launch.

* each molecule's J/K from a pack equal the single-molecule drop-in path
  (1e-11) and, for the first molecules, the FP64 oracle;
* a molecule gives the same bits alone and in the mix;
* a captured HIP graph replayed several times (with eager work and syncs in
  between) reproduces the eager result every time;
* the exchange chain on its own stream beside the Coulomb chain gives the same
  bits as one stream, eagerly and in graph replays;
* the autotuner returns a configuration per class and results stay identical.
"""

import numpy as np
import pytest

import _util as U
from vlxmi import configs

pytestmark = pytest.mark.gpu
TOL = 1e-11


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def dev():
    from vlxmi.runtime import Device
    return Device()


@pytest.fixture(scope="module")
def mols():
    out = [configs.make_batch_molecule(m, lo=4, hi=12) for m in range(5)]
    out.append(configs.make_lattice4())            # one molecule with d shells
    return out


@pytest.fixture(scope="module")
def pack(dev, mols):
    from vlxmi.batch import MolEntry, Pack
    entries = [MolEntry(mol, dev) for mol, _ in mols]
    return entries, Pack(entries, dev)


def test_batched_equals_single_and_oracle(dev, mols, pack):
    from vlxmi.batch import BatchEngine
    from vlxmi.fock import FockEngine
    entries, pk = pack
    _say("case: pack of %d molecules vs single-molecule path and oracle" % len(entries))
    be = BatchEngine(dev)
    Js, Ks = be.build(pk, [D for _, D in mols])
    for m, (mol, D) in enumerate(mols):
        eng = FockEngine(mol, dev=dev)
        J1, K1 = eng.build(D)
        assert (Js[m] - J1).abs().max().item() <= TOL, m
        assert (Ks[m] - K1).abs().max().item() <= TOL, m
        if m in (0, len(mols) - 1):
            Dc = U.cart_density(mol, D)
            Jr, Kr = U.oracle_fock(entries[m].fl, Dc, "component")
            T = mol.cart2sph()
            assert np.abs(Js[m].cpu().numpy() - T @ Jr @ T.T).max() <= TOL
            assert np.abs(Ks[m].cpu().numpy() - T @ Kr @ T.T).max() <= TOL


def test_alone_equals_in_mix_bitwise(dev, mols, pack):
    from vlxmi.batch import BatchEngine, Pack
    entries, pk = pack
    _say("case: molecule alone vs in the mix (bitwise)")
    be = BatchEngine(dev)
    Js, Ks = be.build(pk, [D for _, D in mols])
    for m in (1, len(mols) - 1):
        alone = Pack([entries[m]], dev)
        J1, K1 = be.build(alone, [mols[m][1]])
        assert bool((J1[0] == Js[m]).all()) and bool((K1[0] == Ks[m]).all()), m


def test_graph_replay_matches_eager(dev, mols, pack):
    from vlxmi.batch import BatchEngine
    entries, pk = pack
    _say("case: HIP graph capture + 3 replays vs eager")
    be = BatchEngine(dev)
    dens = [D for _, D in mols]
    replay = be.capture(pk, dens)
    rng = np.random.default_rng(3)
    for it in range(3):
        d2 = [D + 1e-3 * it * (lambda E: E + E.T)(rng.normal(size=D.shape)) for D in dens]
        Je, Ke = be.build(pk, d2)                    # eager work between replays
        dev.torch.cuda.synchronize()
        Jg, Kg = replay(d2)
        dev.torch.cuda.synchronize()
        for m in range(len(dens)):
            assert (Jg[m] - Je[m]).abs().max().item() <= TOL
            assert (Kg[m] - Ke[m]).abs().max().item() <= TOL


def test_concurrent_chains_bitwise(dev, mols, pack):
    from vlxmi.batch import BatchEngine
    entries, pk = pack
    _say("case: exchange launches on a second stream vs one stream, eager and 3 graph replays (bitwise)")
    dens = [D for _, D in mols]
    seq = BatchEngine(dev)
    seq.concurrent = False
    Js, Ks = seq.build(pk, dens)
    con = BatchEngine(dev)
    assert con.concurrent
    Jc, Kc = con.build(pk, dens)
    for m in range(len(dens)):
        assert bool((Jc[m] == Js[m]).all()) and bool((Kc[m] == Ks[m]).all()), m
    replay = con.capture(pk, dens)
    for it in range(3):
        Je, Ke = seq.build(pk, dens)                 # eager work between replays
        dev.torch.cuda.synchronize()
        Jg, Kg = replay(dens)
        dev.torch.cuda.synchronize()
        for m in range(len(dens)):
            assert bool((Jg[m] == Js[m]).all()) and bool((Kg[m] == Ks[m]).all()), (it, m)


def test_autotune_keeps_results(dev, mols, pack):
    from vlxmi.batch import BatchEngine
    entries, pk = pack
    _say("case: autotuner")
    be = BatchEngine(dev)
    dens = [D for _, D in mols]
    J0, K0 = be.build(pk, dens)
    summary = be.autotune(pk, dens, reps=1)
    assert len(summary) > 0
    J1, K1 = be.build(pk, dens)
    for m in range(len(dens)):
        assert (J1[m] - J0[m]).abs().max().item() <= TOL and (K1[m] - K0[m]).abs().max().item() <= TOL
    pk.tuning = {}


@pytest.mark.parametrize("policy", ["all", "measured"])
def test_batched_mixed_error(dev, mols, pack, policy):
    from vlxmi.batch import BatchEngine
    entries, pk = pack
    _say("case: batched mixed precision vs FP64, policy %s" % policy)
    dens = [D for _, D in mols]
    J64, K64 = BatchEngine(dev).build(pk, dens)
    Jm, Km = BatchEngine(dev, precision="mixed", mixed_policy=policy).build(pk, dens)
    for m, D in enumerate(dens):
        Dt = dev.tensor(D)
        dEJ = 0.5 * float((Dt * (Jm[m] - J64[m])).sum().item())
        dEK = 0.5 * float((Dt * (Km[m] - K64[m])).sum().item())
        assert abs(dEJ) <= 1e-8 and abs(dEK) <= 1e-8, (m, dEJ, dEK)


def test_quartet_counters_match_host_counts(dev, mols, pack):
    """The batched kernels count the primitive quartets they evaluate; the count
    must be the N_q of the FLOP model (quartets passing the block screen),
    counted independently on the host from the same lists and densities."""
    from vlxmi import flops
    from vlxmi import pairs as P
    from vlxmi.batch import BatchEngine
    entries, pk = pack
    _say("case: batched quartet counters vs host N_q")
    be = BatchEngine(dev)
    be.build(pk, [D for _, D in mols])
    got = be.quartet_counts(pk)
    want = {}
    for e, (mol, D) in zip(entries, mols):
        Dc = U.cart_density(mol, D)
        for ab, cd in P.j_classes():
            if ab in e.fl.jlists and cd in e.fl.jlists:
                want["J:" + ab + cd] = want.get("J:" + ab + cd, 0) + flops.count_j(e.fl, Dc, ab, cd)
        for cls in P.k_classes():
            mx = P.max_abs_D_block(Dc, e.fl.prims, P.LVAL[cls[1]], P.LVAL[cls[3]])
            want["K:" + cls] = want.get("K:" + cls, 0) + flops.count_k(e.fl, Dc, cls, mx)
    for name, n in want.items():
        assert got.get(name, (0, 0)) == (n, 0), (name, got.get(name), n)


def test_batch_workload_large_molecules(dev):
    """Molecules of the 512-molecule batch workload up to its largest (47
    atoms, nbf 235) packed with small ones: each molecule's J/K from the pack
    against the single-molecule FP64 path (1e-11) and the size-independent
    properties (J, K symmetric; E_J > E_K > 0)."""
    from vlxmi.batch import BatchEngine, MolEntry, Pack
    from vlxmi.fock import FockEngine
    ids = [207, 5, 22, 0, 1, 2]
    mols = [configs.make_batch_molecule(m) for m in ids]
    _say("case: batch workload molecules %s (atoms %s) packed vs single path"
         % (ids, [m.natoms for m, _ in mols]))
    assert max(m.natoms for m, _ in mols) >= 47
    entries = [MolEntry(mol, dev) for mol, _ in mols]
    be = BatchEngine(dev)
    Js, Ks = be.build(Pack(entries, dev), [D for _, D in mols])
    for k, (mol, D) in enumerate(mols):
        J1, K1 = FockEngine(mol, dev=dev, lists=entries[k].fl).build(D)
        assert (Js[k] - J1).abs().max().item() <= TOL, ids[k]
        assert (Ks[k] - K1).abs().max().item() <= TOL, ids[k]
        J, K = Js[k].cpu().numpy(), Ks[k].cpu().numpy()
        assert np.abs(J - J.T).max() <= 1e-11 and np.abs(K - K.T).max() <= 1e-11
        EJ, EK = 0.5 * np.sum(D * J), 0.5 * np.sum(D * K)
        assert EJ > EK > 0
