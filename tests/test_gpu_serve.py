"""Continuous-batching server on the GPU.

* Bursts of arrivals with two batches in flight: requests admitted while
  another batch is on the GPU, and each request's J/K equal (bitwise) to the
  same request served alone.
* The serving workload's generator (molecules of 8-200 atoms at the 8-atom
  workload's number density, 2 s + 1 p per atom, densities perturbed per SCF
  iteration) on a slice with a 128-atom molecule: every mixed result's report
  gives its errors against the FP64 build, each within 1e-8 Eh in E_J and E_K;
  every result is checked against the plain FP64 single-molecule path
  (1e-8 Eh), and the small molecules against the FP64 oracle.
"""

import numpy as np
import pytest

import _util as U
from vlxmi import configs

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def dev():
    from vlxmi.runtime import Device
    return Device()


def test_served_equals_alone(dev):
    from vlxmi.batch import BatchEngine, MolEntry, Pack
    from vlxmi.serve import FockServer, Request
    _say("case: serving 6 molecules x 3 SCF iterations in bursts, two batches in flight")
    mols, dens = {}, []
    for m in range(6):
        mol, D = configs.make_batch_molecule(m, lo=4, hi=14)
        mols[m] = mol
        dens.append(D)
    # every molecule's iteration k arrives together at 30 ms * k: each burst is
    # more than one batch (max_batch 3), so batches mix molecules and the
    # second batch of a burst is launched while the first is on the GPU
    reqs = [Request(rid=len(mols) * it + m, mol_id=m, iteration=it, arrival=0.03 * it,
                    D=configs.serve_iteration_density(dens[m], m, it)) for it in range(3) for m in range(len(mols))]
    srv = FockServer(mols, dev=dev, hbm_budget=4e9, max_batch=3, precision="mixed", keep_results=True)
    st = srv.serve(reqs)
    _say("  batches %d, max in flight %d, launched while another ran %d"
         % (len(srv.batch_log), st.max_inflight, st.overlapped))
    assert st.n == 18 and st.p99_ms >= st.p50_ms > 0
    assert max(len(b) for b in srv.batch_log) > 1          # requests were actually mixed
    assert st.overlapped > 0                               # admitted while another batch was on the GPU
    be = BatchEngine(dev, precision="mixed")
    for r in reqs:
        e = MolEntry(mols[r.mol_id], dev)
        Ja, Ka = be.build(Pack([e], dev), [r.D])
        assert bool((Ja[0] == r.J).all()) and bool((Ka[0] == r.K).all()), r.rid
        assert 0.0 <= r.report["fp32_quartet_share_of_batch"] <= 1.0


def test_serving_slice_with_large_molecule(dev):
    from vlxmi.fock import FockEngine
    from vlxmi.serve import FockServer, poisson_requests
    ids = [0, 4, 17, 1, 6]                     # 9, 10, 8, 16 and 128 atoms
    mols, dens = {}, []
    for k, m in enumerate(ids):
        mol, D = configs.make_serve_molecule(m)
        mols[k] = mol
        dens.append(D)
    assert max(mol.natoms for mol in mols.values()) >= 100
    _say("case: serving slice, molecules of %s atoms x 2 iterations, mixed, verified against FP64"
         % [mols[k].natoms for k in range(len(ids))])
    reqs = poisson_requests(dens, 2, rate=50.0, seed=3,
                            perturb=lambda D0, m, it: configs.serve_iteration_density(D0, ids[m], it))
    srv = FockServer(mols, dev=dev, hbm_budget=16e9, max_batch=4, precision="mixed", keep_results=True,
                     verify="fp64")
    st = srv.serve(reqs)
    assert st.n == 10
    engs = {}
    for r in reqs:
        rep = r.report
        _say("  request %d (%d atoms): dE_J %.2e dE_K %.2e max|dJ| %.2e max|dK| %.2e vs FP64"
             % (r.rid, mols[r.mol_id].natoms, rep["dE_J"], rep["dE_K"], rep["max_abs_dJ"], rep["max_abs_dK"]))
        assert abs(rep["dE_J"]) <= 1e-8 and abs(rep["dE_K"]) <= 1e-8
        if r.mol_id not in engs:
            engs[r.mol_id] = FockEngine(mols[r.mol_id], dev=dev)
        eng = engs[r.mol_id]
        J64, K64 = eng.build(r.D)
        Dd = dev.tensor(r.D)
        EJ, EK = eng.energies(Dd, r.J, r.K)
        EJ64, EK64 = eng.energies(Dd, J64, K64)
        assert abs(EJ - EJ64) <= 1e-8 and abs(EK - EK64) <= 1e-8, r.rid
        if mols[r.mol_id].natoms <= 10:
            Dc = U.cart_density(mols[r.mol_id], r.D)
            Jr, Kr = U.oracle_fock(eng.fl, Dc, "component")
            T = mols[r.mol_id].cart2sph()
            Jr, Kr = T @ Jr @ T.T, T @ Kr @ T.T
            EJr, EKr = 0.5 * np.sum(r.D * Jr), 0.5 * np.sum(r.D * Kr)
            assert abs(EJ - EJr) <= 1e-8 and abs(EK - EKr) <= 1e-8, r.rid
            assert np.abs(r.J.cpu().numpy() - Jr).max() <= 1e-8 and np.abs(r.K.cpu().numpy() - Kr).max() <= 1e-8
