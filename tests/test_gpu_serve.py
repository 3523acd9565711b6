"""Continuous-batching server on the GPU: Poisson arrivals, residency, and each
request's J/K equal (bitwise) to the same request served alone."""

import numpy as np
import pytest

from vlxmi import configs

pytestmark = pytest.mark.gpu


def test_served_equals_alone(capsys):
    from vlxmi.batch import BatchEngine, MolEntry, Pack
    from vlxmi.runtime import Device
    from vlxmi.serve import FockServer, poisson_requests
    print("case: serving 4 molecules x 3 SCF iterations, Poisson arrivals", flush=True)
    dev = Device()
    mols, dens = {}, []
    for m in range(4):
        mol, D = configs.make_batch_molecule(m, lo=4, hi=10)
        mols[m] = mol
        dens.append(D)
    reqs = poisson_requests(dens, 3, rate=200.0, seed=1,
                            perturb=lambda D0, m, it: configs.serve_iteration_density(D0, m, it))
    srv = FockServer(mols, dev=dev, hbm_budget=4e9, max_batch=8, precision="mixed", keep_results=True)
    st = srv.serve(reqs)
    assert st.n == 12 and st.p99_ms >= st.p50_ms > 0
    assert max(len(b) for b in srv.batch_log) > 1          # requests were actually mixed
    be = BatchEngine(dev, precision="mixed")
    for r in reqs:
        e = MolEntry(mols[r.mol_id], dev)
        Ja, Ka = be.build(Pack([e], dev), [r.D])
        assert bool((Ja[0] == r.J).all()) and bool((Ka[0] == r.K).all()), r.rid
