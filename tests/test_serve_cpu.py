"""Scheduling, admission and eviction logic of the request server, exercised on
the CPU with a fake device entry and batch runner (no kernels run)."""

import pytest

from vlxmi.serve import FockServer, Request


class FakeEntry:
    def __init__(self, mol):
        self.nbytes = mol["bytes"]


def _runner(entries, densities):
    return [("J", d) for d in densities], [("K", d) for d in densities]


def _server(budget, max_batch=4, sizes=(10, 10, 10, 10)):
    mols = {m: {"bytes": b} for m, b in enumerate(sizes)}
    return FockServer(mols, hbm_budget=budget, max_batch=max_batch, keep_results=True,
                      entry_factory=FakeEntry, batch_runner=_runner, synchronize=lambda: None)


def _reqs(nmol, iters):
    return [Request(rid=k, mol_id=k % nmol, iteration=k // nmol, arrival=0.0, D=k) for k in range(nmol * iters)]


def test_one_request_per_molecule_per_batch_in_iteration_order():
    srv = _server(budget=1e9)
    reqs = _reqs(3, 4)
    st = srv.serve(reqs, realtime=False)
    assert st.n == 12
    seen = {}
    for batch in srv.batch_log:
        mols = [reqs[r].mol_id for r in batch]
        assert len(mols) == len(set(mols))
        for r in batch:
            m, it = reqs[r].mol_id, reqs[r].iteration
            assert seen.get(m, -1) == it - 1
            seen[m] = it
    assert all(r.J == ("J", r.D) for r in reqs)


def test_budget_admission_and_lru_eviction():
    srv = _server(budget=25, max_batch=4)          # two molecules fit at a time
    reqs = _reqs(4, 2)
    st = srv.serve(reqs, realtime=False)
    assert st.n == 8 and st.max_resident <= 2 and st.evictions > 0
    for batch in srv.batch_log:
        assert len(batch) <= 2


def test_request_that_never_fits_is_an_error():
    srv = _server(budget=5)
    with pytest.raises(RuntimeError):
        srv.serve(_reqs(1, 1), realtime=False)


def test_realtime_latency_accounting():
    srv = _server(budget=1e9)
    reqs = [Request(rid=k, mol_id=k % 2, iteration=k // 2, arrival=0.002 * k, D=k) for k in range(6)]
    st = srv.serve(reqs, realtime=True)
    assert st.n == 6 and st.p99_ms >= st.p50_ms >= 0.0 and st.wall_s >= 0.01


def test_time_breakdown_and_batch_order_rule():
    """Setup, packing and build times are reported; a molecule whose earlier
    iteration is still pending is held back even when it arrives first in the
    queue order (requests given out of iteration order)."""
    import time as _t

    def slow_entry(mol):
        _t.sleep(0.01)
        return FakeEntry(mol)

    mols = {m: {"bytes": 10} for m in range(2)}
    srv = FockServer(mols, hbm_budget=1e9, max_batch=4, keep_results=True, entry_factory=slow_entry,
                     batch_runner=_runner, synchronize=lambda: None)
    reqs = [Request(rid=0, mol_id=0, iteration=1, arrival=0.0, D=0),
            Request(rid=1, mol_id=0, iteration=0, arrival=0.0, D=1),
            Request(rid=2, mol_id=1, iteration=0, arrival=0.0, D=2)]
    st = srv.serve(reqs, realtime=False)
    assert st.setups == 2 and st.setup_s >= 0.02 and st.pack_s == 0.0 and st.build_s >= 0.0
    first = srv.batch_log[0]
    assert 0 not in first and 1 in first and 2 in first        # iteration 1 waits for iteration 0
    assert srv.batch_log[1] == [0]
