"""Scheduling, admission and eviction logic of the request server, exercised on
the CPU with a fake device entry and batch runner (no kernels run)."""

import time

import pytest

from vlxmi.serve import FockServer, Request


class FakeEntry:
    def __init__(self, mol):
        self.nbytes = mol["bytes"]


def _runner(entries, densities):
    return [("J", d) for d in densities], [("K", d) for d in densities]


class SlowHandle:
    """A batch that stays 'on the GPU' for ``polls`` readiness checks."""
    live = 0
    most = 0

    def __init__(self, densities, polls):
        self.d = densities
        self.left = polls
        SlowHandle.live += 1
        SlowHandle.most = max(SlowHandle.most, SlowHandle.live)

    def ready(self):
        self.left -= 1
        if self.left <= 0:
            SlowHandle.live -= 1
            return True
        return False

    def result(self):
        return [("J", d) for d in self.d], [("K", d) for d in self.d], [{"n": len(self.d)}] * len(self.d)


def _server(budget, max_batch=4, sizes=(10, 10, 10, 10), **kw):
    mols = {m: {"bytes": b} for m, b in enumerate(sizes)}
    return FockServer(mols, hbm_budget=budget, max_batch=max_batch, keep_results=True, entry_factory=FakeEntry,
                      batch_runner=kw.pop("runner", _runner), estimate_bytes=lambda mol: mol["bytes"], **kw)


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


def test_admission_never_builds_what_does_not_fit():
    """Admission uses the size estimate before building: every entry built is
    kept (no build-and-discard), and the budget is never exceeded."""
    built = []

    def factory(mol):
        built.append(mol["bytes"])
        return FakeEntry(mol)

    mols = {m: {"bytes": 10} for m in range(5)}
    srv = FockServer(mols, hbm_budget=25, max_batch=4, keep_results=True, entry_factory=factory,
                     batch_runner=_runner, estimate_bytes=lambda mol: mol["bytes"])
    peak = [0.0]
    orig = srv._make_room

    def watch(need, keep):
        ok = orig(need, keep)
        peak[0] = max(peak[0], srv.used + sum(srv.reserved.values()) + (need if ok else 0))
        return ok

    srv._make_room = watch
    st = srv.serve(_reqs(5, 3), realtime=False)
    assert st.n == 15 and st.setups == len(built)
    assert peak[0] <= 25


def test_request_that_never_fits_is_an_error():
    srv = _server(budget=5)
    with pytest.raises(RuntimeError):
        srv.serve(_reqs(1, 1), realtime=False)


def test_realtime_latency_accounting():
    srv = _server(budget=1e9)
    reqs = [Request(rid=k, mol_id=k % 2, iteration=k // 2, arrival=0.002 * k, D=k) for k in range(6)]
    st = srv.serve(reqs, realtime=True)
    assert st.n == 6 and st.p99_ms >= st.p50_ms >= 0.0 and st.wall_s >= 0.01


def test_batches_overlap_and_molecule_order_holds():
    """With two batch slots, a second batch is launched while the first is
    still on the GPU; a molecule is never in two batches at once and its
    iterations complete in order; every result comes back with its report."""
    SlowHandle.live = SlowHandle.most = 0
    order = []

    def runner(entries, densities):
        order.append(list(densities))
        return SlowHandle(densities, polls=3)

    srv = _server(budget=1e9, max_batch=2, sizes=(10,) * 6, runner=runner, inflight=2)
    reqs = _reqs(6, 3)
    st = srv.serve(reqs, realtime=False)
    assert st.n == 18 and st.max_inflight == 2 and st.overlapped > 0 and SlowHandle.most == 2
    assert all(r.J == ("J", r.D) and r.report["n"] in (1, 2) for r in reqs)
    done_at = {}
    for r in sorted(reqs, key=lambda q: q.t_done):
        assert done_at.get(r.mol_id, -1) == r.iteration - 1
        done_at[r.mol_id] = r.iteration


def test_setup_runs_off_the_loop_and_is_timed():
    """Molecule set-up runs on the worker thread while batches of resident
    molecules go on; set-up time is reported; an iteration waits for the
    previous one of its molecule even when it is earlier in the queue."""
    def slow_entry(mol):
        time.sleep(0.02 if mol["bytes"] == 11 else 0.0)
        return FakeEntry(mol)

    mols = {0: {"bytes": 10}, 1: {"bytes": 11}}
    srv = FockServer(mols, hbm_budget=1e9, max_batch=4, keep_results=True, entry_factory=slow_entry,
                     batch_runner=_runner, estimate_bytes=lambda mol: mol["bytes"])
    reqs = [Request(rid=0, mol_id=0, iteration=1, arrival=0.0, D=0),
            Request(rid=1, mol_id=0, iteration=0, arrival=0.0, D=1),
            Request(rid=2, mol_id=1, iteration=0, arrival=0.0, D=2)]
    st = srv.serve(reqs, realtime=False)
    assert st.setups == 2 and st.setup_s >= 0.02 and st.pack_s == 0.0
    # molecule 0's two iterations were served before the slow set-up of molecule 1 finished
    first_two = srv.batch_log[0] + srv.batch_log[1]
    assert srv.batch_log[0] == [1] and 0 in first_two
    assert srv.batch_log[-1] == [2]
