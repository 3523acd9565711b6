"""Continuous-batching J/K request server on one GPU.

Requests (molecule id, SCF iteration, density) arrive over time.  The server
loop repeatedly

1. moves every request that has arrived into the pending FIFO;
2. forms the next batch from the FIFO: at most ``max_batch`` requests, at most
   one per molecule (a molecule's SCF iterations are served in order), each
   admitted only if its molecule is resident or can be made resident within the
   HBM budget (idle molecules are evicted least-recently-used first);
3. packs the batch's resident molecules and runs one batched launch per class
   (``BatchEngine``); requests that arrive while a batch is on the GPU are
   picked up by the next batch.

Pair data, Schwarz tables and the K CSR lists of a molecule stay resident
across its SCF iterations; only the density changes per request.  Latency is
measured from arrival to the moment the request's J and K are complete on the
device.  Because the batched kernels give each molecule the same work items in
the same order regardless of the other molecules, a request's J/K are
bit-identical alone and in any mix.
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from .batch import BatchEngine, MolEntry, Pack
from .runtime import Device


@dataclass
class Request:
    rid: int
    mol_id: int
    iteration: int
    arrival: float                 # seconds from the start of the run
    D: object                      # density (spherical basis), host array or device tensor
    t_done: float = None
    J: object = None
    K: object = None


@dataclass
class ServeStats:
    n: int
    wall_s: float
    builds_per_s: float
    p50_ms: float
    p99_ms: float
    mean_batch: float
    max_resident: int
    evictions: int
    setups: int
    setup_s: float = 0.0           # making molecules resident (lists, Schwarz, plans)
    pack_s: float = 0.0            # concatenating resident molecules into batches
    build_s: float = 0.0           # density preparation, launches and results (synchronised)
    latencies_ms: list = field(default_factory=list)

    def as_dict(self):
        d = dict(self.__dict__)
        d.pop("latencies_ms")
        return d


class FockServer:
    def __init__(self, molecules, dev: Device = None, hbm_budget=16e9, max_batch=32, precision="mixed",
                 thr=1e-12, fp32_threshold=1e-7, keep_results=False, entry_factory=None, batch_runner=None,
                 synchronize=None, mixed_policy="measured"):
        """``entry_factory(mol) -> entry with .nbytes`` and ``batch_runner(entries,
        densities) -> (Js, Ks)`` default to the GPU MolEntry / BatchEngine; they are
        injectable so the scheduling logic can be exercised without a GPU."""
        self.molecules = molecules                 # mol_id -> Molecule
        self.budget = float(hbm_budget)
        self.max_batch = int(max_batch)
        if entry_factory is None or batch_runner is None:
            self.dev = dev or Device()
            self.engine = BatchEngine(self.dev, thr=thr, precision=precision, fp32_threshold=fp32_threshold,
                                      mixed_policy=mixed_policy)
        else:
            self.dev = dev
            self.engine = None
        self.entry_factory = entry_factory or (lambda mol: MolEntry(mol, self.dev, thr=self.thr))
        self.batch_runner = batch_runner or self._gpu_batch
        self.synchronize = synchronize or (lambda: self.dev.torch.cuda.synchronize())
        self.thr = thr
        self.resident = collections.OrderedDict()  # mol_id -> MolEntry (LRU order)
        self.used = 0
        self.evictions = 0
        self.setups = 0
        self.setup_s = self.pack_s = self.build_s = 0.0
        self.keep_results = keep_results
        self.batch_log = []

    # ------------------------------------------------------------ residency
    def _evict_for(self, need, busy):
        for mid in list(self.resident.keys()):
            if self.used + need <= self.budget:
                return True
            if mid in busy:
                continue
            e = self.resident.pop(mid)
            self.used -= e.nbytes
            self.evictions += 1
            del e
        return self.used + need <= self.budget

    def admit(self, mol_id, busy):
        """Make ``mol_id`` resident if it fits the HBM budget; False if not now."""
        if mol_id in self.resident:
            self.resident.move_to_end(mol_id)
            return True
        t0 = time.perf_counter()
        e = self.entry_factory(self.molecules[mol_id])
        self.setup_s += time.perf_counter() - t0
        self.setups += 1
        if not self._evict_for(e.nbytes, busy):
            del e
            return False
        self.resident[mol_id] = e
        self.used += e.nbytes
        return True

    # ------------------------------------------------------------ one batch
    def _gpu_batch(self, entries, densities):
        t0 = time.perf_counter()
        pack = Pack(entries, self.dev)
        self.pack_s += time.perf_counter() - t0
        return self.engine.build(pack, densities)

    def run_batch(self, reqs):
        entries = [self.resident[r.mol_id] for r in reqs]
        return self.batch_runner(entries, [r.D for r in reqs])

    # ------------------------------------------------------------ loop
    def serve(self, requests, realtime=True, progress=None, progress_every=100):
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` the loop
        follows the wall clock (requests become visible at their arrival time);
        otherwise every request is visible at once (offline batching).
        ``progress(msg)`` is called each time another ``progress_every``
        requests have completed."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = collections.deque()
        nxt = 0
        done = 0
        batches = []
        max_res = 0
        t0 = time.perf_counter()
        while done < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and (not realtime or reqs[nxt].arrival <= now):
                pending.append(reqs[nxt])
                nxt += 1
            if not pending:
                if nxt < len(reqs):
                    time.sleep(max(0.0, reqs[nxt].arrival - now))
                continue
            batch, chosen = [], set()
            busy = set(r.mol_id for r in pending)
            first_it = {}
            for q in pending:
                if q.iteration < first_it.get(q.mol_id, q.iteration + 1):
                    first_it[q.mol_id] = q.iteration
            for r in list(pending):
                if len(batch) >= self.max_batch:
                    break
                if r.mol_id in chosen:
                    continue
                # a molecule's earlier iteration still pending keeps order
                if r.iteration > first_it[r.mol_id]:
                    continue
                if not self.admit(r.mol_id, busy | chosen):
                    continue
                batch.append(r)
                chosen.add(r.mol_id)
            if not batch:
                raise RuntimeError("no pending request fits the HBM budget (%.3g bytes)" % self.budget)
            for r in batch:
                pending.remove(r)
            tb = time.perf_counter()
            Js, Ks = self.run_batch(batch)
            self.synchronize()
            self.build_s += time.perf_counter() - tb
            t_done = time.perf_counter() - t0
            for r, J, K in zip(batch, Js, Ks):
                r.t_done = t_done
                if self.keep_results:
                    r.J, r.K = (J.clone(), K.clone()) if hasattr(J, "clone") else (J, K)
            if progress is not None and (done + len(batch)) // progress_every > done // progress_every:
                progress("served %d of %d requests" % (done + len(batch), len(reqs)))
            done += len(batch)
            batches.append(len(batch))
            self.batch_log.append([r.rid for r in batch])
            max_res = max(max_res, len(self.resident))
        wall = time.perf_counter() - t0
        lat = np.array([(r.t_done - r.arrival) * 1e3 for r in reqs])
        return ServeStats(n=len(reqs), wall_s=wall, builds_per_s=len(reqs) / wall,
                          p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                          mean_batch=float(np.mean(batches)), max_resident=max_res,
                          evictions=self.evictions, setups=self.setups, setup_s=self.setup_s,
                          pack_s=self.pack_s, build_s=self.build_s - self.pack_s, latencies_ms=lat.tolist())


def poisson_requests(mol_densities, iterations, rate, seed=0, perturb=None):
    """Requests for molecules m = 0..M-1, each repeated ``iterations`` times,
    interleaved round-robin (request k is molecule k mod M, iteration k // M),
    with exponential inter-arrival times of mean 1/rate.  ``perturb(D0, m, it)``
    gives iteration it's density."""
    rng = np.random.default_rng(seed)
    M = len(mol_densities)
    n = M * iterations
    gaps = rng.exponential(1.0 / rate, size=n) if rate > 0 else np.zeros(n)
    t = np.cumsum(gaps) - gaps[0]
    out = []
    for k in range(n):
        m, it = k % M, k // M
        D0 = mol_densities[m]
        D = perturb(D0, m, it) if perturb else D0
        out.append(Request(rid=k, mol_id=m, iteration=it, arrival=float(t[k]), D=D))
    return out
