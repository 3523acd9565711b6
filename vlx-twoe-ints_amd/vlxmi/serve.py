"""Continuous-batching J/K request server on one GPU.

Requests (molecule id, SCF iteration, density) arrive over time.  The server
loop repeatedly

1. moves every request that has arrived into the pending list;
2. starts making the molecules of pending requests resident on a set-up worker
   thread (pair lists, Schwarz factors, assembly plans: ``MolEntry``), each
   admitted only when its estimated size fits the HBM budget beside the
   resident molecules and the set-ups under way (idle molecules are evicted
   least-recently-used first, and only when that frees enough room);
3. whenever fewer than ``inflight`` batches are on the GPU, forms the next
   batch from the pending requests whose molecule is resident and not in a
   batch on the GPU (at most ``max_batch`` requests, one per molecule, a
   molecule's SCF iterations in order), packs it and launches it on a stream
   of its own (one persistent launch per class, ``BatchEngine``) without
   waiting for the batches already on the GPU: their kernels and this batch's
   host preparation and kernels overlap;
4. completes every batch whose GPU work has finished (an event on its stream).

Pair data, Schwarz tables and the K CSR lists of a molecule stay resident
across its SCF iterations; only the density changes per request.  Latency is
measured from arrival to the moment the request's J and K are complete on the
device.  Because the batched kernels give each molecule the same work items in
the same order regardless of the other molecules, a request's J/K are
bit-identical alone and in any mix.

Each result carries a report: the share of the batch's primitive quartets the
mixed kernels evaluated in FP32 and, with ``verify="fp64"``, its errors against
the plain FP64 build of the same request (dE_J, dE_K, max|dJ|, max|dK|; the
FP64 build runs on the same stream right after the mixed one, so it doubles
the GPU work: a validation mode, not a serving mode).
"""

from __future__ import annotations

import collections
import concurrent.futures as cf
import threading
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
    report: dict = None


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
    setup_s: float = 0.0           # making molecules resident (worker thread, lists, Schwarz, plans)
    pack_s: float = 0.0            # concatenating resident molecules into batches
    build_s: float = 0.0           # host time forming and launching batches (preparation included)
    max_inflight: int = 0          # most batches on the GPU at once
    overlapped: int = 0            # batches launched while another batch was on the GPU
    latencies_ms: list = field(default_factory=list)

    def as_dict(self):
        d = dict(self.__dict__)
        d.pop("latencies_ms")
        return d


class _Done:
    """Handle of a batch that finished when it was launched (host runners)."""

    def __init__(self, result):
        self._r = result

    def ready(self):
        return True

    def result(self):
        Js, Ks = self._r[0], self._r[1]
        reps = self._r[2] if len(self._r) > 2 else [{} for _ in Js]
        return Js, Ks, reps


class _GpuBatch:
    """One batch on its own stream: pack, prepare and launch on entry, an event
    marks the end of its GPU work."""

    def __init__(self, srv, entries, densities, stream):
        t = srv.dev.torch
        self.stream = stream
        with t.cuda.stream(stream):
            t0 = time.perf_counter()
            self.pack = Pack(entries, srv.dev)
            srv.pack_s += time.perf_counter() - t0
            self.Js, self.Ks = srv.engine.build(self.pack, densities)
            self.qcount = self.pack.qcount.clone()
            self.err = None
            if srv.verify == "fp64":
                J64, K64 = srv.engine64.build(self.pack, densities)
                rows = []
                for D, J, K, Jf, Kf in zip(densities, self.Js, self.Ks, J64, K64):
                    Dd = D if t.is_tensor(D) else srv.dev.tensor(np.asarray(D, dtype=np.float64))
                    rows.append(t.stack([0.5 * (Dd * (J - Jf)).sum(), 0.5 * (Dd * (K - Kf)).sum(),
                                         (J - Jf).abs().max(), (K - Kf).abs().max()]))
                self.err = t.stack(rows)
            self.event = t.cuda.Event()
            self.event.record(stream)

    def ready(self):
        return self.event.query()

    def result(self):
        self.event.synchronize()
        c = self.qcount.sum(dim=0).cpu().tolist()
        share = c[1] / max(c[0] + c[1], 1)
        reps = [{"fp32_quartet_share_of_batch": share} for _ in self.Js]
        if self.err is not None:
            for rep, row in zip(reps, self.err.cpu().tolist()):
                rep.update(dE_J=row[0], dE_K=row[1], max_abs_dJ=row[2], max_abs_dK=row[3], reference="fp64")
        return self.Js, self.Ks, reps


class FockServer:
    def __init__(self, molecules, dev: Device = None, hbm_budget=16e9, max_batch=32, precision="mixed",
                 thr=1e-12, fp32_threshold=1e-7, keep_results=False, entry_factory=None, batch_runner=None,
                 estimate_bytes=None, mixed_policy="measured", inflight=2, verify=None):
        """``entry_factory(mol) -> entry with .nbytes`` (run on the set-up worker
        thread), ``estimate_bytes(mol)`` (an upper bound used for admission
        before the entry exists) and ``batch_runner(entries, densities)`` (a
        handle with ready() / result() -> (Js, Ks, reports), or a finished
        (Js, Ks) tuple) default to the GPU MolEntry / BatchEngine; they are
        injectable so the scheduling logic can be exercised without a GPU.
        ``verify="fp64"`` also builds every batch in FP64 and reports each
        result's errors against it."""
        self.molecules = molecules                 # mol_id -> Molecule
        self.budget = float(hbm_budget)
        self.max_batch = int(max_batch)
        self.inflight = max(1, int(inflight))
        self.thr = thr
        self.verify = verify
        gpu = entry_factory is None or batch_runner is None
        if gpu:
            self.dev = dev or Device()
            t = self.dev.torch
            self.engine = BatchEngine(self.dev, thr=thr, precision=precision, fp32_threshold=fp32_threshold,
                                      mixed_policy=mixed_policy)
            self.engine64 = BatchEngine(self.dev, thr=thr, precision="fp64") if verify == "fp64" else None
            # default-priority streams: one per batch slot, one for molecule set-up
            self.streams = [t.cuda.Stream(device=self.dev.device) for _ in range(self.inflight)]
            self.setup_stream = t.cuda.Stream(device=self.dev.device)
        else:
            self.dev = dev
            self.engine = self.engine64 = None
            self.streams = [None] * self.inflight
            self.setup_stream = None
        self.entry_factory = entry_factory or self._gpu_entry
        self.batch_runner = batch_runner
        self.estimate_bytes = estimate_bytes or (lambda mol: MolEntry.estimate_nbytes(mol))
        self.resident = collections.OrderedDict()  # mol_id -> MolEntry (LRU order)
        self.used = 0.0                            # bytes of resident molecules
        self.reserved = {}                         # mol_id -> estimated bytes of a set-up under way
        self.evictions = 0
        self.setups = 0
        self.setup_s = self.pack_s = self.build_s = 0.0
        self._setup_lock = threading.Lock()
        self.keep_results = keep_results
        self.batch_log = []

    # ------------------------------------------------------------ residency
    def _gpu_entry(self, mol):
        t = self.dev.torch
        with t.cuda.stream(self.setup_stream):
            e = MolEntry(mol, self.dev, thr=self.thr)
        self.setup_stream.synchronize()
        return e

    def _timed_entry(self, mol):
        t0 = time.perf_counter()
        e = self.entry_factory(mol)
        with self._setup_lock:
            self.setup_s += time.perf_counter() - t0
        return e

    def _make_room(self, need, keep):
        """Evict idle molecules (LRU first, none in ``keep``) so that ``need``
        more bytes fit; evicts nothing unless that is enough."""
        held = self.used + sum(self.reserved.values())
        if held + need <= self.budget:
            return True
        idle = [m for m in self.resident if m not in keep]
        free = 0.0
        plan = []
        for m in idle:
            if held - free + need <= self.budget:
                break
            plan.append(m)
            free += self.resident[m].nbytes
        if held - free + need > self.budget:
            return False
        for m in plan:
            e = self.resident.pop(m)
            self.used -= e.nbytes
            self.evictions += 1
            del e
        return True

    # ------------------------------------------------------------ batches
    def _launch(self, reqs, slot):
        entries = [self.resident[r.mol_id] for r in reqs]
        dens = [r.D for r in reqs]
        if self.batch_runner is not None:
            h = self.batch_runner(entries, dens)
            return h if hasattr(h, "ready") else _Done(h)
        return _GpuBatch(self, entries, dens, self.streams[slot])

    # ------------------------------------------------------------ loop
    def serve(self, requests, realtime=True, progress=None, progress_every=100, poll_s=0.0005):
        """Serve ``requests``.  With ``realtime`` the loop follows the wall
        clock (requests become visible at their arrival time); otherwise every
        request is visible at once (offline batching).  ``progress(msg)`` is
        called each time another ``progress_every`` requests have completed."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        n = len(reqs)
        pending = []
        nxt = done = 0
        inflight = []                  # [reqs, handle, slot]
        setups = {}                    # mol_id -> Future
        batches, max_res, max_fl, overlapped = [], 0, 0, 0
        pool = cf.ThreadPoolExecutor(max_workers=1, thread_name_prefix="vlxmi-setup")
        t0 = time.perf_counter()
        try:
            while done < n:
                now = time.perf_counter() - t0
                acted = False
                while nxt < n and (not realtime or reqs[nxt].arrival <= now):
                    pending.append(reqs[nxt])
                    nxt += 1
                # batches whose GPU work has finished
                for item in list(inflight):
                    if not item[1].ready():
                        continue
                    Js, Ks, reps = item[1].result()
                    t_done = time.perf_counter() - t0
                    for r, J, K, rep in zip(item[0], Js, Ks, reps):
                        r.t_done = t_done
                        r.report = rep
                        if self.keep_results:
                            r.J, r.K = J, K
                    inflight.remove(item)
                    if progress is not None and (done + len(item[0])) // progress_every > done // progress_every:
                        progress("served %d of %d requests" % (done + len(item[0]), n))
                    done += len(item[0])
                    acted = True
                # finished set-ups become resident
                for m, fu in list(setups.items()):
                    if fu.done():
                        e = fu.result()
                        del setups[m]
                        self.reserved.pop(m)
                        self.resident[m] = e
                        self.used += e.nbytes
                        self.setups += 1
                        acted = True
                max_res = max(max_res, len(self.resident))
                on_gpu = set(r.mol_id for item in inflight for r in item[0])
                waiting = set(r.mol_id for r in pending)
                # start set-ups for molecules of pending requests (in arrival order)
                for r in pending:
                    m = r.mol_id
                    if m in self.resident or m in setups:
                        continue
                    need = float(self.estimate_bytes(self.molecules[m]))
                    if need > self.budget:
                        raise RuntimeError("molecule %r needs about %.3g bytes, above the HBM budget (%.3g)"
                                           % (m, need, self.budget))
                    if not self._make_room(need, on_gpu | waiting | set(setups)):
                        break
                    self.reserved[m] = need
                    setups[m] = pool.submit(self._timed_entry, self.molecules[m])
                    acted = True
                # the next batch, when a slot is free
                if len(inflight) < self.inflight and pending:
                    first_it = {}
                    for q in pending:
                        first_it[q.mol_id] = min(first_it.get(q.mol_id, q.iteration), q.iteration)
                    batch, chosen = [], set()
                    for r in pending:
                        if len(batch) >= self.max_batch:
                            break
                        m = r.mol_id
                        if m in chosen or m in on_gpu or m not in self.resident or r.iteration > first_it[m]:
                            continue
                        batch.append(r)
                        chosen.add(m)
                    if batch:
                        for r in batch:
                            pending.remove(r)
                            self.resident.move_to_end(r.mol_id)
                        used_slots = set(item[2] for item in inflight)
                        slot = min(s for s in range(self.inflight) if s not in used_slots)
                        if inflight:
                            overlapped += 1
                        tb = time.perf_counter()
                        inflight.append([batch, self._launch(batch, slot), slot])
                        self.build_s += time.perf_counter() - tb
                        max_fl = max(max_fl, len(inflight))
                        batches.append(len(batch))
                        self.batch_log.append([r.rid for r in batch])
                        acted = True
                if not acted:
                    if not inflight and not setups and pending and nxt >= n:
                        raise RuntimeError("no pending request can be served within the HBM budget (%.3g bytes)"
                                           % self.budget)
                    wait = poll_s
                    if realtime and nxt < n and not pending and not inflight and not setups:
                        wait = max(0.0, reqs[nxt].arrival - (time.perf_counter() - t0))
                    time.sleep(wait)
        finally:
            pool.shutdown(wait=True)
        wall = time.perf_counter() - t0
        lat = np.array([(r.t_done - r.arrival) * 1e3 for r in reqs])
        return ServeStats(n=n, wall_s=wall, builds_per_s=n / wall,
                          p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                          mean_batch=float(np.mean(batches)), max_resident=max_res,
                          evictions=self.evictions, setups=self.setups, setup_s=self.setup_s,
                          pack_s=self.pack_s, build_s=self.build_s - self.pack_s, max_inflight=max_fl,
                          overlapped=overlapped, latencies_ms=lat.tolist())


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
