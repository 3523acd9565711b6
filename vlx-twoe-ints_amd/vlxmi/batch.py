"""Many molecules in one persistent launch per class.

``MolEntry`` keeps one molecule's density-independent data resident on the
device with molecule-local indices.  ``Pack`` concatenates a set of entries on
the device (index rebasing by per-type primitive offsets, AO offsets and
per-molecule density slots), builds the J work list (bra tiles of 64 blocks x
that molecule's ket range) and the per-molecule tables the batched K kernels
read.  ``BatchEngine.build(pack, densities)`` then runs, per class, one
``computeCoulombFock{AB}{CD}_batched[_mixed]`` and one
``computeExchangeFock{ABCD}_batched[_mixed]`` launch for all molecules of the
pack, and returns each molecule's J and K.  Results of a molecule do not depend
on the other molecules of the pack (same work items, same order, no atomics),
so a molecule gives the same bits alone and in any mix.

The exchange launches run on a second stream beside the Coulomb launches (a
fork/join of two chains that write different buffers), so latency-bound
Coulomb kernels and gather-bound exchange kernels share the CUs; eager builds
timed per launch (``timer`` set) stay on one stream.

``BatchEngine.capture(pack)`` records the whole build in a HIP graph (torch
CUDA-graph API over the HIP runtime) with static density/output buffers;
``replay`` refills the density slots and replays.

``autotune(pack)`` times, per class, a few persistent grid sizes and work-item
orders on the actual pack and keeps the fastest.
"""

from __future__ import annotations

import time

import numpy as np

from . import pairs as P
from .assemble import Assembler
from .basis import NCART
from .policy import use_mixed
from .runtime import Device, check, lib

LV = P.LVAL


class MolEntry:
    """Device-resident, density-independent data of one molecule."""

    def __init__(self, mol, dev: Device, thr=1e-12, d_bound=100.0, builder="device"):
        self.mol = mol
        self.dev = dev
        t = dev.torch
        if builder == "device":
            from .devpairs import build_lists_device
            self.fl = fl = build_lists_device(mol, dev, thr=thr, d_bound=d_bound)
        else:
            self.fl = fl = P.build_lists(mol, dev.schwarz, thr=thr, d_bound=d_bound)
        self.naos = mol.naos_cart
        self.nbf = mol.nbf
        Tm = mol.cart2sph()
        self.identity_T = (Tm.shape[0] == Tm.shape[1]) and np.allclose(Tm, np.eye(Tm.shape[0]))
        self.T = None if self.identity_T else dev.tensor(Tm)
        d = dev.tensor
        self.count = [p.count if p is not None else 0 for p in fl.prims]
        self.prim_info = [d(p.info) if p is not None else None for p in fl.prims]
        self.prim_aoinds = [d(p.aoinds) if p is not None else None for p in fl.prims]
        n = self.naos
        self.jl = {}
        for pc, jl in fl.jlists.items():
            self.jl[pc] = dict(first=d(jl.first), second=d(jl.second), pair_data=d(jl.pair_data),
                               mat_Q=d(jl.mat_Q), Didx=d((jl.ao_first * n + jl.ao_second).astype(np.int64)),
                               keep=d((~jl.lower_diag).astype(np.float64)), nblk=jl.nblk, count=jl.count)
        self.kc = {}
        for pc, cs in fl.kcsr.items():
            self.kc[pc] = dict(displs=d(cs.displs), counts=d(cs.counts), Q=d(cs.Q), D_inds=d(cs.D_inds),
                               pair_data=d(cs.pair_data), nnz=int(cs.Q.size), nblkcsr=int(cs.blk_max.size))
        self.ik = {}
        for pc, ik in fl.iklists.items():
            self.ik[pc] = dict(i=d(ik.i), k=d(ik.k), nblk=ik.nblk, count=int(ik.i.size))
        self._assembly_plans(fl)
        # AO type of every Cartesian AO (for the per-molecule maxD_bd table)
        aotype = np.zeros(n, dtype=np.int64)
        for p in fl.prims:
            if p is not None:
                aotype[p.aoinds.ravel()] = p.l
        self.pair_type = d((3 * aotype[:, None] + aotype[None, :]).ravel())
        self.nbytes = self._nbytes()

    def _assembly_plans(self, fl):
        """Sorted assembly plans of this molecule alone (value index -> AO
        element, ordered by output then value index), computed once.  Value
        indices are in the molecule's own layout (pair classes in PAIR_CLASSES
        order, each class's entries consecutive); ``pcid`` records each entry's
        pair class so a pack can move it to the class's slice of the batch
        buffer.  Concatenating these plans with offsets gives the batch plan in
        the same per-output order, so a molecule's sums are the same in any batch."""
        n = self.naos
        pcs = list(P.PAIR_CLASSES)
        plans = {}
        for kind in ("J", "K"):
            vi, oi, pi = [], [], []
            off = 0
            for ipc, pc in enumerate(pcs):
                if kind == "J":
                    if pc not in fl.jlists or fl.jlists[pc].nblk == 0:
                        continue
                    jl = fl.jlists[pc]
                    use = ~jl.lower_diag
                    offd = (jl.first != jl.second) if jl.la == jl.lb else np.ones(jl.count, dtype=bool)
                    sel1, sel2 = np.nonzero(use)[0], np.nonzero(use & offd)[0]
                    out1 = (jl.ao_first * n + jl.ao_second)[sel1]
                    out2 = (jl.ao_second * n + jl.ao_first)[sel2]
                    cnt = jl.count
                else:
                    if pc not in fl.iklists or fl.iklists[pc].nblk == 0:
                        continue
                    ik = fl.iklists[pc]
                    na, nc = NCART[ik.la], NCART[ik.lc]
                    ai = fl.prims[ik.la].aoinds[ik.i % na, ik.i // na].astype(np.int64)
                    ak = fl.prims[ik.lc].aoinds[ik.k % nc, ik.k // nc].astype(np.int64)
                    sel1 = np.arange(ik.i.size)
                    sel2 = (sel1 if ik.la != ik.lc
                            else np.nonzero(np.repeat(ik.I, na * nc) != np.repeat(ik.K, na * nc))[0])
                    out1, out2 = ai * n + ak, ak[sel2] * n + ai[sel2]
                    cnt = int(ik.i.size)
                vi += [sel1 + off, sel2 + off]
                oi += [out1, out2]
                pi += [np.full(sel1.size + sel2.size, ipc, dtype=np.uint8)]
                off += cnt
            vi = np.concatenate(vi).astype(np.int64) if vi else np.zeros(0, np.int64)
            oi = np.concatenate(oi).astype(np.int64) if oi else np.zeros(0, np.int64)
            pi = np.concatenate(pi) if pi else np.zeros(0, np.uint8)
            order = np.lexsort((vi, oi))
            vo, oo, po = vi[order], oi[order], pi[order]
            starts = np.nonzero(np.r_[True, oo[1:] != oo[:-1]])[0] if oo.size else np.zeros(0, np.int64)
            d = self.dev.tensor
            plans[kind] = dict(vo=d(vo.astype(np.int32)), pcid=d(po.astype(np.int32)),
                               seg_start=d(starts.astype(np.int32)), seg_out=d(oo[starts].astype(np.int32)),
                               nval=int(vo.size), nseg=int(starts.size))
        self.jplan, self.kplan = plans["J"], plans["K"]

    @staticmethod
    def estimate_nbytes(mol):
        """Upper bound of ``nbytes`` before the lists exist (every candidate
        pair kept), for admission by HBM budget: per J entry 40 bytes (indices,
        S, Q, density index, mask), per K CSR entry 20, per (i,k) entry 8, the
        assembly plans (16 bytes per contribution), the density / J / K /
        pair-type slots."""
        counts = [0, 0, 0]
        for sh in mol.shells:
            counts[sh.l] += sh.nprim
        nc = [NCART[l] for l in range(3)]
        jent = kent = 0
        for la in range(3):
            for lb in range(3):
                blocks = counts[la] * counts[lb]
                kent += blocks * nc[la] * nc[lb]
                if la <= lb:
                    jent += (counts[la] * (counts[la] + 1) // 2 if la == lb else blocks) * nc[la] * nc[lb]
        ikent = jent
        n = mol.naos_cart
        return 40 * jent + 20 * kent + 8 * ikent + 16 * 2 * (jent + ikent) + 8 * 3 * sum(counts) * 6 \
            + 8 * n * n * 4

    def _nbytes(self):
        tot = 0
        for v in list(self.jl.values()) + list(self.kc.values()) + list(self.ik.values()) + [self.jplan, self.kplan]:
            for x in v.values():
                if hasattr(x, "element_size"):
                    tot += x.element_size() * x.numel()
        for x in self.prim_info + self.prim_aoinds:
            if x is not None:
                tot += x.element_size() * x.numel()
        tot += 8 * self.naos * self.naos * 4        # D, J, K slots + pair-type map
        return tot


class Pack:
    """Concatenation of molecule entries for one batched build."""

    def __init__(self, entries, dev: Device):
        t = dev.torch
        self.dev = dev
        self.entries = list(entries)
        E = self.entries
        nm = len(E)
        self.nmol = nm
        dv = dev.device
        i32 = t.int32
        # offsets
        self.poff = [[0] * nm for _ in range(3)]
        for l in range(3):
            acc = 0
            for m, e in enumerate(E):
                self.poff[l][m] = acc
                acc += e.count[l]
        self.count = [sum(e.count[l] for e in E) for l in range(3)]
        ao_off, D_off = [], []
        a = o = 0
        for e in E:
            ao_off.append(a)
            D_off.append(o)
            a += e.naos
            o += e.naos * e.naos
        self.ao_off, self.D_off, self.Dsize, self.naos_total = ao_off, D_off, o, a
        self.prim_info = []
        self.prim_aoinds = []
        for l in range(3):
            if self.count[l] == 0:
                self.prim_info.append(None)
                self.prim_aoinds.append(None)
                continue
            self.prim_info.append(t.cat([e.prim_info[l] for e in E if e.count[l]], dim=1).contiguous())
            self.prim_aoinds.append(t.cat([e.prim_aoinds[l] + ao_off[m] for m, e in enumerate(E) if e.count[l]],
                                          dim=1).contiguous())
        # J lists
        self.jl = {}
        for pc in P.PAIR_CLASSES:
            la, lb = LV[pc[0]], LV[pc[1]]
            na, nb = NCART[la], NCART[lb]
            ms = [m for m, e in enumerate(E) if pc in e.jl and e.jl[pc]["nblk"] > 0]
            if not ms:
                continue
            boff, eoff = {}, {}
            b = c = 0
            for m in ms:
                boff[m], eoff[m] = b, c
                b += E[m].jl[pc]["nblk"]
                c += E[m].jl[pc]["count"]
            g = lambda key: t.cat([E[m].jl[pc][key] for m in ms]).contiguous()
            first = t.cat([E[m].jl[pc]["first"] + self.poff[la][m] * na for m in ms]).contiguous()
            second = t.cat([E[m].jl[pc]["second"] + self.poff[lb][m] * nb for m in ms]).contiguous()
            Didx = t.cat([E[m].jl[pc]["Didx"] + D_off[m] for m in ms]).contiguous()
            blk_mol = t.cat([t.full((E[m].jl[pc]["nblk"],), m, dtype=t.int64, device=dv) for m in ms])
            self.jl[pc] = dict(first=first, second=second, pair_data=g("pair_data"), mat_Q=g("mat_Q"), Didx=Didx,
                               keep=g("keep"), blk_mol=blk_mol, nblk=b, count=c, nab=na * nb, boff=boff, eoff=eoff,
                               nblk_m={m: E[m].jl[pc]["nblk"] for m in ms})
        # J work lists per class
        self.jwork = {}
        for ab, cd in P.j_classes():
            if ab not in self.jl or cd not in self.jl:
                continue
            A, Cc = self.jl[ab], self.jl[cd]
            items, cost = [], []
            for m, nbra in A["nblk_m"].items():
                if m not in Cc["nblk_m"]:
                    continue
                kb, ke = Cc["boff"][m], Cc["boff"][m] + Cc["nblk_m"][m]
                b0 = A["boff"][m] + np.arange(0, nbra, 64, dtype=np.int64)
                b1 = np.minimum(A["boff"][m] + nbra, b0 + 64)
                items.append(np.stack([b0, b1, np.full_like(b0, kb), np.full_like(b0, ke)], axis=1))
                cost.append((b1 - b0).astype(np.float64) * (ke - kb))
            if items:
                self.jwork[(ab, cd)] = (np.concatenate(items).astype(np.uint32), np.concatenate(cost))
        # K CSR
        self.kc = {}
        for la in range(3):
            for lb in range(3):
                pc = "spd"[la] + "spd"[lb]
                na, nb = NCART[la], NCART[lb]
                ms = [m for m, e in enumerate(E) if pc in e.kc]
                if not ms:
                    continue
                # every molecule that has type-a primitives must contribute displs/counts rows
                rows = [m for m, e in enumerate(E) if e.count[la]]
                nnz_off, o2 = {}, 0
                for m in ms:
                    nnz_off[m] = o2
                    o2 += E[m].kc[pc]["nnz"]
                displs, counts = [], []
                for m in rows:
                    if pc in E[m].kc:
                        displs.append(E[m].kc[pc]["displs"] + nnz_off[m])
                        counts.append(E[m].kc[pc]["counts"])
                    else:
                        z = t.zeros(E[m].count[la] * na, dtype=i32, device=dv)
                        displs.append(z)
                        counts.append(z)
                self.kc[pc] = dict(displs=t.cat(displs).contiguous(), counts=t.cat(counts).contiguous(),
                                   Q=t.cat([E[m].kc[pc]["Q"] for m in ms]).contiguous(),
                                   D_inds=t.cat([E[m].kc[pc]["D_inds"] + self.poff[lb][m] * nb for m in ms]).contiguous(),
                                   pair_data=t.cat([E[m].kc[pc]["pair_data"] for m in ms]).contiguous())
        # (i,k) lists
        self.ik = {}
        for pc in P.PAIR_CLASSES:
            la, lc = LV[pc[0]], LV[pc[1]]
            na, nc = NCART[la], NCART[lc]
            ms = [m for m, e in enumerate(E) if pc in e.ik and e.ik[pc]["nblk"] > 0]
            if not ms:
                continue
            eoff, o3 = {}, 0
            for m in ms:
                eoff[m] = o3
                o3 += E[m].ik[pc]["count"]
            self.ik[pc] = dict(
                i=t.cat([E[m].ik[pc]["i"] + self.poff[la][m] * na for m in ms]).contiguous(),
                k=t.cat([E[m].ik[pc]["k"] + self.poff[lc][m] * nc for m in ms]).contiguous(),
                blk_mol=t.cat([t.full((E[m].ik[pc]["nblk"],), m, dtype=i32, device=dv) for m in ms]).contiguous(),
                nblk=sum(E[m].ik[pc]["nblk"] for m in ms), count=o3, eoff=eoff)
            self.ik[pc]["blk_list"] = t.arange(self.ik[pc]["nblk"], dtype=i32, device=dv)
        self.mol_D_off = t.tensor(D_off, dtype=t.int64, device=dv)
        self.mol_naos = t.tensor([e.naos for e in E], dtype=i32, device=dv)
        self.mol_ao_off = t.tensor(ao_off, dtype=i32, device=dv)
        self.pair_type = t.cat([e.pair_type + 9 * m for m, e in enumerate(E)])
        self.jwork_dev = {k: dev.tensor(v[0].ravel()) for k, v in self.jwork.items()}
        self.tuning = {}
        # static per-class output buffers (views of one buffer) and assembly plans
        self.j_order = [pc for pc in P.PAIR_CLASSES if pc in self.jl]
        self.k_order = [pc for pc in P.PAIR_CLASSES if pc in self.ik]
        joff, koff = {}, {}
        o = 0
        for pc in self.j_order:
            joff[pc] = o
            o += self.jl[pc]["count"]
        self.jbuf = t.zeros(max(o, 1), dtype=t.float64, device=dv)
        o = 0
        for pc in self.k_order:
            koff[pc] = o
            o += self.ik[pc]["count"]
        self.kbuf = t.zeros(max(o, 1), dtype=t.float64, device=dv)
        self.matJ = {pc: self.jbuf[joff[pc]: joff[pc] + self.jl[pc]["count"]] for pc in self.j_order}
        self.matK = {pc: self.kbuf[koff[pc]: koff[pc] + self.ik[pc]["count"]] for pc in self.k_order}
        self.jasm = _pack_assembler(self, dev, joff, "jl", "jplan")
        self.kasm = _pack_assembler(self, dev, koff, "ik", "kplan")
        self.Jflat = t.zeros(self.Dsize, dtype=t.float64, device=dv)
        self.Kflat = t.zeros(self.Dsize, dtype=t.float64, device=dv)
        # static density-dependent inputs (filled by BatchEngine.prepare)
        self.Dflat = t.zeros(self.Dsize, dtype=t.float64, device=dv)
        self.kets = {pc: dict(first=t.zeros_like(j["first"]), second=t.zeros_like(j["second"]),
                              pair_data=t.zeros_like(j["pair_data"]), mat_Q=t.zeros_like(j["mat_Q"]),
                              mat_D=t.zeros_like(j["pair_data"]), count=j["count"]) for pc, j in self.jl.items()}
        self.maxD = t.zeros((9, max(nm, 1)), dtype=t.float64, device=dv)
        # per class launch (J classes in j_classes() order, then K classes): primitive
        # quartets evaluated in FP64 / FP32 by the last build, all molecules together
        self.launch_names = (["J:" + ab + cd for ab, cd in P.j_classes()] + ["K:" + c for c in P.k_classes()])
        self.qcount = t.zeros((len(self.launch_names), 2), dtype=t.int64, device=dv)


def _pack_assembler(pack, dev, class_off, lists, plan_attr):
    """The batch assembly plan: every molecule's own sorted plan, its value
    indices moved to the batch buffer (class offset + the molecule's offset in
    the class slice - the class offset in the molecule's own layout), its
    outputs to the molecule's density slot.  Device-side concatenation."""
    t = dev.torch
    dv = dev.device
    pcs = list(P.PAIR_CLASSES)
    perm, seg_start, seg_out = [], [], []
    vbase = 0
    for m, e in enumerate(pack.entries):
        plan = getattr(e, plan_attr)
        if plan["nval"] == 0:
            continue
        own = {}
        o = 0
        src = e.jl if lists == "jl" else e.ik
        for pc in pcs:
            if pc in src and src[pc]["nblk"] > 0:
                own[pc] = o
                o += src[pc]["count"]
        delta = np.zeros(len(pcs), dtype=np.int64)
        for pc, o in own.items():
            delta[pcs.index(pc)] = class_off[pc] + getattr(pack, lists)[pc]["eoff"][m] - o
        perm.append(plan["vo"].long() + t.as_tensor(delta, device=dv)[plan["pcid"].long()])
        seg_start.append(plan["seg_start"].long() + vbase)
        seg_out.append(plan["seg_out"].long() + pack.D_off[m])
        vbase += plan["nval"]
    if not perm:
        z = t.zeros(0, dtype=t.int64, device=dv)
        perm, seg_start, seg_out = [z], [z], [z]
    starts = t.cat(seg_start + [t.tensor([vbase], dtype=t.int64, device=dv)])
    return Assembler.from_device(dev, t.cat(perm), starts, t.cat(seg_out), pack.Dsize)


class BatchEngine:
    def __init__(self, dev: Device = None, thr=1e-12, precision="fp64", fp32_threshold=1e-7, omega=0.0,
                 mixed_policy="measured"):
        self.dev = dev or Device()
        self.torch = self.dev.torch
        self.thr = float(thr)
        self.precision = precision
        self.fp32_threshold = float(fp32_threshold)
        self.omega = float(omega)
        # which classes launch their _mixed kernel in mixed precision (vlxmi.policy)
        self.mixed_policy = mixed_policy
        self.timer = None
        # run the exchange launches on a second stream beside the Coulomb ones
        # (they write different buffers); eager timed builds stay on one stream
        self.concurrent = True
        self._kstreams = {}

    # ------------------------------------------------------------------ helpers
    def density_flat(self, pack: Pack, densities):
        t = self.torch
        parts, stats = [], []
        for e, D in zip(pack.entries, densities):
            Ds = D if t.is_tensor(D) else self.dev.tensor(np.asarray(D, dtype=np.float64))
            Dd = Ds if e.identity_T else e.T.t() @ Ds @ e.T
            stats.append(t.stack([Dd.abs().amax(), (Ds - Ds.t()).abs().amax(), Ds.abs().amax()]))
            # exactly symmetric: some exchange classes read D_jl as D[l][j]
            Dd = 0.5 * (Dd + Dd.t())
            parts.append(Dd.reshape(-1))
        # one host read for the whole batch: every density must fit the bound its
        # molecule's lists were screened for and be symmetric to rounding
        if stats:
            for e, st in zip(pack.entries, t.stack(stats).cpu().tolist()):
                P.check_density(e.fl, *st)
        return t.cat(parts).contiguous()

    def _maxD(self, pack, Dflat):
        t = self.torch
        out = t.zeros(pack.nmol * 9, dtype=t.float64, device=Dflat.device)
        out.scatter_reduce_(0, pack.pair_type, Dflat.abs(), reduce="amax", include_self=True)
        return out.view(pack.nmol, 9).t().contiguous()        # [bd][mol]

    def prepare(self, pack, densities):
        """Density-dependent inputs into the pack's static buffers (eager torch
        ops): Cartesian densities, per-molecule maxD_bd, and the ket copies of
        every J list ordered, within each molecule, by non-increasing block bound."""
        t = self.torch
        pack.Dflat.copy_(self.density_flat(pack, densities))
        Dflat = pack.Dflat
        pack.maxD.copy_(self._maxD(pack, Dflat))
        for pc, j in pack.jl.items():
            mD = Dflat[j["Didx"]] * j["keep"]
            nab = j["nab"]
            beta = (j["mat_Q"] * mD).abs().view(-1, nab).amax(dim=1)
            o1 = t.sort(beta, descending=True, stable=True).indices
            o2 = t.sort(j["blk_mol"][o1], stable=True).indices
            order = o1[o2]
            ent = (order[:, None] * nab + t.arange(nab, device=Dflat.device)[None, :]).reshape(-1)
            k = pack.kets[pc]
            k["first"].copy_(j["first"][ent])
            k["second"].copy_(j["second"][ent])
            k["pair_data"].copy_(j["pair_data"][ent])
            k["mat_Q"].copy_(j["mat_Q"][ent])
            k["mat_D"].copy_(mD[ent])

    def _ev(self, name, fn, *args):
        if self.timer is None:
            check(fn(*args), name)
            return
        t = self.torch
        e0 = t.cuda.Event(enable_timing=True)
        e1 = t.cuda.Event(enable_timing=True)
        e0.record()
        check(fn(*args), name)
        e1.record()
        self.timer.append((name, e0, e1))

    def launch(self, pack: Pack):
        """Zero the outputs, run every class launch for the pack and assemble
        J/K into pack.Jflat / pack.Kflat.  Only this library's kernels run here
        (safe to capture in a HIP graph)."""
        from .assemble import zero_
        L = lib()
        kets = pack.kets
        matJ = pack.matJ
        for buf in (pack.jbuf, pack.kbuf, pack.Jflat, pack.Kflat, pack.qcount):
            zero_(self.dev, buf)
        t = self.torch
        cur = t.cuda.current_stream(self.dev.device)
        side = None
        if self.concurrent and self.timer is None:
            # fork: the exchange chain waits for the zeroing above, then runs
            # beside the Coulomb chain; the join below orders the rest
            key = (self.dev.device, cur.cuda_stream)
            if key not in self._kstreams:
                self._kstreams[key] = t.cuda.Stream(device=self.dev.device)
            side = self._kstreams[key]
            side.wait_stream(cur)
        st = self.dev.stream_ptr()
        for (ab, cd), (items, cost) in pack.jwork.items():
            bra, ket = pack.jl[ab], kets[cd]
            cls = ab + cd
            args = [matJ[ab].data_ptr()]
            for x in "spd":
                if x in cls:
                    l = LV[x]
                    args += [pack.prim_info[l].data_ptr(), pack.count[l]]
            args += [ket["mat_D"].data_ptr(), bra["mat_Q"].data_ptr(), ket["mat_Q"].data_ptr(),
                     bra["first"].data_ptr(), bra["second"].data_ptr(), bra["pair_data"].data_ptr(), bra["count"],
                     ket["first"].data_ptr(), ket["second"].data_ptr(), ket["pair_data"].data_ptr(), ket["count"],
                     self.dev.boys_table.data_ptr(), self.dev.boys_ft.data_ptr(), self.thr]
            tun = pack.tuning.get(("J", cls), {})
            work = tun.get("work", pack.jwork_dev[(ab, cd)])
            args += [work.data_ptr(), len(items)]
            mixed = use_mixed("J", cls, self.precision, self.mixed_policy)
            qc = pack.qcount[pack.launch_names.index("J:" + cls)].data_ptr()
            args += [self.fp32_threshold, qc] if mixed else [qc]
            args += [tun.get("wg", 0), st]
            name = "computeCoulombFock" + cls.upper() + "_batched" + ("_mixed" if mixed else "")
            self._ev(name, getattr(L, "vlxmi_launch_" + name), *args)
        if side is not None:
            with t.cuda.stream(side):
                self._launch_k(pack, L, side.cuda_stream)
            cur.wait_stream(side)
        else:
            self._launch_k(pack, L, st)
        pack.jasm.run(pack.jbuf, pack.Jflat)
        pack.kasm.run(pack.kbuf, pack.Kflat)

    def _launch_k(self, pack, L, st):
        """The exchange class launches of ``launch`` on stream ``st``."""
        Dflat = pack.Dflat
        matK = pack.matK
        maxD = pack.maxD
        for cls in P.k_classes():
            a, b, c, d = cls
            ac, ab, cd, bd = a + c, a + b, c + d, b + d
            if ac not in pack.ik or ab not in pack.kc or cd not in pack.kc:
                continue
            ik = pack.ik[ac]
            args = [matK[ac].data_ptr(), ik["i"].data_ptr(), ik["k"].data_ptr(), ik["count"]]
            for x in "spd":
                if x in cls:
                    l = LV[x]
                    args += [pack.prim_info[l].data_ptr(), pack.prim_aoinds[l].data_ptr(), pack.count[l]]
            args += [1.0, Dflat.data_ptr(), 0]
            A, Cc = pack.kc[ab], pack.kc[cd]
            for nm in ("Q", "D_inds", "displs", "counts", "pair_data"):
                args.append(A[nm].data_ptr())
                if cd != ab:
                    args.append(Cc[nm].data_ptr())
            args += [self.dev.boys_table.data_ptr(), self.dev.boys_ft.data_ptr(), self.omega, self.thr]
            tun = pack.tuning.get(("K", cls), {})
            blk_list = tun.get("list", ik["blk_list"])
            bdi = 3 * LV[b] + LV[d]
            args += [blk_list.data_ptr(), ik["nblk"], ik["blk_mol"].data_ptr(), maxD[bdi].data_ptr(),
                     pack.mol_D_off.data_ptr(), pack.mol_naos.data_ptr(), pack.mol_ao_off.data_ptr()]
            mixed = use_mixed("K", cls, self.precision, self.mixed_policy)
            qc = pack.qcount[pack.launch_names.index("K:" + cls)].data_ptr()
            args += [self.fp32_threshold, qc] if mixed else [qc]
            args += [tun.get("wg", 0), st]
            name = "computeExchangeFock" + cls.upper() + "_batched" + ("_mixed" if mixed else "")
            self._ev(name, getattr(L, "vlxmi_launch_" + name), *args)

    def quartet_counts(self, pack: Pack):
        """{launch name ("J:abcd" / "K:abcd"): (FP64, FP32) primitive quartets}
        of the last build of ``pack`` (reads the device)."""
        c = pack.qcount.cpu().tolist()
        return {n: (int(a), int(b)) for n, (a, b) in zip(pack.launch_names, c) if a or b}

    def unpack(self, pack: Pack, flat):
        out = []
        for m, e in enumerate(pack.entries):
            n = e.naos
            X = flat[pack.D_off[m]: pack.D_off[m] + n * n].view(n, n)
            if not e.identity_T:
                X = e.T @ X @ e.T.t()
            out.append(X)
        return out

    def build(self, pack: Pack, densities):
        self.prepare(pack, densities)
        self.launch(pack)
        return [x.clone() for x in self.unpack(pack, pack.Jflat)], [x.clone() for x in self.unpack(pack, pack.Kflat)]

    # ------------------------------------------------------------------ graphs
    def capture(self, pack: Pack, warm_densities):
        """Record ``launch(pack)`` (every class launch + assembly) in a HIP graph.
        Returns replay(densities) -> (J list, K list): eager ``prepare`` into the
        pack's static buffers, then one graph replay."""
        t = self.torch
        dv = self.dev.device
        self.prepare(pack, warm_densities)
        s = t.cuda.Stream(device=dv)
        s.wait_stream(t.cuda.current_stream(dv))
        with t.cuda.stream(s):
            self.launch(pack)                 # warm-up outside the graph
        t.cuda.current_stream(dv).wait_stream(s)
        t.cuda.synchronize(dv)
        g = t.cuda.CUDAGraph()
        with t.cuda.graph(g):
            self.launch(pack)
        t.cuda.synchronize(dv)

        def replay(densities):
            self.prepare(pack, densities)
            g.replay()
            return ([x.clone() for x in self.unpack(pack, pack.Jflat)],
                    [x.clone() for x in self.unpack(pack, pack.Kflat)])

        replay.graph = g
        return replay

    # ------------------------------------------------------------------ autotuner
    def autotune(self, pack: Pack, densities, reps=2, wg_choices=(256, 512, 1024, 2048, 0), verbose=False,
                 progress=None):
        """Per class: persistent grid size x work-item order (natural or
        costliest-first), timed on this pack; the fastest is kept in pack.tuning.
        ``progress(msg)`` is called after each candidate round."""
        t = self.torch
        self.prepare(pack, densities)
        pack.tuning = {}
        # candidates
        cands = {}
        for (ab, cd), (items, cost) in pack.jwork.items():
            order = np.argsort(-cost, kind="stable")
            costly = self.dev.tensor(items[order].ravel())
            cands[("J", ab + cd)] = [dict(wg=w, work=wk) for w in wg_choices
                                     for wk in (pack.jwork_dev[(ab, cd)], costly)]
        for cls in P.k_classes():
            ac = cls[0] + cls[2]
            if ac in pack.ik and cls[0] + cls[1] in pack.kc and cls[2] + cls[3] in pack.kc:
                cands[("K", cls)] = [dict(wg=w) for w in wg_choices]
        best = {k: (None, float("inf")) for k in cands}
        nmax = max(len(v) for v in cands.values()) if cands else 0
        for i in range(nmax):
            pack.tuning = {k: v[i] for k, v in cands.items() if i < len(v)}
            for _ in range(reps):
                self.timer = []
                self.launch(pack)
                t.cuda.synchronize()
                for name, e0, e1 in self.timer:
                    fam = "J" if name.startswith("computeCoulomb") else "K"
                    cls = name.replace("computeCoulombFock", "").replace("computeExchangeFock", "").split("_")[0].lower()
                    key = (fam, cls)
                    if key in pack.tuning:
                        dt = e0.elapsed_time(e1)
                        if dt < best[key][1]:
                            best[key] = (pack.tuning[key], dt)
            if progress is not None:
                progress("autotune: candidate round %d of %d" % (i + 1, nmax))
            self.timer = None
        pack.tuning = {k: v[0] for k, v in best.items() if v[0] is not None}
        summary = {}
        for k, (cfg, dt) in best.items():
            if cfg is None:
                continue
            costly = "work" in cfg and cfg["work"] is not pack.jwork_dev.get((k[1][:2], k[1][2:]))
            summary["%s:%s" % k] = {"workgroups": cfg.get("wg"), "costliest_first": bool(costly), "ms": dt}
            if verbose:
                print(k, summary["%s:%s" % k])
        return summary
