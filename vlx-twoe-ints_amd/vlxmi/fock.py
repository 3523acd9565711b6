"""J/K Fock build driver on one GPU.

``FockEngine(mol)`` keeps everything that does not depend on the density
resident on the device (primitive tables, screened pair lists with Schwarz
factors, K CSR lists, (i,k) lists, Boys tables, assembly indices).  Each
``build(D)`` call:

1. transforms D to the Cartesian AO basis (T^T D T, on the device);
2. gathers ``mat_D`` for every J pair list, computes the block bounds
   ``beta = max|Q D|`` and orders the ket copy of each list by non-increasing
   bound (device sort);
3. launches the 36 ``computeCoulombFock{AB}{CD}`` and 54
   ``computeExchangeFock{A}{B}{C}{D}`` kernels through the C ABI;
4. assembles AO J and K from the per-pair outputs and transforms them back
   (T J T^T), and returns J, K (and E_J = 1/2 tr(D J), E_K = 1/2 tr(D K)).

VeloxChem's own host driver (pair building, sorting, AO assembly) is not in
the reference tree; this is the minimal host the kernels need.
"""

from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import pairs as P
from .basis import NCART
from .policy import use_mixed
from .runtime import Device, check, lib


class FockEngine:
    def __init__(self, mol, thr=1e-12, dev=None, d_bound=100.0, omega=0.0, target_waves=32768,
                 precision="fp64", fp32_threshold=1e-7, mixed_policy="measured", lists=None, builder="device"):
        """``lists``: the molecule's FockLists if already built (e.g. by a
        batched-engine entry of the same molecule); otherwise they are built
        here, on the device (``builder="device"``, vlxmi.devpairs) or by the
        host builder with the GPU Schwarz kernels (``"host"``)."""
        self.dev = dev or Device()
        t = self.dev.torch
        self.torch = t
        self.mol = mol
        self.thr = float(thr)
        self.omega = float(omega)
        self.target_waves = target_waves
        self.timer = None          # list of (kernel name, start event, end event) when set
        self.progress = None       # progress(msg): called after each class launch has finished (synchronizes)
        # exchange launches on a second stream beside the Coulomb launches (they
        # write different buffers); builds timed per launch stay on one stream
        self.concurrent = True
        self._kstreams = {}
        # "fp64": plain path; "mixed": quartets whose Schwarz x |D| bound is below
        # fp32_threshold run the VRR/HRR in FP32 (Boys, prefactors, sums in FP64)
        assert precision in ("fp64", "mixed")
        self.precision = precision
        self.fp32_threshold = float(fp32_threshold)
        # which classes launch their _mixed kernel in mixed precision (vlxmi.policy)
        self.mixed_policy = mixed_policy
        # device int64[2]: primitive quartets the _mixed kernels evaluated in FP64 / FP32
        # since the last reset_counts() (every mixed build() resets it first)
        self.qcount = None
        t0 = time.time()
        if lists is not None:
            assert lists.mol is mol and lists.thr == self.thr
            self.fl = lists
        elif builder == "device":
            from .devpairs import build_lists_device
            self.fl = build_lists_device(mol, self.dev, thr=thr, d_bound=d_bound)
        else:
            self.fl = P.build_lists(mol, self.dev.schwarz, thr=thr, d_bound=d_bound)
        self.setup_lists_s = time.time() - t0
        fl = self.fl
        d = self.dev
        self.naos = fl.naos
        self.T = d.tensor(mol.cart2sph())
        self.prim_info = [d.tensor(p.info) if p is not None else None for p in fl.prims]
        self.prim_aoinds = [d.tensor(p.aoinds) if p is not None else None for p in fl.prims]
        self.prim_count = [p.count if p is not None else 0 for p in fl.prims]
        self.jl = {}
        for pc, jl in fl.jlists.items():
            self.jl[pc] = dict(
                first=d.tensor(jl.first), second=d.tensor(jl.second),
                pair_data=d.tensor(jl.pair_data), mat_Q=d.tensor(jl.mat_Q),
                ao_first=d.tensor(jl.ao_first), ao_second=d.tensor(jl.ao_second),
                keep=d.tensor(~jl.lower_diag), count=jl.count, nblk=jl.nblk,
                nab=NCART[jl.la] * NCART[jl.lb])
        self.kc = {}
        for pc, cs in fl.kcsr.items():
            self.kc[pc] = dict(displs=d.tensor(cs.displs), counts=d.tensor(cs.counts), Q=d.tensor(cs.Q),
                               D_inds=d.tensor(cs.D_inds), pair_data=d.tensor(cs.pair_data),
                               blk_max=d.tensor(cs.blk_max), blk_displs=d.tensor(cs.blk_displs),
                               blk_counts=d.tensor(cs.blk_counts))
        self.ik = {}
        for pc, ik in fl.iklists.items():
            self.ik[pc] = dict(i=d.tensor(ik.i), k=d.tensor(ik.k), count=int(ik.i.size),
                               I=d.tensor(ik.I.astype(np.uint32)), K=d.tensor(ik.K.astype(np.uint32)),
                               nblk=int(ik.I.size))
        self._build_assembly()
        # AO index sets per angular type (for maxD_bd)
        self.ao_of_type = []
        for p in fl.prims:
            self.ao_of_type.append(d.tensor(np.unique(p.aoinds.ravel()).astype(np.int64)) if p is not None else None)
        self.setup_s = time.time() - t0

    # ------------------------------------------------------------------ assembly
    def _build_assembly(self):
        """Deterministic assembly plans (value index -> Cartesian AO element)."""
        from .assemble import Assembler
        fl, d, n = self.fl, self.dev, self.fl.naos
        vi, oi, off = [], [], 0
        self.j_order = [pc for pc in P.PAIR_CLASSES if pc in fl.jlists]
        for pc in self.j_order:
            jl = fl.jlists[pc]
            use = ~jl.lower_diag
            f, s = jl.ao_first, jl.ao_second
            offd = (jl.first != jl.second) if jl.la == jl.lb else np.ones(jl.count, dtype=bool)
            sel1 = np.nonzero(use)[0]
            sel2 = np.nonzero(use & offd)[0]
            vi += [off + sel1, off + sel2]
            oi += [f[sel1] * n + s[sel1], s[sel2] * n + f[sel2]]
            off += jl.count
        self.jasm = Assembler(d, np.concatenate(vi), np.concatenate(oi), n * n)
        self._jmap = (np.concatenate(vi), np.concatenate(oi), off)
        vi, oi, off = [], [], 0
        self.k_order = [pc for pc in P.PAIR_CLASSES if pc in fl.iklists]
        for pc in self.k_order:
            ik = fl.iklists[pc]
            la, lc = ik.la, ik.lc
            na, nc = NCART[la], NCART[lc]
            ai = fl.prims[la].aoinds[ik.i % na, ik.i // na].astype(np.int64)
            ak = fl.prims[lc].aoinds[ik.k % nc, ik.k // nc].astype(np.int64)
            if la != lc:
                sel2 = np.arange(ik.i.size)
            else:
                sel2 = np.nonzero(np.repeat(ik.I, na * nc) != np.repeat(ik.K, na * nc))[0]
            vi += [off + np.arange(ik.i.size), off + sel2]
            oi += [ai * n + ak, ak[sel2] * n + ai[sel2]]
            off += int(ik.i.size)
        self.kasm = Assembler(d, np.concatenate(vi), np.concatenate(oi), n * n)
        self._kmap = (np.concatenate(vi), np.concatenate(oi))
        self.fasm = None           # FockAssembler, made on the first fock() call

    def _cat(self, m, order, counts):
        t = self.torch
        parts = [m[pc] if pc in m else t.zeros(counts[pc], dtype=t.float64, device=self.dev.device) for pc in order]
        return t.cat(parts) if parts else t.zeros(0, dtype=t.float64, device=self.dev.device)

    def assemble_J(self, matJ):
        t = self.torch
        n = self.naos
        vals = self._cat(matJ, self.j_order, {pc: self.jl[pc]["count"] for pc in self.j_order})
        J = t.zeros(n * n, dtype=t.float64, device=self.dev.device)
        self.jasm.run(vals, J)
        return J.view(n, n)

    def assemble_K(self, matK):
        t = self.torch
        n = self.naos
        vals = self._cat(matK, self.k_order, {pc: self.ik[pc]["count"] for pc in self.k_order})
        K = t.zeros(n * n, dtype=t.float64, device=self.dev.device)
        self.kasm.run(vals, K)
        return K.view(n, n)

    # ------------------------------------------------------------------ per build
    def cart_density(self, D_sph):
        t = self.torch
        if not t.is_tensor(D_sph):
            D_sph = self.dev.tensor(np.asarray(D_sph, dtype=np.float64))
        Dc = self.T.t() @ D_sph @ self.T
        # one device reduction, one host read: the density must fit the bound
        # the lists were screened for and be symmetric to rounding
        st = t.stack([Dc.abs().amax(), (D_sph - D_sph.t()).abs().amax(), D_sph.abs().amax()]).cpu().tolist()
        P.check_density(self.fl, *st)
        # exactly symmetric: some exchange classes read D_jl as D[l][j]
        return (0.5 * (Dc + Dc.t())).contiguous()

    def prepare_kets(self, Dc):
        """mat_D per J list and the ket-ordered copies (device); ``mat_D_bra``
        keeps mat_D in the resident (bra) order."""
        t = self.torch
        kets = {}
        for pc, j in self.jl.items():
            mD = Dc[j["ao_first"], j["ao_second"]] * j["keep"].to(t.float64)
            nab = j["nab"]
            beta = (j["mat_Q"] * mD).abs().view(-1, nab).amax(dim=1)
            order = t.sort(beta, descending=True, stable=True).indices
            ent = (order[:, None] * nab + t.arange(nab, device=Dc.device)[None, :]).reshape(-1)
            kets[pc] = dict(first=j["first"][ent].contiguous(), second=j["second"][ent].contiguous(),
                            pair_data=j["pair_data"][ent].contiguous(), mat_Q=j["mat_Q"][ent].contiguous(),
                            mat_D=mD[ent].contiguous(), count=j["count"], mat_D_bra=mD)
        return kets

    def _ev_start(self):
        if self.timer is None:
            return None
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.torch.cuda.current_stream(self.dev.device))
        return e

    def _ev_end(self, e0, name):
        if self.progress is not None:
            self.torch.cuda.synchronize(self.dev.device)
            self.progress("%s done" % name)
        if e0 is None:
            return
        e1 = self.torch.cuda.Event(enable_timing=True)
        e1.record(self.torch.cuda.current_stream(self.dev.device))
        self.timer.append((name, e0, e1))

    def _ket_split(self, nbra_blocks, nket_blocks):
        """Ket-list chunks per Coulomb launch: enough workgroups (one wavefront
        each) for about ``target_waves`` waves, at least 64 ket blocks per
        chunk.  The light classes wait on the Boys-table gathers; 32768 waves
        (128 per CU over the launch) hide that latency where 2048 did not
        (lattice64 J build 1.40 s -> 1.11 s)."""
        tiles = max(1, (nbra_blocks + 63) // 64)
        s = max(1, -(-self.target_waves // tiles))
        return int(min(s, max(1, nket_blocks // 64)))

    def bra_slice(self, ab, blocks):
        """The bra list of pair class ``ab`` restricted to the primitive-pair
        blocks ``blocks`` (indices into the resident list): the reference's
        ``_local`` arguments (gen_coulomb.py:81-86), the bra share one GPU of
        several would take.  The ket list stays whole."""
        t = self.torch
        j = self.jl[ab]
        nab = j["nab"]
        b = t.as_tensor(np.asarray(blocks, dtype=np.int64), device=self.dev.device)
        ent = (b[:, None] * nab + t.arange(nab, device=self.dev.device)[None, :]).reshape(-1)
        return dict(first=j["first"][ent].contiguous(), second=j["second"][ent].contiguous(),
                    pair_data=j["pair_data"][ent].contiguous(), mat_Q=j["mat_Q"][ent].contiguous(),
                    count=int(ent.numel()), nblk=int(b.numel()), nab=nab)

    def ik_slice(self, ac, blocks):
        """The (i,k) list of class ``ac`` restricted to the (I,K) blocks
        ``blocks`` (the exchange share of one GPU of several)."""
        t = self.torch
        ik = self.ik[ac]
        nac = ik["count"] // max(ik["nblk"], 1)
        b = t.as_tensor(np.asarray(blocks, dtype=np.int64), device=self.dev.device)
        ent = (b[:, None] * nac + t.arange(nac, device=self.dev.device)[None, :]).reshape(-1)
        return dict(i=ik["i"][ent].contiguous(), k=ik["k"][ent].contiguous(), count=int(ent.numel()),
                    nblk=int(b.numel()))

    def run_J(self, Dc, kets=None, classes=None, bra_blocks=None):
        """Per bra class mat_J (bra list order), summed over ket classes.
        ``bra_blocks`` ({ab: block indices}) runs each bra class on that slice of
        its list only; mat_J then follows the slice's order."""
        t = self.torch
        L = lib()
        kets = kets or self.prepare_kets(Dc)
        out = {}
        slices = {}
        for ab, cd in (classes or P.j_classes()):
            if ab not in self.jl or cd not in self.jl:
                continue
            if bra_blocks is not None and ab not in slices:
                slices[ab] = self.bra_slice(ab, bra_blocks[ab])
            bra, ket = slices.get(ab, self.jl[ab]), kets[cd]
            if ab not in out:
                out[ab] = t.zeros(bra["count"], dtype=t.float64, device=Dc.device)
            cls = ab + cd
            args = [out[ab].data_ptr()]
            for x in "spd":
                if x in cls:
                    l = P.LVAL[x]
                    args += [self.prim_info[l].data_ptr(), self.prim_count[l]]
            args += [ket["mat_D"].data_ptr(), bra["mat_Q"].data_ptr(), ket["mat_Q"].data_ptr(),
                     bra["first"].data_ptr(), bra["second"].data_ptr(), bra["pair_data"].data_ptr(), bra["count"],
                     ket["first"].data_ptr(), ket["second"].data_ptr(), ket["pair_data"].data_ptr(), ket["count"],
                     self.dev.boys_table.data_ptr(), self.dev.boys_ft.data_ptr(), self.thr]
            split = self._ket_split(bra["nblk"], self.jl[cd]["nblk"])
            mixed = use_mixed("J", cls, self.precision, self.mixed_policy)
            name = "computeCoulombFock" + cls.upper() + ("_mixed" if mixed else "")
            if mixed:
                args += [self.fp32_threshold, self._qcount_ptr()]
            fn = getattr(L, "vlxmi_launch_" + name)
            ev = self._ev_start()
            check(fn(*args, split, self.dev.stream_ptr()), name)
            self._ev_end(ev, name)
        return out

    def max_D(self, Dc):
        out = {}
        for lb in range(3):
            for ld in range(3):
                ab, ad = self.ao_of_type[lb], self.ao_of_type[ld]
                if ab is None or ad is None:
                    continue
                out["spd"[lb] + "spd"[ld]] = float(Dc[ab][:, ad].abs().max().item())
        return out

    def run_K(self, Dc, classes=None, maxD=None, ik_blocks=None):
        """Per (a,c) class mat_K ((i,k) list order), summed over (b,d).
        ``ik_blocks`` ({ac: block indices}) runs each class on that slice of its
        (i,k) list only; the CSR lists stay whole."""
        t = self.torch
        L = lib()
        maxD = maxD or self.max_D(Dc)
        out = {}
        slices = {}
        for cls in (classes or P.k_classes()):
            a, b, c, d = cls
            ac, ab, cd, bd = a + c, a + b, c + d, b + d
            if ac not in self.ik or ab not in self.kc or cd not in self.kc:
                continue
            if ik_blocks is not None and ac not in slices:
                slices[ac] = self.ik_slice(ac, ik_blocks[ac])
            ik = slices.get(ac, self.ik[ac])
            if ac not in out:
                out[ac] = t.zeros(ik["count"], dtype=t.float64, device=Dc.device)
            args = [out[ac].data_ptr(), ik["i"].data_ptr(), ik["k"].data_ptr(), ik["count"]]
            for x in "spd":
                if x in cls:
                    l = P.LVAL[x]
                    args += [self.prim_info[l].data_ptr(), self.prim_aoinds[l].data_ptr(), self.prim_count[l]]
            args += [maxD[bd], Dc.data_ptr(), self.naos]
            A, Cc = self.kc[ab], self.kc[cd]
            for nm in ("Q", "D_inds", "displs", "counts", "pair_data"):
                args.append(A[nm].data_ptr())
                if cd != ab:
                    args.append(Cc[nm].data_ptr())
            args += [self.dev.boys_table.data_ptr(), self.dev.boys_ft.data_ptr(), self.omega, self.thr]
            mixed = use_mixed("K", cls, self.precision, self.mixed_policy)
            name = "computeExchangeFock" + cls.upper() + ("_mixed" if mixed else "")
            if mixed:
                args += [self.fp32_threshold, self._qcount_ptr()]
            fn = getattr(L, "vlxmi_launch_" + name)
            ev = self._ev_start()
            check(fn(*args, self.dev.stream_ptr()), name)
            self._ev_end(ev, name)
        return out

    def count_k_quartets(self, cls, maxD):
        """N_q of one K class (primitive shell quartets passing the block screen)."""
        t = self.torch
        a, b, c, d = cls
        ik, A, Cc = self.ik.get(a + c), self.kc.get(a + b), self.kc.get(c + d)
        if ik is None or A is None or Cc is None or not maxD > 0:
            return 0
        out = t.zeros(1, dtype=t.int64, device=self.dev.device)
        check(lib().vlxmi_count_k_quartets(ik["I"].data_ptr(), ik["K"].data_ptr(), ik["nblk"],
                                           A["blk_max"].data_ptr(), A["blk_displs"].data_ptr(), A["blk_counts"].data_ptr(),
                                           Cc["blk_max"].data_ptr(), Cc["blk_displs"].data_ptr(), Cc["blk_counts"].data_ptr(),
                                           self.thr / maxD, out.data_ptr(), self.dev.stream_ptr()), "count_k_quartets")
        return int(out.item())

    def run_JK(self, Dc):
        """run_J and run_K; with ``concurrent`` (and no per-launch timer or
        progress callback) the exchange launches run on a second stream that
        forks from and joins the current one."""
        t = self.torch
        if not (self.concurrent and self.timer is None and self.progress is None):
            return self.run_J(Dc), self.run_K(Dc)
        cur = t.cuda.current_stream(self.dev.device)
        key = cur.cuda_stream
        if key not in self._kstreams:
            self._kstreams[key] = t.cuda.Stream(device=self.dev.device)
        side = self._kstreams[key]
        side.wait_stream(cur)
        mJ = self.run_J(Dc)
        with t.cuda.stream(side):
            maxD = self.max_D(Dc)
            mK = self.run_K(Dc, maxD=maxD)
        cur.wait_stream(side)
        for v in mK.values():
            v.record_stream(cur)       # allocated on the side stream, read on this one
        return mJ, mK

    def fock(self, D_sph, H=None, cJ=2.0, cK=-1.0):
        """F = H + c_J J + c_K K in the output basis for density D (device
        tensors; H optional, the core Hamiltonian is not computed here).  The
        Coulomb and exchange per-pair outputs are assembled into c_J J + c_K K
        by one deterministic pass (FockAssembler), then transformed to the
        output basis; the defaults give the closed-shell 2J - K for D without
        the spin factor 2."""
        from .assemble import FockAssembler
        t = self.torch
        n = self.naos
        if self.fasm is None:
            jv, jo, nj = self._jmap
            kv, ko = self._kmap
            self.fasm = FockAssembler(self.dev, jv, jo, nj, kv, ko, n * n)
        if self.precision == "mixed":
            self.reset_counts()
        Dc = self.cart_density(D_sph)
        mJ, mK = self.run_JK(Dc)
        vals = t.cat([self._cat(mJ, self.j_order, {pc: self.jl[pc]["count"] for pc in self.j_order}),
                      self._cat(mK, self.k_order, {pc: self.ik[pc]["count"] for pc in self.k_order})])
        G = t.zeros(n * n, dtype=t.float64, device=self.dev.device)
        self.fasm.run(vals, G, cJ, cK)
        F = self.T @ G.view(n, n) @ self.T.t()
        if H is not None:
            F = F + (H if t.is_tensor(H) else self.dev.tensor(np.asarray(H, dtype=np.float64)))
        return F

    def reset_counts(self):
        """Zero the FP64/FP32 quartet counters of the _mixed kernels."""
        t = self.torch
        if self.qcount is None:
            self.qcount = t.zeros(2, dtype=t.int64, device=self.dev.device)
        else:
            self.qcount.zero_()

    def _qcount_ptr(self):
        if self.qcount is None:
            self.reset_counts()
        return self.qcount.data_ptr()

    def quartet_counts(self):
        """(FP64, FP32) primitive quartets evaluated by the _mixed kernels since
        the last reset (a mixed build() resets first); classes that ran their
        plain FP64 kernel are not counted.  Reads the device (synchronizes)."""
        if self.qcount is None:
            return 0, 0
        n64, n32 = self.qcount.cpu().tolist()
        return int(n64), int(n32)

    def build(self, D_sph, want_k=True):
        """J and K in the output basis for density D (device tensors)."""
        if self.precision == "mixed":
            self.reset_counts()
        Dc = self.cart_density(D_sph)
        if not want_k:
            return self.T @ self.assemble_J(self.run_J(Dc)) @ self.T.t(), None
        mJ, mK = self.run_JK(Dc)
        J = self.T @ self.assemble_J(mJ) @ self.T.t()
        K = self.T @ self.assemble_K(mK) @ self.T.t()
        return J, K

    def build_with_report(self, D_sph):
        """J, K and a report of the build: in mixed precision the quartet
        counts (FP64 / FP32, classes that ran their _mixed kernel) and the
        errors against the plain FP64 build of the same density (dE_J, dE_K,
        max|dJ|, max|dK|), which this call also runs.  Synchronizes."""
        t = self.torch
        if not t.is_tensor(D_sph):
            D_sph = self.dev.tensor(np.asarray(D_sph, dtype=np.float64))
        J, K = self.build(D_sph)
        rep = {"precision": self.precision}
        if self.precision == "mixed":
            n64, n32 = self.quartet_counts()
            rep.update(fp32_quartets=n32, mixed_kernel_quartets=n64 + n32, fp32_threshold=self.fp32_threshold)
            self.precision = "fp64"
            try:
                J64, K64 = self.build(D_sph)
            finally:
                self.precision = "mixed"
            EJ, EK = self.energies(D_sph, J, K)
            EJ64, EK64 = self.energies(D_sph, J64, K64)
            rep.update(reference="fp64", dE_J=EJ - EJ64, dE_K=EK - EK64,
                       max_abs_dJ=float((J - J64).abs().max()), max_abs_dK=float((K - K64).abs().max()))
        return J, K, rep

    def energies(self, D_sph, J, K):
        t = self.torch
        if not t.is_tensor(D_sph):
            D_sph = self.dev.tensor(np.asarray(D_sph, dtype=np.float64))
        EJ = 0.5 * float((D_sph * J).sum().item())
        EK = 0.5 * float((D_sph * K).sum().item()) if K is not None else None
        return EJ, EK
