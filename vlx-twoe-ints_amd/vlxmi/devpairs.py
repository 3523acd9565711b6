"""The pair lists of ``pairs.build_lists``, built on the GPU.

Every step that grows with the molecule runs on the device: candidate blocks
and their overlap factors (``vlxmi_pair_candidates``), Schwarz factors
(``vlxmi_schwarz_{ab}``), block maxima (``vlxmi_block_absmax``), the screening
against Qglob * d_bound, the bra order (non-increasing block max), the exchange
CSR (device sorts by primitive, block max and partner, then
``vlxmi_kcsr_fill``) and the (i,k) lists.  Only the primitive tables (one row
per primitive) are made on the host.  The result is the same ``FockLists``
(host arrays, copied back once) that the engines and assembly plans take, in
the same order as the host builder produces (the same stable sort keys), so
``FockEngine(mol, lists=build_lists_device(mol, dev))`` is a drop-in for the
host path.
"""

from __future__ import annotations

import numpy as np

from . import pairs as P
from .basis import NCART
from .runtime import check, lib


def _np(x, dtype=None):
    a = x.cpu().numpy()
    return a.astype(dtype) if dtype is not None else a


def build_lists_device(mol, dev, thr=1e-12, d_bound=100.0, keep_device=False) -> P.FockLists:
    """``pairs.build_lists`` on the device (see the module docstring).  With
    ``keep_device`` the device tensors of every list are kept in
    ``fl.device`` as well."""
    t = dev.torch
    dv = dev.device
    L = lib()
    st = dev.stream_ptr()
    i32, i64, f64 = t.int32, t.int64, t.float64
    prims = P.build_prim_tables(mol)
    fl = P.FockLists(mol=mol, prims=prims, thr=thr, d_bound=d_bound)
    info = [dev.tensor(p.info) if p is not None else None for p in prims]
    fac = [dev.tensor(np.ascontiguousarray(p.comp_factor())) if p is not None else None for p in prims]
    aoi = [dev.tensor(p.aoinds.astype(np.int64)) if p is not None else None for p in prims]
    # candidates, S, Schwarz, block maxima
    raw = {}
    for pc in P.PAIR_CLASSES:
        la, lb = P.LVAL[pc[0]], P.LVAL[pc[1]]
        ta, tb = prims[la], prims[lb]
        if ta is None or tb is None:
            continue
        na, nb = NCART[la], NCART[lb]
        n = ta.count * (ta.count + 1) // 2 if la == lb else ta.count * tb.count
        I = t.empty(n, dtype=i32, device=dv)
        J = t.empty(n, dtype=i32, device=dv)
        S = t.empty(n * na * nb, dtype=f64, device=dv)
        check(L.vlxmi_pair_candidates(la, lb, info[la].data_ptr(), ta.count, info[lb].data_ptr(), tb.count,
                                      fac[la].data_ptr(), fac[lb].data_ptr(), I.data_ptr(), J.data_ptr(),
                                      S.data_ptr(), st), "pair_candidates")
        Q = t.empty_like(S)
        check(getattr(L, "vlxmi_schwarz_" + pc)(info[la].data_ptr(), ta.count, info[lb].data_ptr(), tb.count,
                                                 I.data_ptr(), J.data_ptr(), S.data_ptr(), n, Q.data_ptr(),
                                                 dev.boys_table.data_ptr(), st), "schwarz_" + pc)
        qm = t.empty(n, dtype=f64, device=dv)
        check(L.vlxmi_block_absmax(Q.data_ptr(), n, na * nb, qm.data_ptr(), st), "block_absmax")
        raw[pc] = (I, J, S, Q, qm)
    qglob = float(t.stack([v[4].max() for v in raw.values() if v[4].numel()]).max()) if raw else 0.0
    fl.qglob = qglob
    fl.pair_cut = thr / (qglob * d_bound) if qglob > 0 else 0.0
    # screened pair sets in bra order (stable: ties keep candidate order)
    ps_dev = {}
    for pc, (I, J, S, Q, qm) in raw.items():
        la, lb = P.LVAL[pc[0]], P.LVAL[pc[1]]
        nab = NCART[la] * NCART[lb]
        keep = t.nonzero(qm > fl.pair_cut).flatten()
        keep = keep[t.sort(qm[keep], descending=True, stable=True).indices]
        Sk = S.view(-1, nab)[keep].contiguous()
        Qk = Q.view(-1, nab)[keep].contiguous()
        ps_dev[pc] = (I[keep].to(i64), J[keep].to(i64), Sk, Qk, qm[keep].contiguous())
    raw.clear()
    for pc, (I, J, Sk, Qk, qmk) in ps_dev.items():
        la, lb = P.LVAL[pc[0]], P.LVAL[pc[1]]
        na, nb = NCART[la], NCART[lb]
        fl.pairsets[pc] = P.PairSet(la, lb, _np(I), _np(J), _np(Sk).reshape(-1, na, nb), _np(Qk).reshape(-1, na, nb))
        # J list: complete blocks, a-major components
        ca = t.arange(na, device=dv)[None, :, None]
        cb = t.arange(nb, device=dv)[None, None, :]
        first = (I[:, None, None] * na + ca).expand(-1, na, nb)
        second = (J[:, None, None] * nb + cb).expand(-1, na, nb)
        ao_first = aoi[la][ca.expand(I.numel(), na, nb), I[:, None, None].expand(-1, na, nb)]
        ao_second = aoi[lb][cb.expand(J.numel(), na, nb), J[:, None, None].expand(-1, na, nb)]
        if la == lb:
            lower = (I == J)[:, None, None] & (ca > cb)
        else:
            lower = t.zeros((I.numel(), na, nb), dtype=t.bool, device=dv)
        fl.jlists[pc] = P.JList(la, lb, _np(first.reshape(-1), np.uint32), _np(second.reshape(-1), np.uint32),
                                _np(Sk.reshape(-1)), _np(Qk.reshape(-1)), _np(ao_first.reshape(-1)),
                                _np(ao_second.reshape(-1)), _np(lower.reshape(-1)), _np(qmk))
    # exchange CSR per ordered type pair
    for la in range(3):
        for lb in range(3):
            if prims[la] is None or prims[lb] is None:
                continue
            lo, hi = min(la, lb), max(la, lb)
            pc = "spd"[lo] + "spd"[hi]
            if pc not in ps_dev:
                continue
            Ip, Jp, Sk, Qk, qmk = ps_dev[pc]
            nps = Ip.numel()
            src = t.arange(nps, device=dv)
            if la < lb:
                Io, Jo, tr, qo = Ip, Jp, t.zeros(nps, dtype=t.uint8, device=dv), qmk
            elif la > lb:
                Io, Jo, tr, qo = Jp, Ip, t.ones(nps, dtype=t.uint8, device=dv), qmk
            else:
                off = t.nonzero(Ip != Jp).flatten()
                Io, Jo = t.cat([Ip, Jp[off]]), t.cat([Jp, Ip[off]])
                tr = t.cat([t.zeros(nps, dtype=t.uint8, device=dv), t.ones(off.numel(), dtype=t.uint8, device=dv)])
                src = t.cat([src, off])
                qo = t.cat([qmk, qmk[off]])
            # order by I, then non-increasing block max, then J (three stable sorts)
            o = t.sort(Jo, stable=True).indices
            o = o[t.sort(qo[o], descending=True, stable=True).indices]
            o = o[t.sort(Io[o], stable=True).indices]
            Io, Jo, tr, src, qo = Io[o], Jo[o], tr[o].contiguous(), src[o], qo[o]
            na, nb = NCART[la], NCART[lb]
            cnt_a = prims[la].count
            nJ = t.bincount(Io, minlength=cnt_a)
            start = t.cumsum(nJ, 0) - nJ
            rank = t.arange(Io.numel(), device=dv) - start[Io]
            seg = nJ * nb
            base = t.cumsum(seg * na, 0) - seg * na
            displs = (base[:, None] + t.arange(na, device=dv)[None, :] * seg[:, None]).reshape(-1)
            counts = seg.repeat_interleave(na)
            nnz = int((seg * na).sum())
            Qe = t.empty(nnz, dtype=f64, device=dv)
            Se = t.empty(nnz, dtype=f64, device=dv)
            De = t.empty(nnz, dtype=i32, device=dv)
            # the pair set's S and Q, read (transposed where the block is (J, I))
            check(L.vlxmi_kcsr_fill(la, lb, src.contiguous().data_ptr(), tr.data_ptr(), Io.contiguous().data_ptr(),
                                    Jo.contiguous().data_ptr(), rank.contiguous().data_ptr(), base.data_ptr(),
                                    seg.contiguous().data_ptr(), Io.numel(), Sk.data_ptr(), Qk.data_ptr(),
                                    Qe.data_ptr(), Se.data_ptr(), De.data_ptr(), st), "kcsr_fill")
            qprim = t.zeros(cnt_a, dtype=f64, device=dv)
            if Io.numel():
                qprim.scatter_reduce_(0, Io, qo, reduce="amax", include_self=True)
            fl.kcsr["spd"[la] + "spd"[lb]] = P.KCSR(
                la, lb, _np(displs, np.uint32), _np(counts, np.uint32), _np(Qe), _np(De).view(np.uint32),
                _np(Se), _np(qprim), _np(qo), _np(start, np.uint32), _np(nJ, np.uint32))
    # (i,k) lists
    qprim = []
    for la in range(3):
        if prims[la] is None:
            qprim.append(None)
            continue
        q = t.zeros(prims[la].count, dtype=f64, device=dv)
        for lb in range(3):
            cs = fl.kcsr.get("spd"[la] + "spd"[lb])
            if cs is not None:
                q = t.maximum(q, dev.tensor(cs.qmax_prim))
        qprim.append(q)
    for pc in P.PAIR_CLASSES:
        la, lc = P.LVAL[pc[0]], P.LVAL[pc[1]]
        if prims[la] is None or prims[lc] is None:
            continue
        ca_, cc_ = prims[la].count, prims[lc].count
        if la == lc:
            IK = t.triu_indices(ca_, ca_, device=dv)
            I, K = IK[0], IK[1]
        else:
            I = t.arange(ca_, device=dv).repeat_interleave(cc_)
            K = t.arange(cc_, device=dv).repeat(ca_)
        keep = t.nonzero(qprim[la][I] * qprim[lc][K] * d_bound > thr).flatten()
        I, K = I[keep], K[keep]
        na, nc = NCART[la], NCART[lc]
        ii = (I[:, None, None] * na + t.arange(na, device=dv)[None, :, None]).expand(-1, na, nc)
        kk = (K[:, None, None] * nc + t.arange(nc, device=dv)[None, None, :]).expand(-1, na, nc)
        fl.iklists[pc] = P.IKList(la, lc, _np(ii.reshape(-1), np.uint32), _np(kk.reshape(-1), np.uint32),
                                  _np(I), _np(K))
    if keep_device:
        fl.device = ps_dev
    return fl
