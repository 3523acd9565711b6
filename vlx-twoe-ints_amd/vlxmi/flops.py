"""Algorithmic FLOP model of the J/K build (implementation independent).

W = sum over classes of N_q(class) * F(class) where

* N_q = number of primitive shell quartets that pass the reference screening
  for at least one Cartesian component pair: J: max_ij|Q_ij| * max_kl|Q_kl D_kl|
  > thr; K: QKmax_IJ * QKmax_KL * maxD_bd > thr.
* F(class) = FLOPs of one canonical Obara-Saika/HGP evaluation of a primitive
  quartet, FMA = 2:  prefactors 40; Boys 20 + 4(L+1); VRR 2 per term of each
  intermediate's recursion (the minimal intermediate set of the class);
  HRR 2 per intermediate; digestion 2 per Cartesian component quartet.

``python -m vlxmi.flops`` writes ``flop_model.json`` next to this file.  The
table depends only on the class, never on tuning.
"""

from __future__ import annotations

import json
import os

import numpy as np

from .gen.osdag import QuartetDAG, tri
from .pairs import LVAL, j_classes, k_classes

MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flop_model.json")


def _canonical_vrr_hrr_flops(la, lb, lc, ld):
    g = QuartetDAG(la, lb, lc, ld, rule_vrr="first", rule_hrr="first")
    for c in tri(lc):
        for d in tri(ld):
            for a in tri(la):
                for b in tri(lb):
                    g.X(a, b, c, d)
    vrr = hrr = 0
    for ln in g.lines:
        rhs = ln.split("=", 1)[1]
        if "CD_" in rhs or "AB_" in rhs:
            hrr += 2
        else:
            # one multiply(-add) per '*' outside the (c - r*d) pair counts as a term
            vrr += 2 * rhs.count("*")
    return vrr, hrr


def quartet_flops(la, lb, lc, ld):
    L = la + lb + lc + ld
    vrr, hrr = _canonical_vrr_hrr_flops(la, lb, lc, ld)
    ncomp = ((la + 1) * (la + 2) // 2) * ((lb + 1) * (lb + 2) // 2) * ((lc + 1) * (lc + 2) // 2) * ((ld + 1) * (ld + 2) // 2)
    return {"prefactors": 40, "boys": 20 + 4 * (L + 1), "vrr": vrr, "hrr": hrr, "digest": 2 * ncomp,
            "total": 40 + 20 + 4 * (L + 1) + vrr + hrr + 2 * ncomp}


def model():
    out = {"J": {}, "K": {}}
    for ab, cd in j_classes():
        cls = ab + cd
        out["J"][cls] = quartet_flops(*(LVAL[x] for x in cls))
    for cls in k_classes():
        out["K"][cls] = quartet_flops(*(LVAL[x] for x in cls))
    return out


def load_model():
    with open(MODEL_PATH) as f:
        return json.load(f)


# --------------------------------------------------------------------------
# surviving-quartet counts N_q
# --------------------------------------------------------------------------

def count_j(fl, D_cart, ab, cd):
    """Primitive shell quartets of J class (ab|cd) passing the screen."""
    bra, ket = fl.jlists[ab], fl.jlists[cd]
    ncd = ket.count // ket.nblk if ket.nblk else 1
    mD = ket.mat_D(D_cart)
    beta = np.sort(np.abs(ket.mat_Q * mD).reshape(ket.nblk, ncd).max(axis=1))
    lim = fl.thr / np.maximum(bra.qmax_blk, 1e-300)
    return int((beta.size - np.searchsorted(beta, lim, side="right")).sum())


def count_k(fl, D_cart, cls, maxD):
    """Primitive shell quartets of K class (ab|cd) (a<=c) passing the screen."""
    a, b, c, d = cls
    ik = fl.iklists.get(a + c)
    A = fl.kcsr.get(a + b)
    Cc = fl.kcsr.get(c + d)
    if ik is None or A is None or Cc is None or maxD <= 0:
        return 0
    la, lb, lc, ld = LVAL[a], LVAL[b], LVAL[c], LVAL[d]
    na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc, ld)]

    def block_q(cs, nx, ny, count_x):
        # per primitive X: array of partner block maxima (in CSR order)
        out = []
        for X in range(count_x):
            cnt = int(cs.counts[X * nx])
            if cnt == 0:
                out.append(np.zeros(0))
                continue
            segs = [np.abs(cs.Q[cs.displs[X * nx + t]: cs.displs[X * nx + t] + cnt]).reshape(-1, ny) for t in range(nx)]
            out.append(np.max(np.stack([s.max(axis=1) for s in segs]), axis=0))
        return out

    qa = block_q(A, na, nb, fl.prims[la].count)
    qc = block_q(Cc, nc, nd, fl.prims[lc].count)
    total = 0
    # group ik blocks by I
    order = np.argsort(ik.I, kind="stable")
    Is, Ks = ik.I[order], ik.K[order]
    starts = np.searchsorted(Is, np.unique(Is))
    ends = np.append(starts[1:], Is.size)
    qc_sorted = [np.sort(q) for q in qc]
    for s, e in zip(starts, ends):
        I = Is[s]
        if qa[I].size == 0:
            continue
        lim = np.sort(fl.thr / (qa[I] * maxD))          # threshold on QKmax_KL per J
        for K in Ks[s:e]:
            q = qc_sorted[K]
            if q.size == 0:
                continue
            # pairs (J, L) with q_L > lim_J
            total += int((q.size - np.searchsorted(q, lim, side="right")).sum())
    return total


if __name__ == "__main__":
    m = model()
    with open(MODEL_PATH, "w") as f:
        json.dump(m, f, indent=1, sort_keys=True)
    print("wrote", MODEL_PATH)
    for k in ("ssss", "pppp", "dddd"):
        print(k, m["J"][k])
