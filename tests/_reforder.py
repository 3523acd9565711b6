"""Lists in the reference's order, made from this build's lists (test helper:
it stands for VeloxChem's host, which builds the reference kernels' inputs).

* J list of pair class ab: one entry per prim-Cartesian pair, i <= j for one
  type (the entries below the diagonal of a diagonal block dropped, as the
  weight (k != l) + 1 of gen_coulomb.py:276-281 implies); the bra copy in a
  shuffled order, the ket copy sorted by |Q_kl D_kl| per entry, non-increasing
  (gen_coulomb.py:205-212);
* exchange CSR: per combined index i, its partner entries sorted by |Q_K| on
  their own, non-increasing (gen_exchange.py:189-232);
* (i,k) list: i <= k for one type, shuffled.
"""

import numpy as np

from vlxmi.basis import NCART


def j_lists(jl, D_cart, rng):
    keep = ~jl.lower_diag
    idx = np.nonzero(keep)[0]
    bra = rng.permutation(idx)
    mD = D_cart[jl.ao_first, jl.ao_second]
    key = np.abs(jl.mat_Q[idx] * mD[idx])
    ket = idx[np.argsort(-key, kind="stable")]
    return (dict(first=jl.first[bra], second=jl.second[bra], pair_data=jl.pair_data[bra], mat_Q=jl.mat_Q[bra],
                 index=bra),
            dict(first=jl.first[ket], second=jl.second[ket], pair_data=jl.pair_data[ket], mat_Q=jl.mat_Q[ket],
                 mat_D=mD[ket], index=ket))


def k_csr(cs, count_a):
    na = NCART[cs.la]
    displs, counts, Q, Di, S = [], [], [], [], []
    pos = 0
    for i in range(count_a * na):
        d, c = int(cs.displs[i]), int(cs.counts[i])
        seg = np.arange(d, d + c)
        seg = seg[np.argsort(-np.abs(cs.Q[seg]), kind="stable")]
        displs.append(pos)
        counts.append(c)
        Q.append(cs.Q[seg])
        Di.append(cs.D_inds[seg])
        S.append(cs.pair_data[seg])
        pos += c
    cat = lambda v, dt: np.concatenate(v).astype(dt) if v else np.zeros(0, dt)
    return dict(displs=np.array(displs, np.uint32), counts=np.array(counts, np.uint32), Q=cat(Q, np.float64),
                D_inds=cat(Di, np.uint32), pair_data=cat(S, np.float64))


def ik_list(ik, rng):
    keep = np.ones(ik.i.size, dtype=bool) if ik.la != ik.lc else ik.i <= ik.k
    idx = rng.permutation(np.nonzero(keep)[0])
    return dict(i=ik.i[idx], k=ik.k[idx], index=idx)
