"""Primitive tables and screened pair lists in the reference's data layout.

Layout kept from the reference kernels (argument names in brackets):

* ``{x}_prim_info`` - float64 SoA ``[5][prim_count]``: row 0 exponent, rows
  2..4 centre (gen_coulomb.py:170-180).  Row 1 is never read by a kernel; here it
  holds ``cnorm * c_k * N(a_k; x^l)``.
* combined prim-Cartesian index ``prim * ncart + cart`` (gen_coulomb.py:170,189).
* ``{x}_prim_aoinds`` - uint32 ``[ncart][prim_count]``: Cartesian AO index of each
  primitive component (gen_exchange.py:207-209).
* J lists per pair class ``ab`` in {ss, sp, sd, pp, pd, dd}: parallel arrays
  ``first_inds``, ``second_inds`` (uint32 combined indices), ``pair_data`` (S_ij),
  ``mat_Q`` (Schwarz), ``mat_D`` (density of the pair) (gen_coulomb.py:74-93).
* K lists: the ``(i,k)`` pair list per class ``ac`` plus, per ordered type pair
  ``ab``, a CSR over combined indices i: ``pair_displs_K``, ``pair_counts_K``, and
  per entry ``Q_K``, ``D_inds_K`` (partner combined index), ``pair_data_K``
  (gen_exchange.py:75-115,161-168,199-221).

Ordering contract (what the HIP kernels rely on, stated once here):

* Every J list is a sequence of complete primitive-pair blocks: the
  ``na*nb`` entries of one primitive pair (I, J) are consecutive, component
  order a-major (entry ``ca*nb + cb``).  Same-type lists hold blocks I <= J;
  a diagonal block (I == J) is complete too, and in its ket role the entries
  below the diagonal (ca > cb) carry ``mat_D = 0`` so they contribute nothing
  (the reference lists simply omit them; the sums are identical).
* Ket lists are ordered by non-increasing block bound
  ``beta_KL = max |Q_kl * D_kl|`` over the block, so a scan may stop at the first
  block whose bound fails (the reference orders entries, not blocks, by the same
  key: gen_coulomb.py:209-212).
* The ``(i,k)`` list holds complete (I, K) blocks (``na*nc`` entries, a-major);
  same-type lists hold I <= K.
* CSR segments: the segment of combined index ``i = I*na + ca`` lists the partner
  blocks J of I (``nb`` consecutive entries each, component order) ordered by
  non-increasing ``QKmax_IJ`` (max over the whole (I,J) block); the partner order
  is therefore the same for every component ``ca`` of I.

S_ij (``pair_data``) = ``f_i f_j (pi/zeta)^{3/2} exp(-a_i a_j |AB|^2 / zeta)`` with
``f`` = contraction x normalisation factor of the primitive component, so that
``eri = Lambda * S_ij * S_kl * poly`` as in gen_coulomb.py:270.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .basis import NCART, Molecule, cart_components, component_norm_ratio

PAIR_CLASSES = ["ss", "sp", "sd", "pp", "pd", "dd"]
LVAL = {"s": 0, "p": 1, "d": 2}


@dataclass
class PrimTable:
    l: int
    info: np.ndarray          # (5, count) float64
    aoinds: np.ndarray        # (ncart, count) uint32
    shell: np.ndarray         # (count,) int64 shell index
    ratio: np.ndarray         # (ncart,) component normalisation ratios

    @property
    def count(self) -> int:
        return int(self.info.shape[1])

    @property
    def nc(self) -> int:
        return NCART[self.l]

    def comp_factor(self) -> np.ndarray:
        """(count, ncart) factor of each primitive component."""
        return self.info[1][:, None] * self.ratio[None, :]


def build_prim_tables(mol: Molecule):
    tabs = []
    offs = mol.cart_offsets()
    for l in range(3):
        exps, facs, cen, aos, shl = [], [], [], [], []
        for si, s in enumerate(mol.shells):
            if s.l != l:
                continue
            f = s.prim_factors()
            for k in range(s.nprim):
                exps.append(s.exps[k])
                facs.append(f[k])
                cen.append(s.center)
                aos.append([offs[si] + c for c in range(NCART[l])])
                shl.append(si)
        if not exps:
            tabs.append(None)
            continue
        n = len(exps)
        info = np.zeros((5, n))
        info[0] = exps
        info[1] = facs
        info[2:5] = np.array(cen).T
        aoinds = np.ascontiguousarray(np.array(aos, dtype=np.uint32).T)
        ratio = np.array([component_norm_ratio(l, c) for c in cart_components(l)])
        tabs.append(PrimTable(l, np.ascontiguousarray(info), aoinds, np.array(shl), ratio))
    return tabs


def pair_overlap_factors(ta: PrimTable, tb: PrimTable, I: np.ndarray, J: np.ndarray) -> np.ndarray:
    """S per component pair, shape (nblk, na, nb)."""
    ai = ta.info[0][I]
    aj = tb.info[0][J]
    A = ta.info[2:5][:, I]
    B = tb.info[2:5][:, J]
    z = ai + aj
    ab2 = np.sum((A - B) ** 2, axis=0)
    base = (math.pi / z) ** 1.5 * np.exp(-ai * aj / z * ab2)
    fa = ta.comp_factor()[I]          # (nblk, na)
    fb = tb.comp_factor()[J]          # (nblk, nb)
    return base[:, None, None] * fa[:, :, None] * fb[:, None, :]


@dataclass
class PairSet:
    """Screened primitive-pair blocks of one pair class ab (a <= b)."""
    la: int
    lb: int
    I: np.ndarray      # (nblk,) primitive index of type a
    J: np.ndarray      # (nblk,) primitive index of type b
    S: np.ndarray      # (nblk, na, nb)
    Q: np.ndarray      # (nblk, na, nb)

    @property
    def nblk(self) -> int:
        return int(self.I.size)

    @property
    def qmax(self) -> np.ndarray:
        return np.abs(self.Q).reshape(self.nblk, -1).max(axis=1) if self.nblk else np.zeros(0)


def candidate_blocks(ta, tb, same):
    if same:
        I, J = np.triu_indices(ta.count)
    else:
        I, J = np.meshgrid(np.arange(ta.count), np.arange(tb.count), indexing="ij")
        I, J = I.ravel(), J.ravel()
    return I.astype(np.int64), J.astype(np.int64)


@dataclass
class JList:
    """One J pair list (class ab) in the reference layout."""
    la: int
    lb: int
    first: np.ndarray        # uint32 (n,)
    second: np.ndarray       # uint32 (n,)
    pair_data: np.ndarray    # float64 (n,)
    mat_Q: np.ndarray        # float64 (n,)
    ao_first: np.ndarray     # int64 (n,) Cartesian AO of first
    ao_second: np.ndarray    # int64 (n,)
    lower_diag: np.ndarray   # bool (n,) entry below the diagonal of a diagonal block
    qmax_blk: np.ndarray     # (nblk,)

    @property
    def nblk(self) -> int:
        return int(self.qmax_blk.size)

    @property
    def count(self) -> int:
        return int(self.first.size)

    def mat_D(self, D_cart: np.ndarray) -> np.ndarray:
        d = D_cart[self.ao_first, self.ao_second]
        d = np.where(self.lower_diag, 0.0, d)
        return np.ascontiguousarray(d)

    def ket_order(self, mat_D: np.ndarray) -> np.ndarray:
        """Block permutation that orders blocks by non-increasing beta_KL."""
        nab = NCART[self.la] * NCART[self.lb]
        beta = np.abs(self.mat_Q * mat_D).reshape(self.nblk, nab).max(axis=1) if self.nblk else np.zeros(0)
        return np.argsort(-beta, kind="stable")

    def permuted(self, blk_perm: np.ndarray):
        """Entry permutation for a block permutation."""
        nab = NCART[self.la] * NCART[self.lb]
        return (blk_perm[:, None] * nab + np.arange(nab)[None, :]).ravel()


@dataclass
class KCSR:
    """CSR over combined indices i of type a, partners of type b."""
    la: int
    lb: int
    displs: np.ndarray       # uint32 (count_a * na,)
    counts: np.ndarray       # uint32 (count_a * na,)
    Q: np.ndarray            # float64 (nnz,)
    D_inds: np.ndarray       # uint32 (nnz,)
    pair_data: np.ndarray    # float64 (nnz,)
    qmax_prim: np.ndarray    # (count_a,) max QKmax over partners of I
    blk_max: np.ndarray = None      # (nblocks,) QKmax_IJ of every partner block, CSR order
    blk_displs: np.ndarray = None   # (count_a,) first partner block of I in blk_max
    blk_counts: np.ndarray = None   # (count_a,) number of partner blocks of I


@dataclass
class IKList:
    la: int
    lc: int
    i: np.ndarray            # uint32 (n,)
    k: np.ndarray            # uint32 (n,)
    I: np.ndarray            # (nblk,)
    K: np.ndarray            # (nblk,)

    @property
    def nblk(self) -> int:
        return int(self.I.size)


@dataclass
class FockLists:
    """Everything the J/K kernels read for one molecule (D-independent part)."""
    mol: Molecule
    prims: list
    pairsets: dict = field(default_factory=dict)     # "ab" -> PairSet
    jlists: dict = field(default_factory=dict)       # "ab" -> JList
    kcsr: dict = field(default_factory=dict)         # "ab" (ordered, all 9) -> KCSR
    iklists: dict = field(default_factory=dict)      # "ac" -> IKList
    qglob: float = 0.0
    pair_cut: float = 0.0
    thr: float = 1e-12
    d_bound: float = 0.0

    @property
    def naos(self) -> int:
        return self.mol.naos_cart


class DensityError(ValueError):
    """A density the resident lists cannot serve exactly (see check_density)."""


# relative asymmetry |D - D^T|_max / max|D| accepted as rounding of a symmetric density
SYM_RTOL = 1e-12


def check_density(fl: FockLists, max_abs_cart: float, asym: float, max_abs: float):
    """Reject a density the lists of ``fl`` cannot serve with the reference's
    screening.

    * ``max_abs_cart`` = max|D| in the Cartesian AO basis must not exceed
      ``fl.d_bound``: the lists dropped every pair block whose bound
      Qmax * Qglob * d_bound is below the threshold, which is exact only while
      max|D| <= d_bound (the reference screens per quartet with the actual D,
      gen_coulomb.py:209-212, gen_exchange.py:229-232).
    * ``asym`` = max|D - D^T| must be rounding-level (relative SYM_RTOL of
      ``max_abs`` = max|D|): the engines symmetrise D and some exchange classes
      read D_jl as D[l][j], so a non-symmetric density would silently give
      the exchange matrix of its symmetric part."""
    if not math.isfinite(max_abs_cart) or max_abs_cart > fl.d_bound:
        raise DensityError("max|D| = %.4g exceeds the bound %.4g the pair lists were screened for; "
                           "rebuild the engine with a larger d_bound" % (max_abs_cart, fl.d_bound))
    if not math.isfinite(asym) or asym > SYM_RTOL * max(max_abs, 1e-300):
        raise DensityError("density is not symmetric (max|D - D^T| = %.3g, max|D| = %.3g); "
                           "J/K of a non-symmetric density are not supported" % (asym, max_abs))


def build_lists(mol: Molecule, schwarz_fn, thr=1e-12, d_bound=100.0) -> FockLists:
    """Build the D-independent lists of ``mol``.

    ``schwarz_fn(la, lb, info_a, info_b, I, J, S) -> Q`` evaluates the Schwarz
    factors of the candidate blocks.  A block is kept when
    ``Qmax_IJ * Qglob * d_bound > thr`` (no quartet of a dropped block can pass
    either screening rule while ``max|D| <= d_bound``)."""
    prims = build_prim_tables(mol)
    fl = FockLists(mol=mol, prims=prims, thr=thr, d_bound=d_bound)
    raw = {}
    for pc in PAIR_CLASSES:
        la, lb = LVAL[pc[0]], LVAL[pc[1]]
        ta, tb = prims[la], prims[lb]
        if ta is None or tb is None:
            continue
        I, J = candidate_blocks(ta, tb, la == lb)
        S = pair_overlap_factors(ta, tb, I, J)
        Q = schwarz_fn(la, lb, ta.info, tb.info, I.astype(np.uint32), J.astype(np.uint32),
                       np.ascontiguousarray(S))
        Q = np.asarray(Q, dtype=np.float64).reshape(S.shape)
        raw[pc] = (I, J, S, Q)
    qglob = max((float(np.abs(v[3]).max()) for v in raw.values() if v[3].size), default=0.0)
    fl.qglob = qglob
    fl.pair_cut = thr / (qglob * d_bound) if qglob > 0 else 0.0
    for pc, (I, J, S, Q) in raw.items():
        la, lb = LVAL[pc[0]], LVAL[pc[1]]
        qm = np.abs(Q).reshape(I.size, -1).max(axis=1)
        keep = np.nonzero(qm > fl.pair_cut)[0]
        # bra order: non-increasing Qmax, so that lanes of one wavefront leave
        # their ket scans at similar points (the order is free for the bra role)
        keep = keep[np.argsort(-qm[keep], kind="stable")]
        fl.pairsets[pc] = PairSet(la, lb, I[keep], J[keep], S[keep], Q[keep])
    for pc, ps in fl.pairsets.items():
        fl.jlists[pc] = _make_jlist(ps, prims)
    _make_kcsr(fl)
    _make_iklists(fl)
    return fl


def _make_jlist(ps: PairSet, prims) -> JList:
    la, lb = ps.la, ps.lb
    na, nb = NCART[la], NCART[lb]
    ta, tb = prims[la], prims[lb]
    n = ps.nblk
    ca = np.arange(na)[None, :, None]
    cb = np.arange(nb)[None, None, :]
    first = (ps.I[:, None, None] * na + ca) + 0 * cb
    second = (ps.J[:, None, None] * nb + cb) + 0 * ca
    ao_first = ta.aoinds[ca + 0 * cb, ps.I[:, None, None]].astype(np.int64)
    ao_second = tb.aoinds[cb + 0 * ca, ps.J[:, None, None]].astype(np.int64)
    lower = np.zeros((n, na, nb), dtype=bool)
    if la == lb:
        diag = (ps.I == ps.J)[:, None, None]
        lower = diag & (ca > cb)
    return JList(la, lb,
                 np.ascontiguousarray(first.ravel(), dtype=np.uint32),
                 np.ascontiguousarray(second.ravel(), dtype=np.uint32),
                 np.ascontiguousarray(ps.S.ravel()), np.ascontiguousarray(ps.Q.ravel()),
                 ao_first.ravel(), ao_second.ravel(), lower.ravel(), ps.qmax)


def _ordered_blocks(fl: FockLists, la: int, lb: int):
    """All ordered (I, J) blocks of types (a, b) with S, Q oriented (na, nb)."""
    if la <= lb:
        ps = fl.pairsets.get("spd"[la] + "spd"[lb])
        if ps is None:
            return None
        if la == lb:
            off = ps.I != ps.J
            I = np.concatenate([ps.I, ps.J[off]])
            J = np.concatenate([ps.J, ps.I[off]])
            S = np.concatenate([ps.S, ps.S[off].transpose(0, 2, 1)])
            Q = np.concatenate([ps.Q, ps.Q[off].transpose(0, 2, 1)])
            return I, J, S, Q
        return ps.I, ps.J, ps.S, ps.Q
    ps = fl.pairsets.get("spd"[lb] + "spd"[la])
    if ps is None:
        return None
    return ps.J, ps.I, ps.S.transpose(0, 2, 1), ps.Q.transpose(0, 2, 1)


def _make_kcsr(fl: FockLists):
    prims = fl.prims
    for la in range(3):
        for lb in range(3):
            if prims[la] is None or prims[lb] is None:
                continue
            ob = _ordered_blocks(fl, la, lb)
            na, nb = NCART[la], NCART[lb]
            cnt_a = prims[la].count
            if ob is None:
                continue
            I, J, S, Q = ob
            qm = np.abs(Q).reshape(I.size, -1).max(axis=1) if I.size else np.zeros(0)
            # group by I, order partners by non-increasing qm (ties: J ascending)
            order = np.lexsort((J, -qm, I))
            I, J, S, Q, qm = I[order], J[order], S[order], Q[order], qm[order]
            nJ = np.bincount(I, minlength=cnt_a)
            start = np.concatenate([[0], np.cumsum(nJ)[:-1]])
            # entries: for I: for ca: for J in order: for cb
            seg_len = nJ * nb                       # per component segment
            base = np.concatenate([[0], np.cumsum(seg_len * na)[:-1]])
            displs = (base[:, None] + np.arange(na)[None, :] * seg_len[:, None]).ravel()
            counts = np.repeat(seg_len, na)
            nnz = int((seg_len * na).sum())
            Qe = np.zeros(nnz)
            Se = np.zeros(nnz)
            De = np.zeros(nnz, dtype=np.uint32)
            if I.size:
                pos_in_I = np.arange(I.size) - start[I]          # partner rank within I
                for ca in range(na):
                    for cb in range(nb):
                        idx = base[I] + ca * seg_len[I] + pos_in_I * nb + cb
                        Qe[idx] = Q[:, ca, cb]
                        Se[idx] = S[:, ca, cb]
                        De[idx] = (J * nb + cb).astype(np.uint32)
            qmax_prim = np.zeros(cnt_a)
            if I.size:
                np.maximum.at(qmax_prim, I, qm)
            fl.kcsr["spd"[la] + "spd"[lb]] = KCSR(
                la, lb, np.ascontiguousarray(displs, dtype=np.uint32),
                np.ascontiguousarray(counts, dtype=np.uint32), Qe, De, Se, qmax_prim,
                np.ascontiguousarray(qm), np.ascontiguousarray(start, dtype=np.uint32),
                np.ascontiguousarray(nJ, dtype=np.uint32))


def _make_iklists(fl: FockLists):
    prims = fl.prims
    qprim = []
    for la in range(3):
        if prims[la] is None:
            qprim.append(None)
            continue
        q = np.zeros(prims[la].count)
        for lb in range(3):
            cs = fl.kcsr.get("spd"[la] + "spd"[lb])
            if cs is not None:
                q = np.maximum(q, cs.qmax_prim)
        qprim.append(q)
    for pc in PAIR_CLASSES:
        la, lc = LVAL[pc[0]], LVAL[pc[1]]
        if prims[la] is None or prims[lc] is None:
            continue
        I, K = candidate_blocks(prims[la], prims[lc], la == lc)
        keep = qprim[la][I] * qprim[lc][K] * fl.d_bound > fl.thr
        I, K = I[keep], K[keep]
        na, nc = NCART[la], NCART[lc]
        ii = (I[:, None, None] * na + np.arange(na)[None, :, None]) + 0 * np.arange(nc)[None, None, :]
        kk = (K[:, None, None] * nc + np.arange(nc)[None, None, :]) + 0 * np.arange(na)[None, :, None]
        fl.iklists[pc] = IKList(la, lc, np.ascontiguousarray(ii.ravel(), dtype=np.uint32),
                                np.ascontiguousarray(kk.ravel(), dtype=np.uint32), I, K)


# --------------------------------------------------------------------------
# Class enumeration (gen_coulomb.py:47-56, new_exchange_code/gen_eri_exchange.py:39-58)
# --------------------------------------------------------------------------

def j_classes():
    """36 J classes (ab|cd), ab, cd in {ss,sp,sd,pp,pd,dd}."""
    return [(ab, cd) for ab in PAIR_CLASSES for cd in PAIR_CLASSES]


def k_classes():
    """54 K classes (a b c d) with a <= c and any b, d."""
    out = []
    for ac in PAIR_CLASSES:
        for b in "spd":
            for d in "spd":
                out.append(ac[0] + b + ac[1] + d)
    return out


def max_abs_D_block(D_cart: np.ndarray, prims, lb: int, ld: int) -> float:
    tb, td = prims[lb], prims[ld]
    if tb is None or td is None:
        return 0.0
    ab = np.unique(tb.aoinds.ravel())
    ad = np.unique(td.aoinds.ravel())
    return float(np.abs(D_cart[np.ix_(ab, ad)]).max())


# --------------------------------------------------------------------------
# AO assembly (the step after the kernels; VeloxChem's is not in the reference)
# --------------------------------------------------------------------------

def assemble_J(fl: FockLists, matJ: dict) -> np.ndarray:
    """Cartesian AO J from per-class ``mat_J`` (bra order = jlists order)."""
    n = fl.naos
    J = np.zeros((n, n))
    for pc, jl in fl.jlists.items():
        v = matJ.get(pc)
        if v is None:
            continue
        use = ~jl.lower_diag
        f, s, x = jl.ao_first[use], jl.ao_second[use], v[use]
        np.add.at(J, (f, s), x)
        off = jl.first[use] != jl.second[use] if jl.la == jl.lb else np.ones(f.size, dtype=bool)
        np.add.at(J, (s[off], f[off]), x[off])
    return J


def assemble_K(fl: FockLists, matK: dict) -> np.ndarray:
    """Cartesian AO K from per-class ``mat_K`` (summed over the 9 bd classes)."""
    n = fl.naos
    K = np.zeros((n, n))
    for pc, ik in fl.iklists.items():
        v = matK.get(pc)
        if v is None:
            continue
        la, lc = ik.la, ik.lc
        ta, tc = fl.prims[la], fl.prims[lc]
        na, nc = NCART[la], NCART[lc]
        ai = ta.aoinds[ik.i % na, ik.i // na].astype(np.int64)
        ak = tc.aoinds[ik.k % nc, ik.k // nc].astype(np.int64)
        np.add.at(K, (ai, ak), v)
        if la != lc:
            np.add.at(K, (ak, ai), v)
        else:
            Iblk = np.repeat(ik.I, na * nc)
            Kblk = np.repeat(ik.K, na * nc)
            off = Iblk != Kblk
            np.add.at(K, (ak[off], ai[off]), v[off])
    return K
