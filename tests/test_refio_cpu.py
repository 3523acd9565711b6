"""The host converters from the reference's list order to the block order
(vlxmi_j_list_to_blocks, vlxmi_ik_list_to_blocks, vlxmi_k_csr_to_blocks; host
code of the C-ABI library, no GPU needed).

Reference-order lists are made from this build's lists (tests/_reforder.py);
converted back, they must give the oracle's J and K on the original lists,
entry by entry in the caller's order, and the exchange CSR must come back
identical."""

import numpy as np
import pytest

import _reforder as R
import _util as U
import oracle
from vlxmi import configs
from vlxmi import pairs as P
from vlxmi import refio
from vlxmi.basis import NCART
from vlxmi.runtime import VlxmiError


@pytest.fixture(scope="module")
def lat4():
    mol, D = configs.make_lattice4()
    fl = P.build_lists(mol, U.oracle_schwarz)
    return mol, fl, U.cart_density(mol, D)


def _oracle_J(fl, la, lb, lc, ld, bra, ket, mat_D):
    infos = [t.info if t is not None else None for t in fl.prims]
    counts = [t.count if t is not None else 0 for t in fl.prims]
    out = np.zeros(bra["first"].size)
    oracle.coulomb_class(la, lb, lc, ld, infos, counts, mat_D, bra["mat_Q"], ket["mat_Q"], bra["first"],
                         bra["second"], bra["pair_data"], ket["first"], ket["second"], ket["pair_data"], fl.thr,
                         mode="component", mat_J=out)
    return out


def test_k_csr_comes_back_identical(lat4):
    mol, fl, _ = lat4
    for pc, cs in fl.kcsr.items():
        count_a = fl.prims[cs.la].count
        ref = R.k_csr(cs, count_a)
        got = refio.k_csr_to_blocks(cs.la, cs.lb, count_a, **ref)
        for key, want in (("displs", cs.displs), ("counts", cs.counts), ("Q", cs.Q), ("D_inds", cs.D_inds),
                          ("pair_data", cs.pair_data)):
            assert np.array_equal(got[key], want), (pc, key)


@pytest.mark.parametrize("ab,cd", [("ss", "ss"), ("sp", "pp"), ("pp", "pd"), ("pd", "dd"), ("dd", "dd")])
def test_j_reference_order_gives_oracle_j(lat4, ab, cd):
    mol, fl, Dc = lat4
    rng = np.random.default_rng(7)
    bra_l, ket_l = fl.jlists[ab], fl.jlists[cd]
    la, lb, lc, ld = bra_l.la, bra_l.lb, ket_l.la, ket_l.lb
    bra_ref, _ = R.j_lists(bra_l, Dc, rng)
    _, ket_ref = R.j_lists(ket_l, Dc, rng)
    bra = refio.j_list_to_blocks(la, lb, bra_ref["first"], bra_ref["second"], bra_ref["pair_data"], bra_ref["mat_Q"])
    ket = refio.j_list_to_blocks(lc, ld, ket_ref["first"], ket_ref["second"], ket_ref["pair_data"], ket_ref["mat_Q"],
                                 ket_ref["mat_D"])
    # block order: bra blocks by (I, J); ket blocks by non-increasing bound
    nab, ncd = NCART[la] * NCART[lb], NCART[lc] * NCART[ld]
    beta = np.abs(ket["mat_Q"] * ket["mat_D"]).reshape(-1, ncd).max(axis=1)
    assert np.all(np.diff(beta) <= 0)
    keys = bra["first"][::nab].astype(np.int64) * 10 ** 6 + bra["second"][::nab]
    assert np.all(np.diff(keys) > 0)
    got_blocks = _oracle_J(fl, la, lb, lc, ld, bra, ket, ket["mat_D"])
    got = np.zeros(bra_ref["first"].size)
    sel = bra["ref_index"] >= 0
    got[bra["ref_index"][sel]] = got_blocks[sel]
    want = U.oracle_matJ(fl, Dc, "component", classes=[(ab, cd)])[ab][bra_ref["index"]]
    assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("cls", ["ssss", "spsp", "ppdd", "sdpd", "dddd"])
def test_k_reference_order_gives_oracle_k(lat4, cls):
    mol, fl, Dc = lat4
    rng = np.random.default_rng(8)
    a, b, c, d = cls
    ac, ab, cd = a + c, a + b, c + d
    la, lb, lc, ld = (P.LVAL[x] for x in cls)
    ik_ref = R.ik_list(fl.iklists[ac], rng)
    ik = refio.ik_list_to_blocks(la, lc, ik_ref["i"], ik_ref["k"])
    A = refio.k_csr_to_blocks(la, lb, fl.prims[la].count, **R.k_csr(fl.kcsr[ab], fl.prims[la].count))
    Cc = refio.k_csr_to_blocks(lc, ld, fl.prims[lc].count, **R.k_csr(fl.kcsr[cd], fl.prims[lc].count))
    infos = [t.info if t is not None else None for t in fl.prims]
    aos = [t.aoinds if t is not None else None for t in fl.prims]
    counts = [t.count if t is not None else 0 for t in fl.prims]
    mx = P.max_abs_D_block(Dc, fl.prims, lb, ld)
    out = np.zeros(ik["i"].size)
    oracle.exchange_class(la, lb, lc, ld, ik["i"], ik["k"], infos, aos, counts, mx, Dc, fl.naos,
                          A["Q"], Cc["Q"], A["D_inds"], Cc["D_inds"], A["displs"], Cc["displs"], A["counts"],
                          Cc["counts"], A["pair_data"], Cc["pair_data"], 0.0, fl.thr, mode="component", mat_K=out)
    got = np.zeros(ik_ref["i"].size)
    sel = ik["ref_index"] >= 0
    got[ik["ref_index"][sel]] = out[sel]
    want = U.oracle_matK(fl, Dc, "component", classes=[cls])[ac][ik_ref["index"]]
    assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.abs(want).max())


def test_converters_refuse_other_orders():
    first = np.array([3, 1], np.uint32)          # a p-p list with i > j for one type
    second = np.array([0, 2], np.uint32)
    with pytest.raises(VlxmiError, match="block"):
        refio.j_list_to_blocks(1, 1, first, second, np.ones(2), np.ones(2))
    dup = np.array([0, 0], np.uint32)             # the same entry twice
    with pytest.raises(VlxmiError, match="block"):
        refio.j_list_to_blocks(0, 1, dup, np.array([1, 1], np.uint32), np.ones(2), np.ones(2))
    with pytest.raises(VlxmiError, match="block"):
        refio.ik_list_to_blocks(1, 1, np.array([5], np.uint32), np.array([1], np.uint32))
