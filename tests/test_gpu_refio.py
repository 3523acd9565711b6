"""Lists in the reference's order on the GPU (lattice4: every angular type up
to d).

* With list validation on, the class launchers accept block-order lists and
  refuse reference-order lists (VLXMI_ERR_LAYOUT) instead of reading them
  misaligned.
* Reference-order lists converted by the library's converters, launched, and
  the block-order results added back in the caller's order
  (vlxmi_scatter_add) give the oracle's mat_J / mat_K per reference entry to
  1e-11.
"""

import numpy as np
import pytest

import _reforder as R
import _util as U
from vlxmi import configs
from vlxmi import pairs as P
from vlxmi import refio

pytestmark = pytest.mark.gpu
TOL = 1e-11


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def lat4():
    from vlxmi.fock import FockEngine
    from vlxmi.runtime import Device
    mol, D = configs.make_lattice4()
    eng = FockEngine(mol, dev=Device())
    return eng, U.cart_density(mol, D), eng.cart_density(D)


@pytest.fixture
def validation():
    refio.set_validation(True)
    yield
    refio.set_validation(False)


def _dev_lists(eng, d):
    return {k: eng.dev.tensor(v) for k, v in d.items() if isinstance(v, np.ndarray)}


def _launch_j(eng, cls, bra, ket, out):
    from vlxmi.runtime import lib
    args = [out.data_ptr()]
    for x in "spd":
        if x in cls:
            l = P.LVAL[x]
            args += [eng.prim_info[l].data_ptr(), eng.prim_count[l]]
    args += [ket["mat_D"].data_ptr(), bra["mat_Q"].data_ptr(), ket["mat_Q"].data_ptr(), bra["first"].data_ptr(),
             bra["second"].data_ptr(), bra["pair_data"].data_ptr(), int(bra["first"].numel()),
             ket["first"].data_ptr(), ket["second"].data_ptr(), ket["pair_data"].data_ptr(),
             int(ket["first"].numel()), eng.dev.boys_table.data_ptr(), eng.dev.boys_ft.data_ptr(), eng.thr]
    return getattr(lib(), "vlxmi_launch_computeCoulombFock" + cls.upper())(*args, 4, eng.dev.stream_ptr())


def _launch_k(eng, cls, ik, A, Cc, Dc, maxD, out):
    from vlxmi.runtime import lib
    a, b, c, d = cls
    args = [out.data_ptr(), ik["i"].data_ptr(), ik["k"].data_ptr(), int(ik["i"].numel())]
    for x in "spd":
        if x in cls:
            l = P.LVAL[x]
            args += [eng.prim_info[l].data_ptr(), eng.prim_aoinds[l].data_ptr(), eng.prim_count[l]]
    args += [maxD, Dc.data_ptr(), eng.naos]
    for nm in ("Q", "D_inds", "displs", "counts", "pair_data"):
        args.append(A[nm].data_ptr())
        if c + d != a + b:
            args.append(Cc[nm].data_ptr())
    args += [eng.dev.boys_table.data_ptr(), eng.dev.boys_ft.data_ptr(), 0.0, eng.thr]
    return getattr(lib(), "vlxmi_launch_computeExchangeFock" + cls.upper())(*args, eng.dev.stream_ptr())


@pytest.mark.parametrize("ab,cd", [("pp", "pp"), ("sp", "pd"), ("pd", "dd"), ("dd", "dd")])
def test_reference_order_j(lat4, validation, ab, cd):
    eng, Dc_np, Dc = lat4
    t = eng.torch
    fl = eng.fl
    rng = np.random.default_rng(21)
    bl, kl = fl.jlists[ab], fl.jlists[cd]
    bra_ref, _ = R.j_lists(bl, Dc_np, rng)
    _, ket_ref = R.j_lists(kl, Dc_np, rng)
    _say("case: reference-order J %s%s: direct launch refused, converted launch vs oracle" % (ab, cd))
    direct = t.zeros(bra_ref["first"].size, dtype=t.float64, device=eng.dev.device)
    assert _launch_j(eng, ab + cd, _dev_lists(eng, bra_ref), _dev_lists(eng, ket_ref), direct) == -2
    bra = refio.j_list_to_blocks(bl.la, bl.lb, bra_ref["first"], bra_ref["second"], bra_ref["pair_data"],
                                 bra_ref["mat_Q"])
    ket = refio.j_list_to_blocks(kl.la, kl.lb, ket_ref["first"], ket_ref["second"], ket_ref["pair_data"],
                                 ket_ref["mat_Q"], ket_ref["mat_D"])
    blk = t.zeros(bra["first"].size, dtype=t.float64, device=eng.dev.device)
    assert _launch_j(eng, ab + cd, _dev_lists(eng, bra), _dev_lists(eng, ket), blk) == 0
    got = t.zeros(bra_ref["first"].size, dtype=t.float64, device=eng.dev.device)
    refio.scatter_add(eng.dev, blk, eng.dev.tensor(bra["ref_index"]), got)
    want = U.oracle_matJ(fl, Dc_np, "component", classes=[(ab, cd)])[ab][bra_ref["index"]]
    assert np.max(np.abs(got.cpu().numpy() - want)) <= TOL


@pytest.mark.parametrize("cls", ["spsp", "ppdd", "sdpd", "dddd"])
def test_reference_order_k(lat4, validation, cls):
    eng, Dc_np, Dc = lat4
    t = eng.torch
    fl = eng.fl
    rng = np.random.default_rng(22)
    a, b, c, d = cls
    la, lb, lc, ld = (P.LVAL[x] for x in cls)
    ik_ref = R.ik_list(fl.iklists[a + c], rng)
    A_ref = R.k_csr(fl.kcsr[a + b], fl.prims[la].count)
    C_ref = R.k_csr(fl.kcsr[c + d], fl.prims[lc].count)
    maxD = eng.max_D(Dc)[b + d]
    _say("case: reference-order K %s: direct launch refused, converted launch vs oracle" % cls)
    direct = t.zeros(ik_ref["i"].size, dtype=t.float64, device=eng.dev.device)
    rc = _launch_k(eng, cls, _dev_lists(eng, ik_ref), _dev_lists(eng, A_ref), _dev_lists(eng, C_ref), Dc, maxD,
                   direct)
    assert rc == -2
    ik = refio.ik_list_to_blocks(la, lc, ik_ref["i"], ik_ref["k"])
    A = refio.k_csr_to_blocks(la, lb, fl.prims[la].count, **A_ref)
    Cc = refio.k_csr_to_blocks(lc, ld, fl.prims[lc].count, **C_ref)
    blk = t.zeros(ik["i"].size, dtype=t.float64, device=eng.dev.device)
    assert _launch_k(eng, cls, _dev_lists(eng, ik), _dev_lists(eng, A), _dev_lists(eng, Cc), Dc, maxD, blk) == 0
    got = t.zeros(ik_ref["i"].size, dtype=t.float64, device=eng.dev.device)
    refio.scatter_add(eng.dev, blk, eng.dev.tensor(ik["ref_index"]), got)
    want = U.oracle_matK(fl, Dc_np, "component", classes=[cls])[a + c][ik_ref["index"]]
    assert np.max(np.abs(got.cpu().numpy() - want)) <= TOL


def test_block_lists_pass_validation(lat4, validation):
    """The engine's own lists pass the checks (every class of lattice4)."""
    eng, _, Dc = lat4
    _say("case: validation on, all lattice4 classes through the engine")
    eng.run_J(Dc)
    eng.run_K(Dc)


def test_split_series_equals_class(lat4):
    """VeloxChem launches J PDDD as computeCoulombFockPDDD0..6 and K DDDD as
    computeExchangeFockDDDD0..18 into one buffer; the series' launchers give
    the class launcher's result."""
    from vlxmi.runtime import lib
    eng, _, Dc = lat4
    t = eng.torch
    _say("case: split series J PDDD0..6 and K DDDD0..18 vs the class launchers")
    ref = eng.run_J(Dc, classes=[("pd", "dd")])["pd"]
    kets = eng.prepare_kets(Dc)
    bra, ket = eng.jl["pd"], kets["dd"]
    out = t.zeros_like(ref)
    args = [out.data_ptr(), eng.prim_info[1].data_ptr(), eng.prim_count[1], eng.prim_info[2].data_ptr(),
            eng.prim_count[2], ket["mat_D"].data_ptr(), bra["mat_Q"].data_ptr(), ket["mat_Q"].data_ptr(),
            bra["first"].data_ptr(), bra["second"].data_ptr(), bra["pair_data"].data_ptr(), bra["count"],
            ket["first"].data_ptr(), ket["second"].data_ptr(), ket["pair_data"].data_ptr(), ket["count"],
            eng.dev.boys_table.data_ptr(), eng.dev.boys_ft.data_ptr(), eng.thr]
    split = eng._ket_split(bra["nblk"], eng.jl["dd"]["nblk"])
    for k in range(7):
        assert getattr(lib(), "vlxmi_launch_computeCoulombFockPDDD%d" % k)(*args, split, eng.dev.stream_ptr()) == 0
    assert t.equal(out, ref)
    refK = eng.run_K(Dc, classes=["dddd"])["dd"]
    maxD = eng.max_D(Dc)["dd"]
    outK = t.zeros_like(refK)
    ik, A = eng.ik["dd"], eng.kc["dd"]
    kargs = [outK.data_ptr(), ik["i"].data_ptr(), ik["k"].data_ptr(), ik["count"], eng.prim_info[2].data_ptr(),
             eng.prim_aoinds[2].data_ptr(), eng.prim_count[2], maxD, Dc.data_ptr(), eng.naos]
    kargs += [A[nm].data_ptr() for nm in ("Q", "D_inds", "displs", "counts", "pair_data")]
    kargs += [eng.dev.boys_table.data_ptr(), eng.dev.boys_ft.data_ptr(), 0.0, eng.thr]
    for k in range(19):
        assert getattr(lib(), "vlxmi_launch_computeExchangeFockDDDD%d" % k)(*kargs, eng.dev.stream_ptr()) == 0
    assert t.equal(outK, refK)
