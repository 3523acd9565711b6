"""The allocation-free Coulomb entry point (vlxmi_coulomb_fock_noalloc) against
the class launchers (vlxmi_launch_computeCoulombFock{CLS}) it is a variant of:
same inputs, same ket split, the same bits; two calls in a row, the second on
a smaller bra list, so the static scratch is reused; and a ket split whose
partial sums would not fit the static scratch is refused before any launch."""

import ctypes as C

import numpy as np
import pytest

from vlxmi import configs
from vlxmi import pairs as P

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def lat64():
    from vlxmi.fock import FockEngine
    from vlxmi.runtime import Device
    mol, D = configs.make_lattice64()
    eng = FockEngine(mol, dev=Device())
    Dc = eng.cart_density(D)
    return eng, eng.prepare_kets(Dc)


def _args(eng, bra, ket, cls):
    """The class launcher's argument list (without ket_split / stream)."""
    args = [None]
    for x in "spd":
        if x in cls:
            l = P.LVAL[x]
            args += [eng.prim_info[l].data_ptr(), eng.prim_count[l]]
    args += [ket["mat_D"].data_ptr(), bra["mat_Q"].data_ptr(), ket["mat_Q"].data_ptr(),
             bra["first"].data_ptr(), bra["second"].data_ptr(), bra["pair_data"].data_ptr(), bra["count"],
             ket["first"].data_ptr(), ket["second"].data_ptr(), ket["pair_data"].data_ptr(), ket["count"],
             eng.dev.boys_table.data_ptr(), eng.dev.boys_ft.data_ptr(), eng.thr]
    return args


def _noalloc(L, eng, bra, ket, cls, out, split):
    info = (C.c_void_p * 3)(*[p.data_ptr() if p is not None else None for p in eng.prim_info])
    cnt = (C.c_uint32 * 3)(*eng.prim_count)
    return L.vlxmi_coulomb_fock_noalloc(cls.encode(), out.data_ptr(), info, cnt,
                                        ket["mat_D"].data_ptr(), bra["mat_Q"].data_ptr(), ket["mat_Q"].data_ptr(),
                                        bra["first"].data_ptr(), bra["second"].data_ptr(), bra["pair_data"].data_ptr(),
                                        bra["count"], ket["first"].data_ptr(), ket["second"].data_ptr(),
                                        ket["pair_data"].data_ptr(), ket["count"], eng.dev.boys_table.data_ptr(),
                                        eng.dev.boys_ft.data_ptr(), eng.thr, split, eng.dev.stream_ptr())


@pytest.mark.parametrize("cls", ["ssss", "sdpd", "ppdd"])
def test_noalloc_equals_launcher(lat64, cls):
    from vlxmi.runtime import lib
    eng, kets = lat64
    t = eng.torch
    L = lib()
    ab, cd = cls[:2], cls[2:]
    launcher = getattr(L, "vlxmi_launch_computeCoulombFock" + cls.upper())
    rng = np.random.default_rng(3)
    n = eng.fl.jlists[ab].nblk
    for nsel in (min(n, 40), min(n, 9)):          # the second call on a smaller bra list
        sel = np.sort(rng.choice(n, size=nsel, replace=False))
        bra = eng.bra_slice(ab, sel)
        ket = kets[cd]
        split = 256
        _say("case: noalloc %s, %d bra blocks, ket split %d" % (cls, nsel, split))
        # mat_J is accumulated (+=): both outputs start from the same values
        ref = t.full((bra["count"],), 0.25, dtype=t.float64, device=eng.dev.device)
        got = ref.clone()
        a = _args(eng, bra, ket, cls)
        a[0] = ref.data_ptr()
        assert launcher(*a, split, eng.dev.stream_ptr()) == 0
        assert _noalloc(L, eng, bra, ket, cls, got, split) == 0
        t.cuda.synchronize()
        assert t.equal(got, ref), cls
        assert float((ref - 0.25).abs().max()) > 0


def test_noalloc_refuses_oversized_split(lat64):
    from vlxmi.runtime import lib
    from vlxmi.runtime import check
    eng, kets = lat64
    t = eng.torch
    L = lib()
    cap = int(L.vlxmi_noalloc_scratch_doubles())
    bra = eng.jl["ss"]
    ket = kets["ss"]
    nket = ket["count"]
    split = min(nket, cap // bra["count"] + 1)
    assert split * bra["count"] > cap
    out = t.zeros(bra["count"], dtype=t.float64, device=eng.dev.device)
    _say("case: noalloc refuses ket split %d x %d entries (> %d)" % (split, bra["count"], cap))
    rc = _noalloc(L, eng, bra, ket, "ssss", out, split)
    assert rc == -3
    with pytest.raises(Exception, match="limit"):
        check(rc, "noalloc")
    assert float(out.abs().max()) == 0.0
