"""The device-side pair-list builder (vlxmi.devpairs) against the host builder
(vlxmi.pairs.build_lists with the same GPU Schwarz kernels).

Both builders apply the same screening and the same stable sort keys, so the
lists must hold the same blocks in the same order; the overlap factors are
computed by the device's pow/exp, which may differ from the host's in the last
bit, so values are compared to 1e-14 relative and an order difference is
allowed only between blocks whose sort keys agree to that tolerance.  Then the
J/K build from device-built lists equals the build from host-built lists
(1e-11) and, on the 8-atom workload, the oracle."""

import time

import numpy as np
import pytest

import _util as U
from vlxmi import configs
from vlxmi import pairs as P

pytestmark = pytest.mark.gpu
RTOL = 1e-14


def _say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def dev():
    from vlxmi.runtime import Device
    return Device()


def _close(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return a.shape == b.shape and (a.size == 0 or np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(b), 1e-300) + 1e-300))


def _same_order(keys_h, keys_d, ids_h, ids_d):
    """Identical ids, or differences only inside runs of equal keys."""
    if np.array_equal(ids_h, ids_d):
        return True
    diff = np.nonzero(np.any(ids_h != ids_d, axis=-1) if ids_h.ndim > 1 else ids_h != ids_d)[0]
    return bool(np.all(np.abs(keys_h[diff] - keys_d[diff]) <= 1e-13 * np.abs(keys_h[diff])))


def _compare(fh, fd):
    assert fh.qglob == pytest.approx(fd.qglob, rel=RTOL)
    assert set(fh.jlists) == set(fd.jlists) and set(fh.kcsr) == set(fd.kcsr) and set(fh.iklists) == set(fd.iklists)
    for pc, a in fh.jlists.items():
        b = fd.jlists[pc]
        assert a.count == b.count, pc
        ida = np.stack([a.first, a.second], 1)
        idb = np.stack([b.first, b.second], 1)
        nab = a.count // max(a.nblk, 1)
        assert _same_order(np.repeat(a.qmax_blk, nab), np.repeat(b.qmax_blk, nab), ida, idb), pc
        if np.array_equal(ida, idb):
            assert _close(b.pair_data, a.pair_data) and _close(b.mat_Q, a.mat_Q), pc
            assert np.array_equal(a.ao_first, b.ao_first) and np.array_equal(a.lower_diag, b.lower_diag), pc
    for pc, a in fh.kcsr.items():
        b = fd.kcsr[pc]
        assert np.array_equal(a.displs, b.displs) and np.array_equal(a.counts, b.counts), pc
        # partner blocks in the same order, up to swaps of blocks with equal sort keys
        assert _close(b.blk_max, a.blk_max), pc
        seg = np.repeat(np.arange(a.counts.size), a.counts.astype(np.int64))
        oa, ob = np.lexsort((a.D_inds, seg)), np.lexsort((b.D_inds, seg))
        assert np.array_equal(a.D_inds[oa], b.D_inds[ob]), pc
        if np.array_equal(a.D_inds, b.D_inds):
            assert _close(b.Q, a.Q) and _close(b.pair_data, a.pair_data), pc
        assert _close(b.qmax_prim, a.qmax_prim), pc
    for pc, a in fh.iklists.items():
        b = fd.iklists[pc]
        assert np.array_equal(a.i, b.i) and np.array_equal(a.k, b.k), pc


@pytest.mark.parametrize("wl", ["small8", "lattice4", "lattice64", "batch207"])
def test_device_lists_equal_host_lists(dev, wl):
    from vlxmi.devpairs import build_lists_device
    mol, D = configs.make_batch_molecule(207) if wl == "batch207" else configs.make_workload(wl)
    _say("case: device pair builder vs host builder, %s (%d atoms)" % (wl, mol.natoms))
    t0 = time.perf_counter()
    fh = P.build_lists(mol, dev.schwarz)
    th = time.perf_counter() - t0
    dev.torch.cuda.synchronize()
    t0 = time.perf_counter()
    fd = build_lists_device(mol, dev)
    dev.torch.cuda.synchronize()
    td = time.perf_counter() - t0
    _say("  host builder %.3f s, device builder %.3f s" % (th, td))
    _compare(fh, fd)


@pytest.mark.parametrize("wl", ["small8", "lattice64"])
def test_fock_build_from_device_lists(dev, wl):
    from vlxmi.devpairs import build_lists_device
    from vlxmi.fock import FockEngine
    mol, D = configs.make_workload(wl)
    _say("case: J/K from device-built lists vs host-built lists, %s" % wl)
    eh = FockEngine(mol, dev=dev)
    ed = FockEngine(mol, dev=dev, lists=build_lists_device(mol, dev))
    Jh, Kh = eh.build(D)
    Jd, Kd = ed.build(D)
    assert float((Jh - Jd).abs().max()) <= 1e-11 and float((Kh - Kd).abs().max()) <= 1e-11
    if wl == "small8":
        Dc = U.cart_density(mol, D)
        Jr, Kr = U.oracle_fock(ed.fl, Dc, "component")
        assert np.abs(Jd.cpu().numpy() - Jr).max() <= 1e-11 and np.abs(Kd.cpu().numpy() - Kr).max() <= 1e-11


def test_device_builder_time_lattice256(dev):
    """Set-up time of the 256-atom lattice: host builder vs device builder
    (reported; the lists must agree as above)."""
    from vlxmi.devpairs import build_lists_device
    mol, D = configs.make_lattice256()
    t0 = time.perf_counter()
    fh = P.build_lists(mol, dev.schwarz)
    th = time.perf_counter() - t0
    dev.torch.cuda.synchronize()
    t0 = time.perf_counter()
    fd = build_lists_device(mol, dev)
    dev.torch.cuda.synchronize()
    td = time.perf_counter() - t0
    _say("case: lattice256 lists: host builder %.2f s, device builder %.2f s" % (th, td))
    _compare(fh, fd)
