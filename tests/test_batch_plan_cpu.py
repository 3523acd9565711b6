"""Batch assembly plans built from per-molecule plans (host logic, on the CPU).

A batch's AO assembly plan is the concatenation of each molecule's own sorted
plan with its value indices moved into the batch's per-class buffer slices.
These tests build MolEntry/Pack on a CPU stand-in device (CPU tensors, the
oracle's Schwarz factors; no kernel runs) and check that plan against a direct
whole-batch construction: every output element must receive the same values
in the same order, which is what makes a molecule's J/K the same in any batch."""

import numpy as np
import pytest
import torch

import _util as U
from vlxmi import configs
from vlxmi import pairs as P
from vlxmi.basis import NCART


class CpuDevice:
    """The parts of vlxmi.runtime.Device that MolEntry and Pack use, on the CPU."""

    def __init__(self):
        self.torch = torch
        self.device = torch.device("cpu")

    def tensor(self, a, dtype=None):
        a = np.ascontiguousarray(a)
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        t = torch.from_numpy(a.copy())
        return t.to(dtype) if dtype is not None else t

    def schwarz(self, la, lb, info_a, info_b, I, J, S):
        return U.oracle_schwarz(la, lb, info_a, info_b, I, J, S)


def _direct_plan(pack):
    """(value index, output) pairs of the whole batch, sorted by (output, value):
    the plan a single host sort over the batch gives."""
    E = pack.entries
    out = {}
    for kind in ("J", "K"):
        order = pack.j_order if kind == "J" else pack.k_order
        lists = pack.jl if kind == "J" else pack.ik
        vi, oi = [], []
        base = 0
        for pc in order:
            for m, e in enumerate(E):
                fl = e.fl
                n = e.naos
                if kind == "J":
                    if pc not in fl.jlists or fl.jlists[pc].nblk == 0:
                        continue
                    jl = fl.jlists[pc]
                    use = ~jl.lower_diag
                    offd = (jl.first != jl.second) if jl.la == jl.lb else np.ones(jl.count, dtype=bool)
                    s1, s2 = np.nonzero(use)[0], np.nonzero(use & offd)[0]
                    o1 = (jl.ao_first * n + jl.ao_second)[s1]
                    o2 = (jl.ao_second * n + jl.ao_first)[s2]
                else:
                    if pc not in fl.iklists or fl.iklists[pc].nblk == 0:
                        continue
                    ik = fl.iklists[pc]
                    na, nc = NCART[ik.la], NCART[ik.lc]
                    ai = fl.prims[ik.la].aoinds[ik.i % na, ik.i // na].astype(np.int64)
                    ak = fl.prims[ik.lc].aoinds[ik.k % nc, ik.k // nc].astype(np.int64)
                    s1 = np.arange(ik.i.size)
                    s2 = s1 if ik.la != ik.lc else np.nonzero(np.repeat(ik.I, na * nc) != np.repeat(ik.K, na * nc))[0]
                    o1, o2 = ai * n + ak, ak[s2] * n + ai[s2]
                off = base + lists[pc]["eoff"][m]
                vi += [s1 + off, s2 + off]
                oi += [o1 + pack.D_off[m], o2 + pack.D_off[m]]
            base += lists[pc]["count"]
        vi, oi = np.concatenate(vi), np.concatenate(oi)
        srt = np.lexsort((vi, oi))
        out[kind] = (vi[srt], oi[srt])
    return out


def _expand(asm):
    """(value index, output) pairs of an Assembler's plan, in plan order."""
    perm = asm.perm.numpy().astype(np.int64)
    off = asm.seg_off.numpy().astype(np.int64)
    outs = asm.seg_out.numpy().astype(np.int64)
    return perm, np.repeat(outs, np.diff(off))


@pytest.fixture(scope="module")
def pack():
    from vlxmi.batch import MolEntry, Pack
    dev = CpuDevice()
    mols = [configs.make_batch_molecule(m, hi=12)[0] for m in range(4)]
    mols.append(configs.make_small8()[0])
    entries = [MolEntry(mol, dev, builder="host") for mol in mols]
    return Pack(entries, dev)


def test_pack_plan_equals_whole_batch_sort(pack):
    ref = _direct_plan(pack)
    for kind, asm in (("J", pack.jasm), ("K", pack.kasm)):
        v, o = _expand(asm)
        rv, ro = ref[kind]
        # the batch plan groups molecules; within an output the order must match
        srt = np.lexsort((np.arange(v.size), o))
        assert np.array_equal(o[srt], ro) and np.array_equal(v[srt], rv), kind
        assert asm.nseg == np.unique(ro).size


def test_molecule_plan_independent_of_batch(pack):
    """The same molecule, alone and as the second member of another batch,
    sums the same per-pair outputs into each AO element in the same order."""
    from vlxmi.batch import Pack
    dev = pack.dev
    e = pack.entries[4]
    alone = Pack([e], dev)
    other = Pack([pack.entries[1], e], dev)
    for kind, lists in (("J", "jl"), ("K", "ik")):
        va, oa = _expand(alone.jasm if kind == "J" else alone.kasm)
        vb, ob = _expand(other.jasm if kind == "J" else other.kasm)
        sel = ob >= other.D_off[1]
        vb, ob = vb[sel], ob[sel] - other.D_off[1]
        # map batch value indices back to the molecule's own per-class positions
        order = other.j_order if kind == "J" else other.k_order
        L = getattr(other, lists)
        La = getattr(alone, lists)
        starts = np.cumsum([0] + [L[pc]["count"] for pc in order])
        starts_a = {pc: s for pc, s in zip(alone.j_order if kind == "J" else alone.k_order,
                                           np.cumsum([0] + [La[pc]["count"] for pc in
                                                            (alone.j_order if kind == "J" else alone.k_order)]))}
        mapped = np.empty_like(vb)
        for i, pc in enumerate(order):
            w = (vb >= starts[i]) & (vb < starts[i + 1])
            if w.any():
                mapped[w] = vb[w] - starts[i] - L[pc]["eoff"][1] + starts_a[pc]
        assert np.array_equal(oa, ob) and np.array_equal(va, mapped), kind
