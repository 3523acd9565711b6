"""Deterministic device-side AO assembly.

Every J (K) output element of the AO matrix is the sum of the per-pair kernel
outputs (mat_J / mat_K entries) that map to it.  ``Assembler`` turns the
static map (value index -> output index) into a plan sorted by output index
(ties by value index) once per molecule set; each build then runs one
``vlxmi_segment_sum`` launch: one thread per output element, contributions
summed in plan order, no atomics, so results are bit-reproducible and
independent of what else is assembled in the same launch.
"""

from __future__ import annotations

import numpy as np

from .runtime import check, lib


class Assembler:
    def __init__(self, dev, value_idx: np.ndarray, out_idx: np.ndarray, nout: int):
        value_idx = np.asarray(value_idx, dtype=np.int64)
        out_idx = np.asarray(out_idx, dtype=np.int64)
        assert value_idx.size < 2 ** 32 and nout < 2 ** 32
        order = np.lexsort((value_idx, out_idx))
        vo, oo = value_idx[order], out_idx[order]
        starts = np.nonzero(np.r_[True, oo[1:] != oo[:-1]])[0] if oo.size else np.zeros(0, dtype=np.int64)
        seg_off = np.r_[starts, oo.size].astype(np.uint32)
        seg_out = oo[starts].astype(np.uint32)
        self.dev = dev
        self.nseg = int(starts.size)
        self.nout = int(nout)
        self.perm = dev.tensor(vo.astype(np.uint32))
        self.seg_off = dev.tensor(seg_off)
        self.seg_out = dev.tensor(seg_out)
        # outputs that receive no contribution stay as the caller zeroed them
        self.covers_all = self.nseg == self.nout

    @classmethod
    def from_device(cls, dev, perm, seg_start, seg_out, nout):
        """A plan already sorted: ``perm`` value indices in plan order,
        ``seg_start`` (nseg + 1 entries) where each output's run begins and
        ends, ``seg_out`` the output of each run (device int64 tensors)."""
        t = dev.torch
        assert perm.numel() < 2 ** 31 and nout < 2 ** 31
        self = cls.__new__(cls)
        self.dev = dev
        self.nseg = int(seg_out.numel())
        self.nout = int(nout)
        self.perm = perm.to(t.int32).contiguous()
        self.seg_off = seg_start.to(t.int32).contiguous()
        self.seg_out = seg_out.to(t.int32).contiguous()
        self.covers_all = self.nseg == self.nout
        return self

    def run(self, values, out):
        """out[...] = assembled sums (elements without contributions untouched)."""
        check(lib().vlxmi_segment_sum(values.data_ptr(), self.perm.data_ptr(), self.seg_off.data_ptr(),
                                      self.seg_out.data_ptr(), self.nseg, out.data_ptr(), self.dev.stream_ptr()),
              "segment_sum")
        return out


def zero_(dev, t):
    check(lib().vlxmi_fill_zero(t.data_ptr(), t.numel(), dev.stream_ptr()), "fill_zero")
    return t


class FockAssembler:
    """The two-electron part of the Fock matrix, G = c_J J + c_K K, in one
    deterministic pass (``vlxmi_fock_assemble``) from the per-pair Coulomb and
    exchange outputs: the two value -> output maps are merged into one plan
    sorted by output (Coulomb contributions first, then exchange, each in
    value order), each contribution tagged with its coefficient."""

    def __init__(self, dev, j_value_idx, j_out_idx, nj_values, k_value_idx, k_out_idx, nout):
        vi = np.concatenate([np.asarray(j_value_idx, np.int64), np.asarray(k_value_idx, np.int64) + nj_values])
        oi = np.concatenate([np.asarray(j_out_idx, np.int64), np.asarray(k_out_idx, np.int64)])
        src = np.concatenate([np.zeros(len(j_value_idx), np.uint8), np.ones(len(k_value_idx), np.uint8)])
        assert vi.size < 2 ** 32 and nout < 2 ** 32
        order = np.lexsort((vi, src, oi))
        vo, oo, so = vi[order], oi[order], src[order]
        starts = np.nonzero(np.r_[True, oo[1:] != oo[:-1]])[0] if oo.size else np.zeros(0, dtype=np.int64)
        self.dev = dev
        self.nseg = int(starts.size)
        self.nout = int(nout)
        self.perm = dev.tensor(vo.astype(np.uint32))
        self.src = dev.tensor(so)
        self.seg_off = dev.tensor(np.r_[starts, oo.size].astype(np.uint32))
        self.seg_out = dev.tensor(oo[starts].astype(np.uint32))

    def run(self, values, out, cJ, cK):
        """out[...] = c_J J + c_K K (elements without contributions untouched)."""
        check(lib().vlxmi_fock_assemble(values.data_ptr(), self.perm.data_ptr(), self.src.data_ptr(),
                                        self.seg_off.data_ptr(), self.seg_out.data_ptr(), self.nseg, float(cJ),
                                        float(cK), out.data_ptr(), self.dev.stream_ptr()), "fock_assemble")
        return out
