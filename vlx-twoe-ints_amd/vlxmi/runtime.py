"""ctypes binding of libvlxmi.so over PyTorch-ROCm device tensors.

PyTorch provides device memory and streams only; all arithmetic runs in the
library's HIP kernels.  The library must exist (``vlxmi.gen.build``); there is
no fallback path: a missing library or a missing GPU raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .gen import emit
from .pairs import PAIR_CLASSES, j_classes, k_classes

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libvlxmi.so")

_lib = None


class VlxmiError(RuntimeError):
    pass


def _ctype_of(decl: str):
    decl = decl.replace("const ", "").replace("__restrict__", "").strip()
    ty = decl.rsplit(" ", 1)[0].strip()
    if ty.endswith("*"):
        return C.c_void_p
    return {"uint32_t": C.c_uint32, "double": C.c_double}[ty]


def lib():
    """Load libvlxmi.so and declare every entry point's argument types."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VlxmiError("libvlxmi.so not built: run __graft_entry__.build() (or python -m vlxmi.gen.build)")
    L = C.CDLL(LIB_PATH)
    for ab, cd in j_classes():
        cls = (ab + cd).upper()
        fn = getattr(L, "vlxmi_launch_computeCoulombFock" + cls)
        fn.argtypes = [_ctype_of(s) for s in emit.j_signature(ab + cd)] + [C.c_uint32, C.c_void_p]
        fn.restype = C.c_int
    for cls in k_classes():
        fn = getattr(L, "vlxmi_launch_computeExchangeFock" + cls.upper())
        fn.argtypes = [_ctype_of(s) for s in emit.k_signature(cls)] + [C.c_void_p]
        fn.restype = C.c_int
    mixed_args = [_ctype_of(s) for s in emit.MIXED_ARGS]       # fp32_threshold, quartet_counts
    for ab, cd in j_classes():
        fn = getattr(L, "vlxmi_launch_computeCoulombFock" + (ab + cd).upper() + "_mixed")
        fn.argtypes = [_ctype_of(s) for s in emit.j_signature(ab + cd)] + mixed_args + [C.c_uint32, C.c_void_p]
        fn.restype = C.c_int
    for cls in k_classes():
        fn = getattr(L, "vlxmi_launch_computeExchangeFock" + cls.upper() + "_mixed")
        fn.argtypes = [_ctype_of(s) for s in emit.k_signature(cls)] + mixed_args + [C.c_void_p]
        fn.restype = C.c_int
    kb_extra = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    for mixed in ("", "_mixed"):
        tail = [_ctype_of(s) for s in emit.tail_args(bool(mixed), True)]
        for ab, cd in j_classes():
            fn = getattr(L, "vlxmi_launch_computeCoulombFock" + (ab + cd).upper() + "_batched" + mixed)
            fn.argtypes = [_ctype_of(s) for s in emit.j_signature(ab + cd)] + [C.c_void_p, C.c_uint32] + tail \
                + [C.c_uint32, C.c_void_p]
            fn.restype = C.c_int
        for cls in k_classes():
            fn = getattr(L, "vlxmi_launch_computeExchangeFock" + cls.upper() + "_batched" + mixed)
            fn.argtypes = [_ctype_of(s) for s in emit.k_signature(cls)] + kb_extra + tail + [C.c_uint32, C.c_void_p]
            fn.restype = C.c_int
    from .gen.build import SPLIT_SERIES
    for kind, series in SPLIT_SERIES.items():
        for cls, n in series.items():
            for k in range(n):
                if kind == "J":
                    fn = getattr(L, "vlxmi_launch_computeCoulombFock%s%d" % (cls.upper(), k))
                    fn.argtypes = [_ctype_of(s) for s in emit.j_signature(cls)] + [C.c_uint32, C.c_void_p]
                else:
                    fn = getattr(L, "vlxmi_launch_computeExchangeFock%s%d" % (cls.upper(), k))
                    fn.argtypes = [_ctype_of(s) for s in emit.k_signature(cls)] + [C.c_void_p]
                fn.restype = C.c_int
    from .gen import emit_grad
    for ab, cd in j_classes():
        fn = getattr(L, "vlxmi_launch_computeCoulombGradient" + (ab + cd).upper())
        fn.argtypes = [_ctype_of(s) for s in emit_grad.jg_signature(ab + cd)] + [C.c_uint32, C.c_void_p]
        fn.restype = C.c_int
    for cls in k_classes():
        fn = getattr(L, "vlxmi_launch_computeExchangeGradient" + cls.upper())
        fn.argtypes = [_ctype_of(s) for s in emit_grad.kg_signature(cls)] + [C.c_void_p]
        fn.restype = C.c_int
    for pc in PAIR_CLASSES:
        fn = getattr(L, "vlxmi_schwarz_" + pc)
        fn.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        fn.restype = C.c_int
    L.vlxmi_boys_table_size.argtypes = [C.POINTER(C.c_uint32)] * 3
    L.vlxmi_boys_table_fill.argtypes = [C.c_void_p, C.c_void_p]
    L.vlxmi_boys_eval.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    L.vlxmi_boys_eval_coarse.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    L.vlxmi_status_string.restype = C.c_char_p
    L.vlxmi_status_string.argtypes = [C.c_int]
    L.vlxmi_version.restype = C.c_int
    L.vlxmi_segment_sum.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_void_p, C.c_void_p]
    L.vlxmi_fill_zero.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
    L.vlxmi_fock_assemble.argtypes = [C.c_void_p] * 5 + [C.c_uint32, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    L.vlxmi_ket_workspace.argtypes = [C.c_size_t, C.c_void_p, C.POINTER(C.c_void_p)]
    L.vlxmi_release_workspaces.argtypes = []
    L.vlxmi_ket_split_reduce.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    L.vlxmi_coulomb_fock_noalloc.argtypes = ([C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]
                                             + [C.c_void_p] * 6 + [C.c_uint32] + [C.c_void_p] * 3 + [C.c_uint32]
                                             + [C.c_void_p] * 2 + [C.c_double, C.c_uint32, C.c_void_p])
    P_ = C.c_void_p
    L.vlxmi_j_list_to_blocks.argtypes = [C.c_uint32, C.c_uint32] + [P_] * 5 + [C.c_uint32, C.c_uint32] + [P_] * 7
    L.vlxmi_ik_list_to_blocks.argtypes = [C.c_uint32, C.c_uint32, P_, P_, C.c_uint32, C.c_uint32, P_, P_, P_, P_]
    L.vlxmi_k_csr_to_blocks.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32] + [P_] * 5 + [C.c_uint64] + [P_] * 6
    L.vlxmi_scatter_add.argtypes = [P_, P_, C.c_uint64, P_, P_]
    L.vlxmi_validate_j_list.argtypes = [C.c_uint32, C.c_uint32, P_, P_, P_, P_, C.c_uint32, P_]
    L.vlxmi_validate_ik_list.argtypes = [C.c_uint32, C.c_uint32, P_, P_, C.c_uint32, P_]
    L.vlxmi_validate_k_csr.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, P_, P_, P_, P_, P_]
    L.vlxmi_set_list_validation.argtypes = [C.c_int]
    L.vlxmi_list_validation.argtypes = []
    L.vlxmi_pair_candidates.argtypes = [C.c_uint32, C.c_uint32, P_, C.c_uint32, P_, C.c_uint32] + [P_] * 5 + [P_]
    L.vlxmi_block_absmax.argtypes = [P_, C.c_uint64, C.c_uint32, P_, P_]
    L.vlxmi_kcsr_fill.argtypes = [C.c_uint32, C.c_uint32] + [P_] * 7 + [C.c_uint64] + [P_] * 6
    L.vlxmi_noalloc_scratch_doubles.restype = C.c_uint64
    L.vlxmi_noalloc_scratch_doubles.argtypes = []
    L.vlxmi_count_k_quartets.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32] + [C.c_void_p] * 6 + [C.c_double, C.c_void_p, C.c_void_p]
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().vlxmi_status_string(rc).decode()
        raise VlxmiError("%s failed: %s (%d)" % (what, msg, rc))


def exported_symbols():
    """Names every C-ABI entry point the header declares."""
    names = ["vlxmi_version", "vlxmi_boys_table_size", "vlxmi_boys_table_fill", "vlxmi_status_string",
             "vlxmi_boys_eval", "vlxmi_boys_eval_coarse", "vlxmi_count_k_quartets", "vlxmi_segment_sum",
             "vlxmi_fill_zero", "vlxmi_ket_workspace", "vlxmi_release_workspaces", "vlxmi_ket_split_reduce",
             "vlxmi_coulomb_fock_noalloc", "vlxmi_noalloc_scratch_doubles", "vlxmi_j_list_to_blocks",
             "vlxmi_ik_list_to_blocks", "vlxmi_k_csr_to_blocks", "vlxmi_scatter_add", "vlxmi_validate_j_list",
             "vlxmi_validate_ik_list", "vlxmi_validate_k_csr", "vlxmi_set_list_validation", "vlxmi_list_validation",
             "vlxmi_pair_candidates", "vlxmi_block_absmax", "vlxmi_kcsr_fill", "vlxmi_fock_assemble"]
    names += ["vlxmi_launch_computeCoulombFock" + (ab + cd).upper() for ab, cd in j_classes()]
    names += ["vlxmi_launch_computeExchangeFock" + c.upper() for c in k_classes()]
    names += ["vlxmi_launch_computeCoulombFock" + (ab + cd).upper() + "_mixed" for ab, cd in j_classes()]
    names += ["vlxmi_launch_computeExchangeFock" + c.upper() + "_mixed" for c in k_classes()]
    for mixed in ("", "_mixed"):
        names += ["vlxmi_launch_computeCoulombFock" + (ab + cd).upper() + "_batched" + mixed for ab, cd in j_classes()]
        names += ["vlxmi_launch_computeExchangeFock" + c.upper() + "_batched" + mixed for c in k_classes()]
    names += ["vlxmi_schwarz_" + pc for pc in PAIR_CLASSES]
    names += ["vlxmi_launch_computeCoulombGradient" + (ab + cd).upper() for ab, cd in j_classes()]
    names += ["vlxmi_launch_computeExchangeGradient" + c.upper() for c in k_classes()]
    from .gen.build import SPLIT_SERIES
    for kind, series in SPLIT_SERIES.items():
        base = "vlxmi_launch_computeCoulombFock" if kind == "J" else "vlxmi_launch_computeExchangeFock"
        names += ["%s%s%d" % (base, cls.upper(), k) for cls, n in series.items() for k in range(n)]
    return names


def boys_tables_host():
    nr, nc, nf = C.c_uint32(), C.c_uint32(), C.c_uint32()
    check(lib().vlxmi_boys_table_size(C.byref(nr), C.byref(nc), C.byref(nf)), "boys_table_size")
    table = np.zeros(nr.value * nc.value)
    ft = np.zeros(nf.value)
    check(lib().vlxmi_boys_table_fill(table.ctypes.data, ft.ctypes.data), "boys_table_fill")
    return table, ft


class Device:
    """Device-resident Boys tables and the stream used for launches."""

    def __init__(self, device="cuda:0"):
        import torch
        if not torch.cuda.is_available():
            raise VlxmiError("no HIP device visible: the vlxmi kernels need an MI355X (gfx950)")
        self.torch = torch
        self.device = torch.device(device)
        table, ft = boys_tables_host()
        self.boys_table = torch.from_numpy(table).to(self.device)
        self.boys_ft = torch.from_numpy(ft).to(self.device)

    def stream_ptr(self):
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def tensor(self, a, dtype=None):
        t = self.torch
        a = np.ascontiguousarray(a)
        if dtype is None:
            dtype = {np.dtype(np.float64): t.float64, np.dtype(np.uint32): t.int32,
                     np.dtype(np.int64): t.int64, np.dtype(np.int32): t.int32, np.dtype(bool): t.bool,
                     np.dtype(np.uint8): t.uint8}[a.dtype]
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        return t.from_numpy(a).to(self.device, dtype=dtype)

    def boys_eval(self, T: np.ndarray, L: int, coarse: bool = False) -> np.ndarray:
        """F_0..F_L at T by the device Boys function (coarse: the LDS path)."""
        t = self.torch
        Td = self.tensor(np.asarray(T, dtype=np.float64))
        out = t.zeros(Td.numel() * (L + 1), dtype=t.float64, device=self.device)
        fn = lib().vlxmi_boys_eval_coarse if coarse else lib().vlxmi_boys_eval
        check(fn(Td.data_ptr(), Td.numel(), L, self.boys_table.data_ptr(), out.data_ptr(), self.stream_ptr()),
              "boys_eval")
        return out.cpu().numpy().reshape(-1, L + 1)

    def schwarz(self, la, lb, info_a, info_b, I, J, S):
        """GPU Schwarz factors for candidate pair blocks (host arrays in/out)."""
        t = self.torch
        pc = "spd"[la] + "spd"[lb]
        ia, ib = self.tensor(info_a), self.tensor(info_b)
        Id, Jd = self.tensor(np.asarray(I, dtype=np.uint32)), self.tensor(np.asarray(J, dtype=np.uint32))
        Sd = self.tensor(np.ascontiguousarray(S, dtype=np.float64).ravel())
        Q = t.zeros_like(Sd)
        n = int(np.asarray(I).size)
        check(getattr(lib(), "vlxmi_schwarz_" + pc)(ia.data_ptr(), info_a.shape[1], ib.data_ptr(), info_b.shape[1],
                                                      Id.data_ptr(), Jd.data_ptr(), Sd.data_ptr(), n, Q.data_ptr(),
                                                      self.boys_table.data_ptr(), self.stream_ptr()),
              "schwarz_" + pc)
        return Q.cpu().numpy().reshape(np.shape(S))
