"""Lists in the reference's order, converted to the block order of these kernels.

VeloxChem's host builds the lists the reference kernels take in component-entry
order (gen_coulomb.py:74-93,205-212,276-281; gen_exchange.py:161-232): J pair
lists with i <= j for one type and ket lists sorted by |Q_kl D_kl| per entry,
exchange CSR segments sorted per component by Q_K, (i,k) lists in any order.
The functions below call the library's host converters (vlxmi_*_to_blocks) on
such lists and return the block-order arrays the class launchers need, plus,
for output lists (J bra lists, (i,k) lists), the reference entry of every
block entry so that ``scatter_add`` can put the block-order results back into
the caller's order.  With ``set_validation(True)`` every class launcher checks
its lists on the device first and refuses a list that is not in block order
(VLXMI_ERR_LAYOUT) instead of reading it misaligned.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .runtime import check, lib

ERR_SIZE = -3


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def j_list_to_blocks(la, lb, first, second, pair_data, mat_Q, mat_D=None):
    """A reference-order J list -> block order.  ``mat_D`` None: bra role
    (blocks by (I, J)); else ket role (blocks by non-increasing max |Q D|).
    Returns dict(first, second, pair_data, mat_Q, [mat_D], ref_index)."""
    L = lib()
    first, second = _c(first, np.uint32), _c(second, np.uint32)
    pair_data, mat_Q = _c(pair_data, np.float64), _c(mat_Q, np.float64)
    mat_D = None if mat_D is None else _c(mat_D, np.float64)
    n = int(first.size)
    need = C.c_uint32(0)
    rc = L.vlxmi_j_list_to_blocks(la, lb, _p(first), _p(second), _p(pair_data), _p(mat_Q), _p(mat_D), n, 0,
                                  None, None, None, None, None, None, C.byref(need))
    if rc not in (0, ERR_SIZE):
        check(rc, "j_list_to_blocks")
    m = need.value
    out = dict(first=np.zeros(m, np.uint32), second=np.zeros(m, np.uint32), pair_data=np.zeros(m),
               mat_Q=np.zeros(m), ref_index=np.zeros(m, np.int64))
    if mat_D is not None:
        out["mat_D"] = np.zeros(m)
    check(L.vlxmi_j_list_to_blocks(la, lb, _p(first), _p(second), _p(pair_data), _p(mat_Q), _p(mat_D), n, m,
                                   _p(out["first"]), _p(out["second"]), _p(out["pair_data"]), _p(out["mat_Q"]),
                                   _p(out.get("mat_D")), _p(out["ref_index"]), C.byref(need)), "j_list_to_blocks")
    return out


def ik_list_to_blocks(la, lc, ik_i, ik_k):
    """A reference (i,k) list -> complete (I,K) blocks; dict(i, k, ref_index)."""
    L = lib()
    ik_i, ik_k = _c(ik_i, np.uint32), _c(ik_k, np.uint32)
    n = int(ik_i.size)
    need = C.c_uint32(0)
    rc = L.vlxmi_ik_list_to_blocks(la, lc, _p(ik_i), _p(ik_k), n, 0, None, None, None, C.byref(need))
    if rc not in (0, ERR_SIZE):
        check(rc, "ik_list_to_blocks")
    m = need.value
    out = dict(i=np.zeros(m, np.uint32), k=np.zeros(m, np.uint32), ref_index=np.zeros(m, np.int64))
    check(L.vlxmi_ik_list_to_blocks(la, lc, _p(ik_i), _p(ik_k), n, m, _p(out["i"]), _p(out["k"]),
                                    _p(out["ref_index"]), C.byref(need)), "ik_list_to_blocks")
    return out


def k_csr_to_blocks(la, lb, count_a, displs, counts, Q, D_inds, pair_data):
    """A reference exchange CSR (segments sorted per component) -> block CSR;
    dict(displs, counts, Q, D_inds, pair_data)."""
    L = lib()
    displs, counts = _c(displs, np.uint32), _c(counts, np.uint32)
    Q, D_inds, pair_data = _c(Q, np.float64), _c(D_inds, np.uint32), _c(pair_data, np.float64)
    need = C.c_uint64(0)
    rc = L.vlxmi_k_csr_to_blocks(la, lb, count_a, _p(displs), _p(counts), _p(Q), _p(D_inds), _p(pair_data), 0,
                                 None, None, None, None, None, C.byref(need))
    if rc not in (0, ERR_SIZE):
        check(rc, "k_csr_to_blocks")
    m = need.value
    nseg = count_a * ((la + 1) * (la + 2) // 2)
    out = dict(displs=np.zeros(nseg, np.uint32), counts=np.zeros(nseg, np.uint32), Q=np.zeros(m),
               D_inds=np.zeros(m, np.uint32), pair_data=np.zeros(m))
    check(L.vlxmi_k_csr_to_blocks(la, lb, count_a, _p(displs), _p(counts), _p(Q), _p(D_inds), _p(pair_data), m,
                                  _p(out["displs"]), _p(out["counts"]), _p(out["Q"]), _p(out["D_inds"]),
                                  _p(out["pair_data"]), C.byref(need)), "k_csr_to_blocks")
    return out


def scatter_add(dev, src, ref_index, dst):
    """dst[ref_index[e]] += src[e] where ref_index[e] >= 0 (device tensors)."""
    check(lib().vlxmi_scatter_add(src.data_ptr(), ref_index.data_ptr(), int(src.numel()), dst.data_ptr(),
                                  dev.stream_ptr()), "scatter_add")
    return dst


def set_validation(on: bool):
    """Switch the launchers' device list checks on or off (process-wide)."""
    check(lib().vlxmi_set_list_validation(1 if on else 0), "set_list_validation")
