"""FP64 CPU oracle for the Obara-Saika J/K Fock path.

TEST INFRASTRUCTURE ONLY.  This package is the parity reference: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only to check results (or as the timed CPU baseline), never as the
thing measured or shipped.  The product (``vlx-twoe-ints_amd/``) never imports it.

The arithmetic lives in ``oracle.c`` (built by ``oracle/Makefile`` into
``oracle/_build/liboracle.so``); this module wraps it with ctypes and adds the
pure-numpy pieces the parity tests need (AO-level J/K from the 4-index tensor).

Which reference lines each routine restates is listed in ``oracle.c``'s header.
Pins (what checks this oracle independently of the reference, which has no
numeric fixtures): Boys function vs scipy/mpmath closed forms; (ss|ss) closed
form; H2 STO-3G textbook integrals; an independent McMurchie-Davidson
evaluation for every class (s,p,d)^4; 8-fold permutation symmetry, translation
invariance; brute-force 4-index J/K by einsum.  See ``tests/test_oracle_*.py``.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_boys.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_boys.restype = None
        L.orc_eri_block.argtypes = [C.c_int] * 4 + [C.c_double, _dp, C.c_double, _dp,
                                                    C.c_double, _dp, C.c_double, _dp, C.c_double, _dp]
        L.orc_eri_block.restype = None
        L.orc_eri_deriv_block.argtypes = [C.c_int] * 5 + [C.c_double, _dp, C.c_double, _dp,
                                                          C.c_double, _dp, C.c_double, _dp, C.c_double, _dp]
        L.orc_eri_deriv_block.restype = None
        L.orc_coulomb_grad_class.restype = C.c_int
        L.orc_exchange_grad_class.restype = C.c_int
        L.orc_schwarz.argtypes = [C.c_int, C.c_int, _dp, C.c_uint32, _dp, C.c_uint32,
                                  _up, _up, C.c_uint32, _dp, _dp]
        L.orc_schwarz.restype = None
        L.orc_coulomb_class.restype = C.c_int
        L.orc_exchange_class.restype = C.c_int
        L.orc_ao_eri.restype = None
        _lib = L
    return _lib


def boys(mmax: int, T: float) -> np.ndarray:
    F = np.zeros(mmax + 1)
    lib().orc_boys(int(mmax), float(T), F)
    return F


def eri_block(la, lb, lc, ld, ai, A, aj, B, ak, Cc, al, D, omega=0.0) -> np.ndarray:
    """Lambda * polynomial for every Cartesian component of one primitive
    quartet (no contraction / normalisation factors); shape (na, nb, nc, nd)."""
    from math import comb
    n = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc, ld)]
    out = np.zeros(4096)
    f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    lib().orc_eri_block(la, lb, lc, ld, ai, f(A), aj, f(B), ak, f(Cc), al, f(D), omega, out)
    return out[: n[0] * n[1] * n[2] * n[3]].reshape(n)


def eri_deriv_block(pos, la, lb, lc, ld, ai, A, aj, B, ak, Cc, al, D, omega=0.0) -> np.ndarray:
    """d/dR_g of eri_block for R the centre of the function in position ``pos``
    (0 a, 1 b, 2 c, 3 d); shape (3, na, nb, nc, nd)."""
    n = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc, ld)]
    out = np.zeros(3 * 4096)
    f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    lib().orc_eri_deriv_block(pos, la, lb, lc, ld, ai, f(A), aj, f(B), ak, f(Cc), al, f(D), omega, out)
    return out[: 3 * n[0] * n[1] * n[2] * n[3]].reshape([3] + n)


def schwarz(la, lb, info_a, info_b, I, J, S):
    I = np.ascontiguousarray(I, dtype=np.uint32)
    J = np.ascontiguousarray(J, dtype=np.uint32)
    S = np.ascontiguousarray(S, dtype=np.float64)
    Q = np.zeros_like(S)
    lib().orc_schwarz(la, lb, np.ascontiguousarray(info_a), np.uint32(info_a.shape[1]),
                      np.ascontiguousarray(info_b), np.uint32(info_b.shape[1]),
                      I, J, np.uint32(I.size), S, Q)
    return Q


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def coulomb_class(la, lb, lc, ld, prim_info, prim_count, cd_mat_D, ab_mat_Q, cd_mat_Q,
                  ab_first, ab_second, ab_pair_data, cd_first, cd_second, cd_pair_data,
                  thr, mode="component", bra_blocks=None, mat_J=None, n_eval=None):
    """mat_J (+=) of one J class over the reference-layout lists.

    ``prim_info``/``prim_count``: sequences indexed by angular momentum.
    ``mode``: "component" (reference screening) or "block" (HIP path's)."""
    arrs = []

    def keep(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        arrs.append(a)
        return a

    infos = [keep(prim_info[l] if prim_info[l] is not None else np.zeros((5, 1)), np.float64) for l in range(3)]
    info_ptrs = (C.c_void_p * 3)(*[i.ctypes.data for i in infos])
    counts = keep(np.array([prim_count[l] for l in range(3)]), np.uint32)
    ab_first = keep(ab_first, np.uint32)
    nab = ab_first.size
    if mat_J is None:
        mat_J = np.zeros(nab)
    assert mat_J.dtype == np.float64 and mat_J.flags.c_contiguous
    if bra_blocks is not None:
        bb = keep(bra_blocks, np.uint32)
        bptr, nb = _ptr(bb), bb.size
    else:
        bptr, nb = None, 0
    rc = lib().orc_coulomb_class(
        C.c_int(la), C.c_int(lb), C.c_int(lc), C.c_int(ld), info_ptrs, _ptr(counts),
        _ptr(keep(cd_mat_D, np.float64)), _ptr(keep(ab_mat_Q, np.float64)), _ptr(keep(cd_mat_Q, np.float64)),
        _ptr(ab_first), _ptr(keep(ab_second, np.uint32)), _ptr(keep(ab_pair_data, np.float64)),
        C.c_uint32(nab),
        _ptr(keep(cd_first, np.uint32)), _ptr(keep(cd_second, np.uint32)), _ptr(keep(cd_pair_data, np.float64)),
        C.c_uint32(np.asarray(cd_first).size),
        C.c_double(thr), C.c_int(0 if mode == "component" else 1), bptr, C.c_uint32(nb), _ptr(mat_J),
        _ptr(n_eval) if n_eval is not None else None)
    if rc != 0:
        raise ValueError("oracle coulomb_class: list sizes are not whole blocks")
    return mat_J


def exchange_class(la, lb, lc, ld, ik_i, ik_k, prim_info, prim_aoinds, prim_count, max_D_bd,
                   D_AO, naos, Q_ab, Q_cd, Dinds_ab, Dinds_cd, displs_ab, displs_cd,
                   counts_ab, counts_cd, pd_ab, pd_cd, omega, thr, mode="component",
                   ik_blocks=None, mat_K=None, max_j=0, n_eval=None):
    """mat_K (+=) of one K class over the reference-layout CSR lists."""
    arrs = []

    def keep(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        arrs.append(a)
        return a

    infos = [keep(prim_info[l] if prim_info[l] is not None else np.zeros((5, 1)), np.float64) for l in range(3)]
    aos = [keep(prim_aoinds[l] if prim_aoinds[l] is not None else np.zeros((1, 1)), np.uint32) for l in range(3)]
    info_ptrs = (C.c_void_p * 3)(*[i.ctypes.data for i in infos])
    ao_ptrs = (C.c_void_p * 3)(*[i.ctypes.data for i in aos])
    counts = keep(np.array([prim_count[l] for l in range(3)]), np.uint32)
    ik_i = keep(ik_i, np.uint32)
    if mat_K is None:
        mat_K = np.zeros(ik_i.size)
    if ik_blocks is not None:
        bb = keep(ik_blocks, np.uint32)
        bptr, nb = _ptr(bb), bb.size
    else:
        bptr, nb = None, 0
    rc = lib().orc_exchange_class(
        C.c_int(la), C.c_int(lb), C.c_int(lc), C.c_int(ld), _ptr(ik_i), _ptr(keep(ik_k, np.uint32)),
        C.c_uint32(ik_i.size), info_ptrs, ao_ptrs, _ptr(counts), C.c_double(max_D_bd),
        _ptr(keep(D_AO, np.float64)), C.c_uint32(naos),
        _ptr(keep(Q_ab, np.float64)), _ptr(keep(Q_cd, np.float64)),
        _ptr(keep(Dinds_ab, np.uint32)), _ptr(keep(Dinds_cd, np.uint32)),
        _ptr(keep(displs_ab, np.uint32)), _ptr(keep(displs_cd, np.uint32)),
        _ptr(keep(counts_ab, np.uint32)), _ptr(keep(counts_cd, np.uint32)),
        _ptr(keep(pd_ab, np.float64)), _ptr(keep(pd_cd, np.float64)),
        C.c_double(omega), C.c_double(thr), C.c_int(0 if mode == "component" else 1),
        bptr, C.c_uint32(nb), _ptr(mat_K), C.c_uint32(max_j), _ptr(n_eval) if n_eval is not None else None)
    if rc != 0:
        raise ValueError("oracle exchange_class: ik list size is not whole blocks")
    return mat_K


def coulomb_grad_class(la, lb, lc, ld, prim_info, prim_count, ab_mat_D, cd_mat_D, ab_mat_Q, cd_mat_Q,
                       ab_first, ab_second, ab_pair_data, cd_first, cd_second, cd_pair_data,
                       thr, bra_blocks=None, grad=None, n_eval=None):
    """Coulomb-energy gradient parts (+=) of one J class, shape (bra blocks, 6):
    [0:3] the bra A-centre part, [3:6] the B-centre part (orc_coulomb_grad_class)."""
    arrs = []

    def keep(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        arrs.append(a)
        return a

    infos = [keep(prim_info[l] if prim_info[l] is not None else np.zeros((5, 1)), np.float64) for l in range(3)]
    info_ptrs = (C.c_void_p * 3)(*[i.ctypes.data for i in infos])
    counts = keep(np.array([prim_count[l] for l in range(3)]), np.uint32)
    ab_first = keep(ab_first, np.uint32)
    nab = (la + 1) * (la + 2) // 2 * ((lb + 1) * (lb + 2) // 2)
    if grad is None:
        grad = np.zeros((ab_first.size // nab, 6))
    assert grad.dtype == np.float64 and grad.flags.c_contiguous and grad.size == 6 * (ab_first.size // nab)
    if bra_blocks is not None:
        bb = keep(bra_blocks, np.uint32)
        bptr, nb = _ptr(bb), bb.size
    else:
        bptr, nb = None, 0
    rc = lib().orc_coulomb_grad_class(
        C.c_int(la), C.c_int(lb), C.c_int(lc), C.c_int(ld), info_ptrs, _ptr(counts),
        _ptr(keep(ab_mat_D, np.float64)), _ptr(keep(cd_mat_D, np.float64)),
        _ptr(keep(ab_mat_Q, np.float64)), _ptr(keep(cd_mat_Q, np.float64)),
        _ptr(ab_first), _ptr(keep(ab_second, np.uint32)), _ptr(keep(ab_pair_data, np.float64)),
        C.c_uint32(ab_first.size),
        _ptr(keep(cd_first, np.uint32)), _ptr(keep(cd_second, np.uint32)), _ptr(keep(cd_pair_data, np.float64)),
        C.c_uint32(np.asarray(cd_first).size),
        C.c_double(thr), bptr, C.c_uint32(nb), _ptr(grad), _ptr(n_eval) if n_eval is not None else None)
    if rc != 0:
        raise ValueError("oracle coulomb_grad_class: list sizes are not whole blocks")
    return grad


def exchange_grad_class(la, lb, lc, ld, ik_i, ik_k, prim_info, prim_aoinds, prim_count, max_D_bd,
                        D_AO, naos, Q_ab, Q_cd, Dinds_ab, Dinds_cd, displs_ab, displs_cd,
                        counts_ab, counts_cd, pd_ab, pd_cd, omega, thr, ik_blocks=None, grad=None, n_eval=None):
    """Exchange-energy gradient parts (+=) of one K class, shape ((i,k) blocks, 6):
    [0:3] the A-centre part, [3:6] the C-centre part (orc_exchange_grad_class)."""
    arrs = []

    def keep(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        arrs.append(a)
        return a

    infos = [keep(prim_info[l] if prim_info[l] is not None else np.zeros((5, 1)), np.float64) for l in range(3)]
    aos = [keep(prim_aoinds[l] if prim_aoinds[l] is not None else np.zeros((1, 1)), np.uint32) for l in range(3)]
    info_ptrs = (C.c_void_p * 3)(*[i.ctypes.data for i in infos])
    ao_ptrs = (C.c_void_p * 3)(*[i.ctypes.data for i in aos])
    counts = keep(np.array([prim_count[l] for l in range(3)]), np.uint32)
    ik_i = keep(ik_i, np.uint32)
    nac = (la + 1) * (la + 2) // 2 * ((lc + 1) * (lc + 2) // 2)
    if grad is None:
        grad = np.zeros((ik_i.size // nac, 6))
    assert grad.dtype == np.float64 and grad.flags.c_contiguous and grad.size == 6 * (ik_i.size // nac)
    if ik_blocks is not None:
        bb = keep(ik_blocks, np.uint32)
        bptr, nb = _ptr(bb), bb.size
    else:
        bptr, nb = None, 0
    rc = lib().orc_exchange_grad_class(
        C.c_int(la), C.c_int(lb), C.c_int(lc), C.c_int(ld), _ptr(ik_i), _ptr(keep(ik_k, np.uint32)),
        C.c_uint32(ik_i.size), info_ptrs, ao_ptrs, _ptr(counts), C.c_double(max_D_bd),
        _ptr(keep(D_AO, np.float64)), C.c_uint32(naos),
        _ptr(keep(Q_ab, np.float64)), _ptr(keep(Q_cd, np.float64)),
        _ptr(keep(Dinds_ab, np.uint32)), _ptr(keep(Dinds_cd, np.uint32)),
        _ptr(keep(displs_ab, np.uint32)), _ptr(keep(displs_cd, np.uint32)),
        _ptr(keep(counts_ab, np.uint32)), _ptr(keep(counts_cd, np.uint32)),
        _ptr(keep(pd_ab, np.float64)), _ptr(keep(pd_cd, np.float64)),
        C.c_double(omega), C.c_double(thr), bptr, C.c_uint32(nb), _ptr(grad),
        _ptr(n_eval) if n_eval is not None else None)
    if rc != 0:
        raise ValueError("oracle exchange_grad_class: ik list size is not whole blocks")
    return grad


def ao_eri_tensor(mol) -> np.ndarray:
    """Full Cartesian AO tensor (mu nu|la si) of a small molecule (naos <= ~60)."""
    import sys
    from vlxmi.basis import NCART, cart_components, component_norm_ratio  # noqa: F401
    shells = mol.shells
    nshell = len(shells)
    shl_l = np.array([s.l for s in shells], dtype=np.int32)
    shl_nprim = np.array([s.nprim for s in shells], dtype=np.int32)
    shl_center = np.ascontiguousarray(np.array([s.center for s in shells]).ravel())
    shl_prim_off = np.zeros(nshell, dtype=np.int32)
    off = 0
    exps, facs = [], []
    for i, s in enumerate(shells):
        shl_prim_off[i] = off
        off += s.nprim
        exps += list(s.exps)
        facs += list(s.prim_factors())
    shl_ao_off = np.array(mol.cart_offsets(), dtype=np.int32)
    ratio = np.zeros(18)
    for l in range(3):
        for c, comp in enumerate(cart_components(l)):
            ratio[l * 6 + c] = component_norm_ratio(l, comp)
    n = mol.naos_cart
    out = np.zeros(n ** 4)
    exps = np.array(exps)
    facs = np.array(facs)
    lib().orc_ao_eri(C.c_int(nshell), _ptr(shl_l), _ptr(shl_nprim), _ptr(shl_center), _ptr(shl_prim_off),
                     _ptr(exps), _ptr(facs), _ptr(shl_ao_off), _ptr(ratio), C.c_int(n), _ptr(out))
    return out.reshape(n, n, n, n)


def jk_from_tensor(eri: np.ndarray, D_cart: np.ndarray):
    """J_mn = sum (mn|ls) D_ls, K_mn = sum (ml|ns) D_ls."""
    J = np.einsum("mnls,ls->mn", eri, D_cart, optimize=True)
    K = np.einsum("mlns,ls->mn", eri, D_cart, optimize=True)
    return J, K
