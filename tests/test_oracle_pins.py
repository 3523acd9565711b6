"""Pins of the FP64 oracle that do not depend on the reference (which holds no
numeric fixtures): Boys function closed forms, (ss|ss) and erf-attenuated
(ss|ss) closed forms, an independent McMurchie-Davidson evaluation of every
class (s,p,d)^4, the H2 STO-3G textbook integrals, permutation symmetry,
translation invariance and the Schwarz inequality.
"""

import math

import mpmath as mp
import numpy as np
import pytest

import md_eri
import oracle


# ----------------------------------------------------------------- Boys
@pytest.mark.parametrize("T", [0.0, 1e-14, 1e-6, 0.01, 0.3, 1.0, 2.5, 7.0, 15.0, 29.99, 30.01, 39.0, 45.0, 80.0, 400.0])
def test_boys_vs_mpmath(T):
    F = oracle.boys(12, T)
    for m in range(13):
        ref = float(md_eri.boys_mp(m, T))
        assert abs(F[m] - ref) <= 2e-15 * ref, (m, T, F[m], ref)


def test_boys_identities():
    rng = np.random.default_rng(3)
    for T in rng.uniform(0.0, 60.0, 200):
        F = oracle.boys(8, T)
        # downward recursion F_m = (2T F_{m+1} + e^-T) / (2m+1)
        for m in range(8):
            assert abs(F[m] - (2 * T * F[m + 1] + math.exp(-T)) / (2 * m + 1)) <= 1e-14 * F[m]
        # F_0 = sqrt(pi/T) erf(sqrt T) / 2
        if T > 0:
            assert abs(F[0] - 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))) <= 1e-15
    # large-T asymptote (2m-1)!!/2^(m+1) sqrt(pi/T^(2m+1))
    T = 200.0
    F = oracle.boys(8, T)
    for m in range(9):
        df = 1
        for k in range(2 * m - 1, 0, -2):
            df *= k
        asym = df / 2 ** (m + 1) * math.sqrt(math.pi / T ** (2 * m + 1))
        assert abs(F[m] / asym - 1) < 1e-14
    assert np.allclose(oracle.boys(8, 0.0), [1.0 / (2 * m + 1) for m in range(9)], rtol=0, atol=0)


# ----------------------------------------------------------------- closed forms
def _rand_quartet(rng):
    pts = rng.uniform(-2, 2, (4, 3))
    exps = np.exp(rng.uniform(math.log(0.1), math.log(20), 4))
    return exps, pts


def test_ssss_closed_form():
    rng = np.random.default_rng(11)
    for _ in range(50):
        (a, b, c, d), (A, B, C, D) = _rand_quartet(rng)
        v = oracle.eri_block(0, 0, 0, 0, a, A, b, B, c, C, d, D)[0, 0, 0, 0]
        p, q = a + b, c + d
        P, Q = (a * A + b * B) / p, (c * C + d * D) / q
        T = p * q / (p + q) * float(np.dot(P - Q, P - Q))
        full = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * float(md_eri.boys_mp(0, T))
        assert abs(v * (math.pi / p) ** 1.5 * (math.pi / q) ** 1.5 - full) <= 1e-14 * full


def test_ssss_erf_attenuated_closed_form():
    """(ss|erf(w r)/r|ss) = 2 pi^2.5/(pq sqrt(p+q)) sqrt(w^2/(w^2+rho)) F0(T w^2/(w^2+rho))."""
    rng = np.random.default_rng(12)
    for w in (0.1, 0.4, 2.0):
        (a, b, c, d), (A, B, C, D) = _rand_quartet(rng)
        v = oracle.eri_block(0, 0, 0, 0, a, A, b, B, c, C, d, D, omega=w)[0, 0, 0, 0]
        p, q = a + b, c + d
        rho = p * q / (p + q)
        P, Q = (a * A + b * B) / p, (c * C + d * D) / q
        T = rho * float(np.dot(P - Q, P - Q))
        s = w * w / (w * w + rho)
        full = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * math.sqrt(s) * float(md_eri.boys_mp(0, T * s))
        assert abs(v * (math.pi / p) ** 1.5 * (math.pi / q) ** 1.5 - full) <= 1e-14 * full
    # omega -> infinity recovers the Coulomb integral for a p class too
    (a, b, c, d), (A, B, C, D) = _rand_quartet(rng)
    v1 = oracle.eri_block(1, 1, 1, 0, a, A, b, B, c, C, d, D, omega=1e7)
    v0 = oracle.eri_block(1, 1, 1, 0, a, A, b, B, c, C, d, D)
    assert np.max(np.abs(v1 - v0)) < 1e-9 * np.max(np.abs(v0))


# ----------------------------------------------------------------- McMurchie-Davidson
CLASSES = [(la, lb, lc, ld) for la in range(3) for lb in range(3) for lc in range(3) for ld in range(3)]


@pytest.mark.parametrize("cls", CLASSES, ids=["".join("spd"[x] for x in c) for c in CLASSES])
def test_oracle_vs_mcmurchie_davidson(cls):
    rng = np.random.default_rng(100 + sum(v * 3 ** i for i, v in enumerate(cls)))
    (a, b, c, d), (A, B, C, D) = _rand_quartet(rng)
    v = oracle.eri_block(*cls, a, A, b, B, c, C, d, D)
    m = np.array(md_eri.eri_block_md(*cls, a, A, b, B, c, C, d, D))
    scale = np.max(np.abs(m))
    assert np.max(np.abs(v - m)) <= 1e-12 * scale, np.max(np.abs(v - m)) / scale


# ----------------------------------------------------------------- symmetry / invariance
def test_permutation_symmetry_dd_pd():
    rng = np.random.default_rng(5)
    (a, b, c, d), (A, B, C, D) = _rand_quartet(rng)
    v = oracle.eri_block(2, 1, 2, 0, a, A, b, B, c, C, d, D)
    # (ab|cd) = (ba|cd) = (ab|dc) = (cd|ab)
    v_ba = oracle.eri_block(1, 2, 2, 0, b, B, a, A, c, C, d, D).transpose(1, 0, 2, 3)
    v_dc = oracle.eri_block(2, 1, 0, 2, a, A, b, B, d, D, c, C).transpose(0, 1, 3, 2)
    v_ket = oracle.eri_block(2, 0, 2, 1, c, C, d, D, a, A, b, B).transpose(2, 3, 0, 1)
    s = np.max(np.abs(v))
    for w in (v_ba, v_dc, v_ket):
        assert np.max(np.abs(v - w)) <= 1e-13 * s


def test_translation_invariance_and_schwarz():
    rng = np.random.default_rng(6)
    (a, b, c, d), (A, B, C, D) = _rand_quartet(rng)
    sh = np.array([3.1, -1.7, 0.4])
    v = oracle.eri_block(2, 2, 1, 1, a, A, b, B, c, C, d, D)
    w = oracle.eri_block(2, 2, 1, 1, a, A + sh, b, B + sh, c, C + sh, d, D + sh)
    assert np.max(np.abs(v - w)) <= 1e-12 * np.max(np.abs(v))
    bra = oracle.eri_block(2, 2, 2, 2, a, A, b, B, a, A, b, B)
    ket = oracle.eri_block(1, 1, 1, 1, c, C, d, D, c, C, d, D)
    qb = np.sqrt(np.abs(np.einsum("abab->ab", bra)))
    qk = np.sqrt(np.abs(np.einsum("cdcd->cd", ket)))
    # Schwarz inequality for the full integrals (the S factors restored)
    full = oracle.eri_block(2, 2, 1, 1, a, A, b, B, c, C, d, D)
    Sab = (math.pi / (a + b)) ** 1.5 * math.exp(-a * b / (a + b) * float(np.dot(A - B, A - B)))
    Scd = (math.pi / (c + d)) ** 1.5 * math.exp(-c * d / (c + d) * float(np.dot(C - D, C - D)))
    lhs = np.abs(full) * Sab * Scd
    rhs = (qb * Sab)[:, :, None, None] * (qk * Scd)[None, None, :, :]
    assert np.all(lhs <= rhs * (1 + 1e-12))


# ----------------------------------------------------------------- H2 STO-3G
def test_h2_sto3g_textbook():
    """Szabo & Ostlund eq. 3.235: H2 STO-3G (zeta = 1.24) at R = 1.4 a.u."""
    from vlxmi.basis import Molecule, Shell
    a0 = np.array([2.227660, 0.405771, 0.109818]) * 1.24 ** 2
    dco = np.array([0.154329, 0.535328, 0.444635])
    R = np.array([[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]])
    mol = Molecule(R, [Shell(0, 0, R[0], a0, dco), Shell(1, 0, R[1], a0, dco)])
    eri = oracle.ao_eri_tensor(mol)
    assert abs(eri[0, 0, 0, 0] - 0.7746) < 5e-5
    assert abs(eri[0, 0, 1, 1] - 0.5697) < 5e-5
    assert abs(eri[1, 0, 0, 0] - 0.4441) < 5e-5
    assert abs(eri[1, 0, 1, 0] - 0.2970) < 5e-5
