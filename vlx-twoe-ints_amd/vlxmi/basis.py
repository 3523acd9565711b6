"""Gaussian basis sets, Cartesian component order and AO bookkeeping.

Conventions fixed here and used by every other module (host, generator,
oracle, tests):

* Cartesian component order of angular momentum ``l`` is x-power major,
  y-power next: ``l=1``: x, y, z; ``l=2``: xx, xy, xz, yy, yz, zz.  The d order
  is the one the reference kernels decode through ``d_cart_inds``
  (gen_coulomb.py:143-148).
* Each Cartesian primitive component carries its own normalisation
  ``N(a; lx,ly,lz) = (2a/pi)^(3/4) (4a)^(l/2) / sqrt((2lx-1)!! (2ly-1)!! (2lz-1)!!)``
  and each contracted shell is normalised to unit self-overlap.  With this
  choice the real solid-harmonic d functions built by :func:`cart2sph_d` are
  normalised too.
* Cartesian AOs are ordered shell by shell in the order shells are listed on the
  molecule (atom-major in the generators), components inside a shell in the
  order above.  ``naos`` in the kernel signatures is the Cartesian AO count.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

ANGMOM = "spd"
NCART = (1, 3, 6)


def cart_components(l: int):
    """Cartesian exponent triples of angular momentum ``l`` in canonical order."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def _dfact(n: int) -> int:
    """Double factorial with (-1)!! = 0!! = 1."""
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def prim_norm(alpha: float, lx: int, ly: int, lz: int) -> float:
    l = lx + ly + lz
    num = (2.0 * alpha / math.pi) ** 0.75 * (4.0 * alpha) ** (0.5 * l)
    den = math.sqrt(_dfact(2 * lx - 1) * _dfact(2 * ly - 1) * _dfact(2 * lz - 1))
    return num / den


def component_norm_ratio(l: int, comp) -> float:
    """N(a; comp) / N(a; x^l): independent of the exponent."""
    lx, ly, lz = comp
    return math.sqrt(_dfact(2 * l - 1) / (_dfact(2 * lx - 1) * _dfact(2 * ly - 1) * _dfact(2 * lz - 1)))


@dataclass
class Shell:
    atom: int
    l: int
    center: np.ndarray          # (3,)
    exps: np.ndarray            # (nprim,)
    coefs: np.ndarray           # (nprim,) raw contraction coefficients
    spherical: bool = False     # only meaningful for l == 2

    def __post_init__(self):
        self.center = np.asarray(self.center, dtype=np.float64).reshape(3)
        self.exps = np.asarray(self.exps, dtype=np.float64).ravel()
        self.coefs = np.asarray(self.coefs, dtype=np.float64).ravel()
        assert self.exps.shape == self.coefs.shape and self.exps.size > 0
        assert 0 <= self.l <= 2

    @property
    def nprim(self) -> int:
        return self.exps.size

    @property
    def ncart(self) -> int:
        return ncart(self.l)

    @property
    def nsph(self) -> int:
        return 2 * self.l + 1 if (self.spherical and self.l >= 2) else ncart(self.l)

    def contraction_norm(self) -> float:
        """Factor that gives the contracted shell unit self-overlap when each
        primitive component is normalised individually."""
        a = self.exps
        c = self.coefs
        s = 0.0
        p = self.l + 1.5
        for k in range(a.size):
            for m in range(a.size):
                s += c[k] * c[m] * (2.0 * math.sqrt(a[k] * a[m]) / (a[k] + a[m])) ** p
        return 1.0 / math.sqrt(s)

    def prim_factors(self) -> np.ndarray:
        """Per-primitive factor cnorm * c_k * N(a_k; x^l) (component x^l)."""
        cn = self.contraction_norm()
        lx = self.l
        return np.array([cn * self.coefs[k] * prim_norm(self.exps[k], lx, 0, 0) for k in range(self.nprim)])


@dataclass
class Molecule:
    coords: np.ndarray                  # (natoms, 3) bohr
    shells: list = field(default_factory=list)
    name: str = ""

    @property
    def natoms(self) -> int:
        return int(self.coords.shape[0])

    @property
    def naos_cart(self) -> int:
        return sum(s.ncart for s in self.shells)

    @property
    def nbf(self) -> int:
        """Number of basis functions in the output basis (spherical d where flagged)."""
        return sum(s.nsph for s in self.shells)

    def cart_offsets(self):
        off, out = 0, []
        for s in self.shells:
            out.append(off)
            off += s.ncart
        return out

    def sph_offsets(self):
        off, out = 0, []
        for s in self.shells:
            out.append(off)
            off += s.nsph
        return out

    def cart2sph(self) -> np.ndarray:
        """T (nbf x naos_cart) with phi_sph = T @ phi_cart."""
        T = np.zeros((self.nbf, self.naos_cart))
        co, so = self.cart_offsets(), self.sph_offsets()
        for s, c0, s0 in zip(self.shells, co, so):
            if s.nsph == s.ncart:
                T[s0:s0 + s.ncart, c0:c0 + s.ncart] = np.eye(s.ncart)
            else:
                T[s0:s0 + 5, c0:c0 + 6] = cart2sph_d()
        return T

    def max_l(self) -> int:
        return max(s.l for s in self.shells)


def cart2sph_d() -> np.ndarray:
    """Rows: real solid harmonics m = -2..2 (xy, yz, z2, xz, x2-y2) in terms of
    individually normalised Cartesian d components (xx, xy, xz, yy, yz, zz)."""
    r3 = math.sqrt(3.0)
    return np.array([
        [0.0, 1.0, 0.0, 0.0, 0.0, 0.0],
        [0.0, 0.0, 0.0, 0.0, 1.0, 0.0],
        [-0.5, 0.0, 0.0, -0.5, 0.0, 1.0],
        [0.0, 0.0, 1.0, 0.0, 0.0, 0.0],
        [0.5 * r3, 0.0, 0.0, -0.5 * r3, 0.0, 0.0],
    ])


def overlap_matrix_cart(mol: Molecule) -> np.ndarray:
    """Cartesian AO overlap (closed-form Obara-Saika 1-D overlaps); used to
    check the normalisation conventions."""
    n = mol.naos_cart
    S = np.zeros((n, n))
    offs = mol.cart_offsets()
    for si, sa in enumerate(mol.shells):
        fa = sa.prim_factors()
        for sj, sb in enumerate(mol.shells):
            fb = sb.prim_factors()
            for ia, ca in enumerate(cart_components(sa.l)):
                ra = component_norm_ratio(sa.l, ca)
                for ib, cb in enumerate(cart_components(sb.l)):
                    rb = component_norm_ratio(sb.l, cb)
                    v = 0.0
                    for k in range(sa.nprim):
                        for m in range(sb.nprim):
                            v += fa[k] * ra * fb[m] * rb * _prim_overlap(
                                sa.exps[k], sa.center, ca, sb.exps[m], sb.center, cb)
                    S[offs[si] + ia, offs[sj] + ib] = v
    return S


def _prim_overlap(a, A, ca, b, B, cb):
    p = a + b
    P = (a * A + b * B) / p
    K = math.exp(-a * b / p * float(np.dot(A - B, A - B)))
    out = K * (math.pi / p) ** 1.5
    for x in range(3):
        out *= _overlap_1d(ca[x], cb[x], P[x] - A[x], P[x] - B[x], p)
    return out


def _overlap_1d(i, j, PA, PB, p):
    # S_ij / S_00 via the Obara-Saika 1-D recursion
    S = np.zeros((i + 2, j + 2))
    S[0, 0] = 1.0
    for a in range(i + 1):
        for b in range(j + 1):
            if a == 0 and b == 0:
                continue
            if a > 0:
                v = PA * S[a - 1, b]
                if a > 1:
                    v += (a - 1) / (2 * p) * S[a - 2, b]
                if b > 0:
                    v += b / (2 * p) * S[a - 1, b - 1]
            else:
                v = PB * S[a, b - 1]
                if b > 1:
                    v += (b - 1) / (2 * p) * S[a, b - 2]
            S[a, b] = v
    return S[i, j]
