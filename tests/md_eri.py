"""Independent McMurchie-Davidson evaluation of primitive Cartesian ERIs
(Helgaker, Jorgensen & Olsen, Molecular Electronic-Structure Theory, 9.5-9.9),
used only to pin the oracle's Obara-Saika recursions.

Returns the same quantity as ``oracle.eri_block``: the integral over
unnormalised Cartesian Gaussians divided by (pi/zeta)^{3/2} K_AB (pi/eta)^{3/2} K_CD,
i.e. Lambda * polynomial, with the Boys function from mpmath.
"""

from __future__ import annotations

import math

import mpmath as mp

mp.mp.dps = 40


def boys_mp(n, T):
    if T == 0:
        return mp.mpf(1) / (2 * n + 1)
    T = mp.mpf(T)
    return mp.gammainc(n + mp.mpf("0.5"), 0, T) / (2 * T ** (n + mp.mpf("0.5")))


def _E(i, j, t, Qx, a, b):
    """Hermite expansion coefficients E^{ij}_t (without the K_AB factor)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return 1.0
    if j == 0:
        return (1 / (2 * p)) * _E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * _E(i - 1, j, t, Qx, a, b) \
            + (t + 1) * _E(i - 1, j, t + 1, Qx, a, b)
    return (1 / (2 * p)) * _E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * _E(i, j - 1, t, Qx, a, b) \
        + (t + 1) * _E(i, j - 1, t + 1, Qx, a, b)


def _R(t, u, v, n, p, PCx, PCy, PCz, RPC, cache):
    key = (t, u, v, n)
    if key in cache:
        return cache[key]
    T = p * RPC * RPC
    if t == u == v == 0:
        val = (-2 * p) ** n * boys_mp(n, T)
    elif t == u == 0:
        val = 0
        if v > 1:
            val += (v - 1) * _R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC, cache)
        val += PCz * _R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC, cache)
    elif t == 0:
        val = 0
        if u > 1:
            val += (u - 1) * _R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC, cache)
        val += PCy * _R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC, cache)
    else:
        val = 0
        if t > 1:
            val += (t - 1) * _R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC, cache)
        val += PCx * _R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC, cache)
    cache[key] = val
    return val


def cart(l):
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def eri_block_md(la, lb, lc, ld, a, A, b, B, c, C, d, D):
    """List-of-lists block [ia][ib][ic][id] (floats)."""
    p, q = a + b, c + d
    alpha = p * q / (p + q)
    P = [(a * A[x] + b * B[x]) / p for x in range(3)]
    Q = [(c * C[x] + d * D[x]) / q for x in range(3)]
    RPQ = math.sqrt(sum((P[x] - Q[x]) ** 2 for x in range(3)))
    cache = {}
    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    # divide out (pi/p)^1.5 (pi/q)^1.5 (K factors are never introduced here)
    norm = (math.pi / p) ** 1.5 * (math.pi / q) ** 1.5
    AB = [A[x] - B[x] for x in range(3)]
    CD = [C[x] - D[x] for x in range(3)]
    out = []
    for ca in cart(la):
        r1 = []
        for cb in cart(lb):
            r2 = []
            Eab = [[_E(ca[x], cb[x], t, AB[x], a, b) for t in range(ca[x] + cb[x] + 1)] for x in range(3)]
            for cc in cart(lc):
                r3 = []
                for cd in cart(ld):
                    Ecd = [[_E(cc[x], cd[x], t, CD[x], c, d) for t in range(cc[x] + cd[x] + 1)] for x in range(3)]
                    val = mp.mpf(0)
                    for t in range(len(Eab[0])):
                        for u in range(len(Eab[1])):
                            for v in range(len(Eab[2])):
                                e1 = Eab[0][t] * Eab[1][u] * Eab[2][v]
                                if e1 == 0:
                                    continue
                                for tau in range(len(Ecd[0])):
                                    for nu in range(len(Ecd[1])):
                                        for phi in range(len(Ecd[2])):
                                            e2 = Ecd[0][tau] * Ecd[1][nu] * Ecd[2][phi]
                                            if e2 == 0:
                                                continue
                                            val += e1 * e2 * (-1) ** (tau + nu + phi) * _R(
                                                t + tau, u + nu, v + phi, 0, alpha,
                                                P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], RPQ, cache)
                    r3.append(float(val * pref / norm))
                r2.append(r3)
            r1.append(r2)
        out.append(r1)
    return out
