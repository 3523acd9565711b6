"""Pins of the oracle's geometric-derivative routines and of the generator's
derivative code (CPU).

* orc_eri_deriv_block (Obara-Saika eq. 5, the operator of eri.py:85-175)
  against central finite differences of the full primitive integral
  (S factors' Gaussian exponentials times orc_eri_block) for every class
  (s,p,d)^4 and every centre, and the translational invariance
  sum over the four centres = 0;
* the generator's derivative expressions (the DAG the gradient kernels are
  emitted from, evaluated in Python) against orc_eri_deriv_block for every
  class, at the centres the kernels use (a and b for J, a and c for K);
* the per-atom J and K energy gradients assembled from the oracle's class
  routines against central finite differences of E_J = 1/2 sum D D (ab|cd) and
  E_K = 1/2 sum D_ac D_bd (ab|cd) from the full AO tensor (no lists, no
  screening), on two- and three-atom molecules with s, p and d shells, and their sum
  over atoms = 0.
"""

import math

import numpy as np
import pytest

import _util as U
import oracle
from vlxmi import configs
from vlxmi import grad as G
from vlxmi import pairs as P
from vlxmi.basis import Molecule, Shell
from vlxmi.gen import emit_grad as EG
from vlxmi.gen.osdag import QuartetDAG, tri

CLASSES = [(la, lb, lc, ld) for la in range(3) for lb in range(3) for lc in range(3) for ld in range(3)]
IDS = ["".join("spd"[x] for x in c) for c in CLASSES]


def _quartet(seed):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-1.2, 1.2, (4, 3))
    ex = np.exp(rng.uniform(math.log(0.2), math.log(6.0), 4))
    return ex, pts


def _full(cls, ex, pts):
    """The primitive integral up to centre-independent constants."""
    a, b, c, d = ex
    A, B, C, D = pts
    e = math.exp(-a * b / (a + b) * float(np.sum((A - B) ** 2)) - c * d / (c + d) * float(np.sum((C - D) ** 2)))
    return e * oracle.eri_block(*cls, a, A, b, B, c, C, d, D)


@pytest.mark.parametrize("cls", CLASSES, ids=IDS)
def test_deriv_block_vs_finite_difference(cls):
    ex, pts = _quartet(11 + sum(v * 3 ** i for i, v in enumerate(cls)))
    a, b, c, d = ex
    A, B, C, D = pts
    e = math.exp(-a * b / (a + b) * float(np.sum((A - B) ** 2)) - c * d / (c + d) * float(np.sum((C - D) ** 2)))
    h = 1e-4
    tot = 0.0
    for pos in range(4):
        an = oracle.eri_deriv_block(pos, *cls, a, A, b, B, c, C, d, D) * e
        tot = tot + an
        for g in range(3):
            pp, pm = pts.copy(), pts.copy()
            pp[pos, g] += h
            pm[pos, g] -= h
            fd = (_full(cls, ex, pp) - _full(cls, ex, pm)) / (2 * h)
            scale = max(np.abs(fd).max(), 1e-300)
            assert np.abs(an[g] - fd).max() <= 2e-7 * scale, (pos, g)
    # the integral does not change when all four centres move together
    assert np.abs(tot).max() <= 1e-11 * max(np.abs(an).max(), 1e-300)


def _dag_env(cls, ex, pts, L):
    a, b, c, d = ex
    A, B, C, D = pts
    z, e = a + b, c + d
    S4 = z + e
    Pp = (a * A + b * B) / z
    Q = (c * C + d * D) / e
    PQ = Q - Pp
    rho = z * e / S4
    F = oracle.boys(L, rho * float(np.dot(PQ, PQ))) * math.sqrt(4 * rho / math.pi)
    env = {"double": float, "two_ai": 2 * a, "two_aj": 2 * b, "two_ak": 2 * c}
    for m in range(L + 1):
        env["F%d" % m] = F[m]
    for i, x in enumerate("xyz"):
        env["PA_" + x] = b / z * (B[i] - A[i])
        env["QC_" + x] = d / e * (D[i] - C[i])
        env["WP_" + x] = e / S4 * PQ[i]
        env["WQ_" + x] = -z / S4 * PQ[i]
        env["AB_" + x] = A[i] - B[i]
        env["CD_" + x] = C[i] - D[i]
    env.update(h1=0.5 / z, h2=0.5 / e, h4=0.5 / S4, rz=e / S4, re=z / S4)
    return env


@pytest.mark.parametrize("cls", CLASSES, ids=IDS)
def test_generated_derivative_terms_match_oracle(cls):
    la, lb, lc, ld = cls
    ex, pts = _quartet(23 + sum(v * 3 ** i for i, v in enumerate(cls)))
    env = _dag_env(cls, ex, pts, sum(cls) + 1)
    g = QuartetDAG(*cls, rule_vrr="first", rule_hrr="first")
    exprs = {}
    two = {0: "two_ai", 1: "two_aj", 2: "two_ak"}
    for pos in (0, 1, 2):
        for ia, aa in enumerate(tri(la)):
            for ib, bb in enumerate(tri(lb)):
                for ic, cc in enumerate(tri(lc)):
                    for idd, dd in enumerate(tri(ld)):
                        comp = (aa, bb, cc, dd)
                        for ax in range(3):
                            exprs[(pos, ax, ia, ib, ic, idd)] = EG._deriv_terms(g, pos, comp[pos], comp, ax, two[pos])
    code = "\n".join(ln.replace("const double ", "").rstrip(";") for ln in g.mult_decls() + g.lines)
    exec(code, env)
    for pos in (0, 1, 2):
        ref = oracle.eri_deriv_block(pos, *cls, ex[0], pts[0], ex[1], pts[1], ex[2], pts[2], ex[3], pts[3])
        got = np.zeros_like(ref)
        for (p, ax, ia, ib, ic, idd), x in exprs.items():
            if p == pos:
                got[ax, ia, ib, ic, idd] = eval(x, env)
        assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300), pos


def _chain_molecule(n=2):
    rng = np.random.default_rng(4242 + n)
    mol = configs.lattice_molecule(rng, (n, 1, 1), name="lattice%d" % n)
    D = configs.density(rng, mol.nbf, mol.nbf // 4, 0.2)
    return mol, U.cart_density(mol, D)


def _displaced(mol, atom, axis, h):
    coords = mol.coords.copy()
    coords[atom, axis] += h
    shells = [Shell(s.atom, s.l, coords[s.atom], s.exps, s.coefs, s.spherical) for s in mol.shells]
    return Molecule(coords, shells, name=mol.name)


def _energies(mol, Dc):
    eri = oracle.ao_eri_tensor(mol)
    EJ = 0.5 * np.einsum("ab,cd,abcd->", Dc, Dc, eri, optimize=True)
    EK = 0.5 * np.einsum("ac,bd,abcd->", Dc, Dc, eri, optimize=True)
    return EJ, EK


@pytest.mark.parametrize("natoms", [2, 3])
def test_energy_gradient_vs_finite_difference(natoms):
    mol, Dc = _chain_molecule(natoms)
    assert {s.l for s in mol.shells} == {0, 1, 2}
    fl = P.build_lists(mol, U.oracle_schwarz, thr=0.0)
    gJ, gK = U.oracle_gradient(fl, Dc)
    h = 1e-4
    fJ, fK = np.zeros_like(gJ), np.zeros_like(gK)
    for at in range(mol.natoms):
        for ax in range(3):
            ep = _energies(_displaced(mol, at, ax, h), Dc)
            em = _energies(_displaced(mol, at, ax, -h), Dc)
            fJ[at, ax] = (ep[0] - em[0]) / (2 * h)
            fK[at, ax] = (ep[1] - em[1]) / (2 * h)
    print("gJ", gJ.ravel(), "fd", fJ.ravel())
    print("gK", gK.ravel(), "fd", fK.ravel())
    assert np.abs(gJ - fJ).max() <= 1e-7 * np.abs(fJ).max()
    assert np.abs(gK - fK).max() <= 1e-7 * np.abs(fK).max()
    # nothing screened: the gradient of a translation-invariant energy sums to zero
    assert np.abs(gJ.sum(axis=0)).max() <= 1e-10 * np.abs(gJ).max()
    assert np.abs(gK.sum(axis=0)).max() <= 1e-10 * np.abs(gK).max()


def test_row_map_equals_host_sum():
    mol, Dc = _chain_molecule()
    fl = P.build_lists(mol, U.oracle_schwarz)
    rows = U.oracle_gradJ(fl, Dc)
    pa = G.prim_atoms(fl)
    parts = [(rows[ab],) + G.j_block_atoms(fl, ab, pa) for ab in rows]
    ref = G.sum_rows(mol.natoms, parts)
    vals, vi, oi, off = [], [], [], 0
    for r, aa, bb in parts:
        v, o = G.row_map(aa, bb, off)
        vals.append(r.reshape(-1))
        vi.append(v)
        oi.append(o)
        off += r.size
    vals, vi, oi = np.concatenate(vals), np.concatenate(vi), np.concatenate(oi)
    got = np.zeros(3 * mol.natoms)
    np.add.at(got, oi, vals[vi])
    assert np.allclose(got.reshape(-1, 3), ref, rtol=1e-14, atol=1e-300)
