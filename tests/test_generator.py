"""The generator's recursion DAG, evaluated in Python, reproduces the oracle's
primitive-quartet blocks for every class (the same straight-line expressions
are emitted into the HIP kernels)."""

import math

import numpy as np
import pytest

import oracle
from vlxmi.gen.osdag import QuartetDAG, tri
from vlxmi.gen import emit

CLASSES = [(la, lb, lc, ld) for la in range(3) for lb in range(3) for lc in range(3) for ld in range(3)]


def _eval_dag(cls, a, A, b, B, c, C, d, D):
    la, lb, lc, ld = cls
    L = sum(cls)
    g = QuartetDAG(*cls, rule_vrr="first", rule_hrr="first")
    outs = {}
    for ic, cc in enumerate(tri(lc)):
        for idd, dd in enumerate(tri(ld)):
            for ia, aa in enumerate(tri(la)):
                for ib, bb in enumerate(tri(lb)):
                    outs[(ia, ib, ic, idd)] = g.X(aa, bb, cc, dd)
    z, e = a + b, c + d
    S4 = z + e
    P = (a * A + b * B) / z
    Q = (c * C + d * D) / e
    PQ = Q - P
    rho = z * e / S4
    T = rho * float(np.dot(PQ, PQ))
    lam = math.sqrt(4 * rho / math.pi)
    F = oracle.boys(L, T) * lam
    env = {"double": float}
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
    code = "\n".join(ln.replace("const double ", "").rstrip(";") for ln in g.mult_decls() + g.lines)
    exec(code, env)
    n = [(l + 1) * (l + 2) // 2 for l in cls]
    out = np.zeros(n)
    for k, name in outs.items():
        out[k] = env[name]
    return out


@pytest.mark.parametrize("cls", CLASSES, ids=["".join("spd"[x] for x in c) for c in CLASSES])
def test_dag_matches_oracle(cls):
    rng = np.random.default_rng(7 + sum(v * 3 ** i for i, v in enumerate(cls)))
    pts = rng.uniform(-2, 2, (4, 3))
    ex = np.exp(rng.uniform(math.log(0.1), math.log(20), 4))
    v = _eval_dag(cls, ex[0], pts[0], ex[1], pts[1], ex[2], pts[2], ex[3], pts[3])
    r = oracle.eri_block(*cls, ex[0], pts[0], ex[1], pts[1], ex[2], pts[2], ex[3], pts[3])
    assert np.max(np.abs(v - r)) <= 1e-12 * np.max(np.abs(r))


def test_signatures_follow_reference_lists():
    # J: mat_J, per-type (info, count), mat_D, Q_local, Q, first/second/pair_data/count (local), ket lists,
    # boys tables, threshold (gen_coulomb.py:74-93)
    names = [s.split()[-1] for s in emit.j_signature("sppd")]
    assert names == ["mat_J", "s_prim_info", "s_prim_count", "p_prim_info", "p_prim_count", "d_prim_info",
                     "d_prim_count", "pd_mat_D", "sp_mat_Q_local", "pd_mat_Q", "sp_first_inds_local",
                     "sp_second_inds_local", "sp_pair_data_local", "sp_prim_pair_count_local", "pd_first_inds",
                     "pd_second_inds", "pd_pair_data", "pd_prim_pair_count", "boys_func_table", "boys_func_ft",
                     "eri_threshold"]
    # K with ab != cd carries both CSR sets, interleaved (gen_exchange.py:75-115)
    names = [s.split()[-1] for s in emit.k_signature("spdp")]
    assert names[:4] == ["mat_K", "pair_inds_i_for_K_sd", "pair_inds_k_for_K_sd", "pair_inds_count_for_K_sd"]
    assert names[-4:] == ["boys_func_table", "boys_func_ft", "omega", "eri_threshold"]
    assert names[names.index("naos") + 1:names.index("naos") + 3] == ["Q_K_sp", "Q_K_dp"]
    # ab == cd: one set only
    names = [s.split()[-1] for s in emit.k_signature("spsp")]
    assert "Q_K_sp" in names and names.count("Q_K_sp") == 1


def _digest_terms(src):
    """(accumulator, factor, mask, density/record factors) of every digestion
    term of a generated Fock kernel, without the recursion value names."""
    import re
    out = []
    for line in src.split("\n"):
        m = re.match(r"\s*((?:acc\[\d+\])|(?:pacc\d+)) \+= (.+?) \* \((.*)\);\s*$", line)
        if m and "Qk" in line:
            acc, sab = m.group(1).replace("pacc", "acc"), m.group(2)
            for t in re.findall(r"\(\(Qk(\d+) > (.+?)\) \? \(\S+ \* Sc(\d+) \* (Dj[\d_]+?)(?:_c\d+)?\)", m.group(3)):
                out.append((acc, sab) + t)
            continue
        m = re.search(r"(acc\[\d+\]) \+= \(Qd(\d+) > tau\[(\d+)\]\) \? \(\S+ \* Dp(\d+)\)", line)
        if m:
            out.append(m.groups())
        m = re.search(r"if \(Qd(\d+) > ltau\[(\d+) \* TILE_SIZE_J \+ lid\]\) __hip_atomic_fetch_add\("
                      r"&lacc\[(\d+) \* TILE_SIZE_J \+ lid\], \S+ \* Dp(\d+)", line)
        if m:
            out.append(m.groups())
    return sorted(out)


@pytest.mark.parametrize("cls,layout", [("pddd", "reg"), ("pddd", "lds"), ("dddd", "reg"), ("pdpd", "lds")])
def test_pair_passes_digest_the_same_terms(cls, layout):
    """One pass per ket component pair changes only where the recursion values
    come from: every kernel digests the same masked terms into the same
    accumulators as with one pass per ket-first-index component."""
    a, b = emit.emit_k(cls, layout=layout), emit.emit_k(cls, layout=layout, kpairs=True)
    ta, tb = _digest_terms(a), _digest_terms(b)
    assert ta and ta == tb
    nc, nd = [(emit.LV[x] + 1) * (emit.LV[x] + 2) // 2 for x in cls[2:]]
    assert b.count("for (; s < nt; ++s)") == nc * nd
    a, b = emit.emit_j(cls, layout=layout), emit.emit_j(cls, layout=layout, jpairs=True)
    ta, tb = _digest_terms(a), _digest_terms(b)
    assert ta and ta == tb
