"""HIP (gfx950) source emitter for the J/K energy-gradient kernels.

The reference carries the geometric derivative of an integral only as a
symbolic operator: d/dA_g (ab|cd) = 2 alpha (a+1_g b|cd) - N_g(a) (a-1_g b|cd)
and its b, c, d counterparts (Obara-Saika 1986 eq. 5; eri.py:85-175,
``write_eri_code(..., gradient=)`` at eriwriter.py:72-86).  It emits no kernel
that uses them.  These kernels contract that derivative with the density the
way the Fock kernels contract the integral, over the same lists and with the
same screening, so that a host holding the J/K lists gets the Coulomb and
exchange energy gradients without new list data:

  computeCoulombGradient{AB}{CD}: for E_J = 1/2 sum D_ab J_ab (D symmetric)
    dE_J/dX = sum_{bra entries ab} w_ab D_ab sum_{ket entries cd} w_cd D_cd
              [dA (ab|cd) if a sits on X + dB (ab|cd) if b sits on X]
    (bra-ket and a<->b symmetry of the ordered sum; w the J weight,
    gen_coulomb.py:276-281); one lane per bra primitive-pair block as in the J
    kernel, six sums per lane: the A-centre and B-centre parts of the block.
  computeExchangeGradient{A}{B}{C}{D}: for E_K = 1/2 sum D_ac K_ac
    dE_K/dX = sum_{(I,K) blocks} m_IK sum_{a,c} D_ac sum_{b,d} D_bd
              [dA (ab|cd) if a sits on X + dC (ab|cd) if c sits on X]
    (m_IK = 1 for a block I = K of one type, else 2); one wavefront per (I,K)
    block, lanes over partner blocks J of I as in the K kernel, six sums
    reduced over the wavefront: the A-centre and C-centre parts of the block.

Screening is the Fock kernels': a component quartet counts iff
|Q_ij Q_kl D_kl| > thr (J, gen_coulomb.py:209-212) or |Q_ij Q_kl maxD_bd| > thr
(K, gen_exchange.py:229-232).  The raised and lowered integrals come from the
same concrete-component DAG as the Fock kernels (one more Boys order), in one
pass or in one pass per ket-first-index component, a per-class choice
(``grad_groups``, ``grad_passes.json``).  FP64 throughout.
"""

from __future__ import annotations

from .emit import (KET_HEADER, KET_TILE, KREC0, LV, NC, _argnames, _dag_lines, _strip_restrict,
                   j_signature, k_signature, ket_loops, launder, quartet_prologue)
from .osdag import QuartetDAG, add, tri

AXES = "xyz"


def jg_signature(cls):
    """The J kernel's parameters with the output replaced by the per-block
    gradient rows and the bra entries' density added after the ket's."""
    ab, cd = cls[:2], cls[2:]
    p = j_signature(cls)
    p[0] = "double* __restrict__ grad_%s" % ab
    i = p.index("const double* __restrict__ %s_mat_D" % cd)
    p.insert(i, "const double* __restrict__ %s_mat_D_local" % ab)
    return p


def kg_signature(cls):
    """The K kernel's parameters with the output replaced by the per-(I,K)-block
    gradient rows."""
    p = k_signature(cls)
    p[0] = "double* __restrict__ grad_%s" % (cls[0] + cls[2])
    return p


def _deriv_terms(dag, pos, e, fixed, g, two_exp):
    """Expression of the derivative along g of one component quartet with
    respect to the centre of position pos (0 a, 1 b, 2 c), as
    2 alpha (.. r+1_g ..) - N_g(r) (.. r-1_g ..)."""
    comps = list(fixed)
    comps[pos] = add(e, g)
    up = dag.X(*comps)
    t = "%s * %s" % (two_exp, up)
    if e[g] > 0:
        comps[pos] = add(e, g, -1)
        dn = dag.X(*comps)
        t += " - %s" % dn if e[g] == 1 else " - %d.0 * %s" % (e[g], dn)
    return t


def grad_groups(la, lb, lc, ld, passes="one"):
    """The passes of a gradient kernel over the ket component pairs
    (t = ic * nd + id): "one" -- a single pass; "component" -- one pass per
    ket-first-index component ic (the recursion recomputed per pass, fewer
    values live at a time); "pair" -- one pass per pair t.  Neither fits the registers for the p/d classes;
    which runs faster is measured per class (``grad_passes.json``)."""
    nc, nd = (lc + 1) * (lc + 2) // 2, (ld + 1) * (ld + 2) // 2
    if passes == "component":
        return [list(range(i * nd, (i + 1) * nd)) for i in range(nc)]
    if passes == "pair":
        return [[t] for t in range(nc * nd)]
    assert passes == "one", passes
    return [list(range(nc * nd))]


def ket_tile(rec_doubles, extra_bytes=0):
    """Largest ket tile (power of two, at most 64 records) whose LDS stays
    within 32 KB, so that four single-wavefront workgroups (the most that
    512-register kernels reach) fit a CU's 160 KB."""
    kt = KET_TILE
    while kt > 8 and kt * (8 * rec_doubles + extra_bytes) > 32 * 1024:
        kt //= 2
    return kt


def emit_jgrad(cls, passes="one"):
    """Source of computeCoulombGradient{CLS} and its launcher.

    Kernel: blockDim 64, gridDim (ceil(bra blocks / 64), ket chunks); chunk y
    stores the six sums of bra block bb at grad[(y * nbra + bb) * 6 + 0..5]
    (A-centre x, y, z, then B-centre x, y, z).  The launcher runs it into the
    library's workspace and adds the chunks to the caller's rows in chunk order
    (compensated), so the result is reproducible bit for bit."""
    a, b, c, d = cls
    la, lb, lc, ld = LV[a], LV[b], LV[c], LV[d]
    na, nb, nc, nd = NC[a], NC[b], NC[c], NC[d]
    ab, cd = a + b, c + d
    nab, ncd = na * nb, nc * nd
    L = la + lb + lc + ld + 1
    rec = KREC0 + 2 * ncd
    rec += rec % 2
    name = "computeCoulombGradient" + cls.upper()
    sig = jg_signature(cls)
    groups = grad_groups(la, lb, lc, ld, passes)
    o = ["__global__ void __launch_bounds__(64)",
         "%s(%s)" % (name, ",\n    ".join(sig)),
         "{",
         "    constexpr int NA = %d, NB = %d, NC = %d, ND = %d, NAB = %d, NCD = %d;" % (na, nb, nc, nd, nab, ncd),
         "    constexpr int KT = %d, REC = %d;" % (ket_tile(rec), rec),
         "    __shared__ double ket[KT * REC];",
         "    const uint32_t lane = threadIdx.x;",
         "    const uint32_t nbra = %s_prim_pair_count_local / NAB;" % ab,
         "    const uint32_t nket = %s_prim_pair_count / NCD;" % cd,
         "    const uint32_t bb = blockIdx.x * 64 + lane;",
         "    const uint32_t chunk = (nket + gridDim.y - 1) / gridDim.y;",
         "    const uint32_t k_begin = blockIdx.y * chunk;",
         "    const uint32_t k_end = (k_begin + chunk < nket) ? (k_begin + chunk) : nket;",
         "    const bool own = bb < nbra;",
         "    const double thr = eri_threshold;",
         "    bool active = own;",
         "    double zeta = 1.0, sq_zeta = 1.0, inv_S1 = 1.0, P_x = 0.0, P_y = 0.0, P_z = 0.0;",
         "    double PA_x = 0.0, PA_y = 0.0, PA_z = 0.0, AB_x = 0.0, AB_y = 0.0, AB_z = 0.0;",
         "    double two_ai = 0.0, two_aj = 0.0, qmax = 0.0;",
         "    double tau[NAB], wab[NAB];",
         "    double gacc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};",
         "    #pragma unroll",
         "    for (int t = 0; t < NAB; ++t) { tau[t] = 0.0; wab[t] = 0.0; }",
         "    if (active) {",
         "        const uint32_t e0 = bb * NAB;",
         "        const uint32_t I = %s_first_inds_local[e0] / NA;" % ab,
         "        const uint32_t J = %s_second_inds_local[e0] / NB;" % ab,
         "        const double ai = %s_prim_info[I];" % a,
         "        const double A_x = %s_prim_info[I + %s_prim_count * 2], A_y = %s_prim_info[I + %s_prim_count * 3], A_z = %s_prim_info[I + %s_prim_count * 4];" % ((a,) * 6),
         "        const double aj = %s_prim_info[J];" % b,
         "        const double B_x = %s_prim_info[J + %s_prim_count * 2], B_y = %s_prim_info[J + %s_prim_count * 3], B_z = %s_prim_info[J + %s_prim_count * 4];" % ((b,) * 6),
         "        two_ai = 2.0 * ai; two_aj = 2.0 * aj;",
         "        zeta = ai + aj; inv_S1 = 1.0 / zeta; sq_zeta = sqrt(zeta);",
         "        P_x = (ai * A_x + aj * B_x) * inv_S1; P_y = (ai * A_y + aj * B_y) * inv_S1; P_z = (ai * A_z + aj * B_z) * inv_S1;",
         "        PA_x = (aj * inv_S1) * (B_x - A_x); PA_y = (aj * inv_S1) * (B_y - A_y); PA_z = (aj * inv_S1) * (B_z - A_z);",
         "        AB_x = A_x - B_x; AB_y = A_y - B_y; AB_z = A_z - B_z;",
         "        #pragma unroll",
         "        for (int t = 0; t < NAB; ++t) {",
         "            const double q = fabs(%s_mat_Q_local[e0 + t]);" % ab,
         "            qmax = fmax(qmax, q);",
         "            tau[t] = thr / q;"]
    if a == b:
        o.append("            const double w = (%s_first_inds_local[e0 + t] != %s_second_inds_local[e0 + t]) ? 2.0 : 1.0;" % (ab, ab))
    else:
        o.append("            const double w = 2.0;")
    o += ["            wab[t] = %s_pair_data_local[e0 + t] * %s_mat_D_local[e0 + t] * w;" % (ab, ab),
          "        }",
          "    }",
          "    for (uint32_t t0 = k_begin; t0 < k_end; t0 += KT) {",
          "        if (!__any(active)) break;",
          "        __syncthreads();",
          "        {",
          "            const uint32_t kb = t0 + lane;",
          "            double* R = ket + lane * REC;",
          "            if (lane < KT && kb < k_end) {",
          "                const uint32_t f0 = kb * NCD;",
          "                const uint32_t k = %s_first_inds[f0] / NC;" % cd,
          "                const uint32_t l = %s_second_inds[f0] / ND;" % cd,
          "                const double ak = %s_prim_info[k];" % c,
          "                const double C_x = %s_prim_info[k + %s_prim_count * 2], C_y = %s_prim_info[k + %s_prim_count * 3], C_z = %s_prim_info[k + %s_prim_count * 4];" % ((c,) * 6),
          "                const double al = %s_prim_info[l];" % d,
          "                const double D_x = %s_prim_info[l + %s_prim_count * 2], D_y = %s_prim_info[l + %s_prim_count * 3], D_z = %s_prim_info[l + %s_prim_count * 4];" % ((d,) * 6),
          "                const double et = ak + al, iet = 1.0 / et;",
          "                R[1] = et; R[2] = iet; R[12] = sqrt(et);",
          "                R[3] = (ak * C_x + al * D_x) * iet; R[4] = (ak * C_y + al * D_y) * iet; R[5] = (ak * C_z + al * D_z) * iet;",
          "                R[6] = (al * iet) * (D_x - C_x); R[7] = (al * iet) * (D_y - C_y); R[8] = (al * iet) * (D_z - C_z);",
          "                R[9] = C_x - D_x; R[10] = C_y - D_y; R[11] = C_z - D_z;",
          "                double beta = 0.0;",
          "                #pragma unroll",
          "                for (int t = 0; t < NCD; ++t) {",
          "                    const double Dkl = %s_mat_D[f0 + t];" % cd]
    if c == d:
        o.append("                    const double w = (%s_first_inds[f0 + t] != %s_second_inds[f0 + t]) ? 2.0 : 1.0;" % (cd, cd))
    else:
        o.append("                    const double w = 2.0;")
    o += ["                    R[%d + t] = %s_pair_data[f0 + t] * Dkl * w;" % (KREC0, cd),
          "                    const double qd = fabs(%s_mat_Q[f0 + t] * Dkl);" % cd,
          "                    R[%d + NCD + t] = qd;" % KREC0,
          "                    beta = fmax(beta, qd);",
          "                }",
          "                R[0] = beta;",
          "            }",
          "        }",
          "        __syncthreads();",
          "        const uint32_t nt = (k_end - t0 < (uint32_t)KT) ? (k_end - t0) : (uint32_t)KT;"]

    def body(tyq, ind):
        lines = quartet_prologue(L, ind=ind)
        for isl, group in enumerate(groups):
            if isl:
                lines += [ind + x for x in launder(L)]
            dag = QuartetDAG(la, lb, lc, ld, prefix="q%d_" % isl, rule_vrr="first", rule_hrr="first")
            dig = []
            for tcd in group:
                ic, idd = divmod(tcd, nd)
                cc, dd = tri(lc)[ic], tri(ld)[idd]
                dig.append("const double Dp%d = R[%d], Qd%d = R[%d];" % (tcd, KREC0 + tcd, tcd, KREC0 + ncd + tcd))
                for ia, aa in enumerate(tri(la)):
                    for ib, bbx in enumerate(tri(lb)):
                        tab = ia * nb + ib
                        comp = (aa, bbx, cc, dd)
                        cf = "cf%d_%d" % (tab, tcd)
                        dag.lines.append("const double %s = (Qd%d > tau[%d]) ? (wab[%d] * Dp%d) : 0.0;"
                                         % (cf, tcd, tab, tab, tcd))
                        for g in range(3):
                            dag.lines.append("gacc[%d] += %s * (%s);" % (g, cf, _deriv_terms(dag, 0, aa, comp, g, "two_ai")))
                            dag.lines.append("gacc[%d] += %s * (%s);" % (3 + g, cf, _deriv_terms(dag, 1, bbx, comp, g, "two_aj")))
            lines += [ind + x for x in dig]
            lines += [ind + x for x in _dag_lines(dag)]
        return lines

    o += ket_loops("        ", "active", "qmax * R[0] * (1.0 + 1e-13) <= thr", "qmax * R[0]", KET_HEADER,
                   body, False, "lo_mode", count=False)
    o += ["    }",
          "    if (own) {",
          "        double* row = grad_%s + ((size_t)blockIdx.y * nbra + bb) * 6;" % ab,
          "        #pragma unroll",
          "        for (int g = 0; g < 6; ++g) row[g] = gacc[g];",
          "    }",
          "}",
          ""]
    lsig = _strip_restrict(sig)
    args = _argnames(sig)
    o += ['extern "C" int vlxmi_launch_%s(%s,\n    uint32_t ket_split, void* stream)' % (name, ",\n    ".join(lsig)),
          "{",
          "    if (%s_prim_pair_count_local %% %d != 0 || %s_prim_pair_count %% %d != 0) return VLXMI_ERR_LAYOUT;"
          % (ab, nab, cd, ncd),
          "    const uint32_t nbra = %s_prim_pair_count_local / %d, nket = %s_prim_pair_count / %d;" % (ab, nab, cd, ncd),
          "    if (nbra == 0 || nket == 0) return VLXMI_OK;",
          "    if (!grad_%s || !boys_func_table) return VLXMI_ERR_ARG;" % ab,
          "    if (ket_split == 0) ket_split = 1;",
          "    if (ket_split > nket) ket_split = nket;",
          "    if (vlxmi_list_validation()) {",
          "        int rc = vlxmi_validate_j_list(%d, %d, %s_first_inds_local, %s_second_inds_local, %s_mat_Q_local, nullptr, %s_prim_pair_count_local, stream);"
          % (la, lb, ab, ab, ab, ab),
          "        if (rc == 0) rc = vlxmi_validate_j_list(%d, %d, %s_first_inds, %s_second_inds, %s_mat_Q, %s_mat_D, %s_prim_pair_count, stream);"
          % (lc, ld, cd, cd, cd, cd, cd),
          "        if (rc != 0) return rc;",
          "    }",
          "    double* part = nullptr;",
          "    int rc = vlxmi_ket_workspace((size_t)ket_split * nbra * 6, stream, &part);",
          "    if (rc != 0) return rc;",
          "    hipLaunchKernelGGL(%s, dim3((nbra + 63) / 64, ket_split), dim3(64), 0, (hipStream_t)stream, %s);"
          % (name, ", ".join(["part"] + args[1:])),
          "    rc = (int)hipGetLastError();",
          "    if (rc != 0) return rc;",
          "    return vlxmi_ket_split_reduce(part, ket_split, nbra * 6, grad_%s, stream);" % ab,
          "}"]
    return "\n".join(o) + "\n"


def emit_kgrad(cls, passes="one", split=False):
    """Source of computeExchangeGradient{CLS} and its launcher.

    Kernel: blockDim 64, gridDim = (i,k) blocks; lane 0 adds the wavefront's
    six sums to grad[blk * 6 + 0..5] (A-centre x, y, z, then C-centre x, y, z);
    one wavefront per block, so one writer per row.  With ``split`` the
    A-centre and C-centre parts are two kernels (``_A``, ``_C``), each with
    one raised recursion (a+1 on the bra, or c+1 on the ket) instead of both,
    which the launcher runs one after the other."""
    a, b, c, d = cls
    la, lb, lc, ld = LV[a], LV[b], LV[c], LV[d]
    na, nb, nc, nd = NC[a], NC[b], NC[c], NC[d]
    ac, ab, cd, bd = a + c, a + b, c + d, b + d
    nac, nab, ncd = na * nc, na * nb, nc * nd
    L = la + lb + lc + ld + 1
    rec = KREC0 + 2 * ncd
    rec += rec % 2
    name = "computeExchangeGradient" + cls.upper()
    sig = kg_signature(cls)
    groups = grad_groups(la, lb, lc, ld, passes)
    kt = ket_tile(rec, 4 * nd)
    def kernel(kname, cents):
        o = ["__global__ void __launch_bounds__(64)",
             "%s(%s)" % (kname, ",\n    ".join(sig)),
             "{",
             "    constexpr int NA = %d, NB = %d, NC = %d, ND = %d, NAC = %d, NAB = %d, NCD = %d;" % (na, nb, nc, nd, nac, nab, ncd),
             "    constexpr int KT = %d, REC = %d;" % (kt, rec),
             "    __shared__ double ket[KT * REC];",
             "    __shared__ uint32_t ket_cg[KT * ND];",
             "    const uint32_t lane = threadIdx.x;",
             "    const uint32_t blk = blockIdx.x;",
             "    if ((blk + 1) * NAC > pair_inds_count_for_K_%s) return;" % ac,
             "    const double maxD = %s_max_D;" % bd,
             "    if (!(maxD > 0.0)) return;",
             "    const double* __restrict__ Dm = mat_D_full_AO;",
             "    const uint32_t nao = naos;",
             "    const double thr = eri_threshold;",
             "    const double thr_eff = thr / maxD;",
             "    const uint32_t e0 = blk * NAC;",
             "    const uint32_t I = pair_inds_i_for_K_%s[e0] / NA;" % ac,
             "    const uint32_t K = pair_inds_k_for_K_%s[e0] / NC;" % ac,
             "    const double ai = %s_prim_info[I];" % a,
             "    const double A_x = %s_prim_info[I + %s_prim_count * 2], A_y = %s_prim_info[I + %s_prim_count * 3], A_z = %s_prim_info[I + %s_prim_count * 4];" % ((a,) * 6),
             "    const double ak = %s_prim_info[K];" % c,
             "    const double C_x = %s_prim_info[K + %s_prim_count * 2], C_y = %s_prim_info[K + %s_prim_count * 3], C_z = %s_prim_info[K + %s_prim_count * 4];" % ((c,) * 6),
             "    const double two_ai = 2.0 * ai, two_ak = 2.0 * ak;",
             "    const double mIK = %s;" % ("(I == K) ? 1.0 : 2.0" if a == c else "2.0"),
             "    double Dac[NAC];",
             "    #pragma unroll",
             "    for (int ca = 0; ca < NA; ++ca) {",
             "        const uint32_t ao_a = %s_prim_aoinds[I + %s_prim_count * ca];" % (a, a),
             "        #pragma unroll",
             "        for (int cc = 0; cc < NC; ++cc)",
             "            Dac[ca * NC + cc] = mIK * Dm[(size_t)ao_a * nao + %s_prim_aoinds[K + %s_prim_count * cc]];" % (c, c),
             "    }",
             "    uint32_t displ_i[NA], displ_k[NC];",
             "    #pragma unroll",
             "    for (int t = 0; t < NA; ++t) displ_i[t] = pair_displs_K_%s[I * NA + t];" % ab,
             "    #pragma unroll",
             "    for (int t = 0; t < NC; ++t) displ_k[t] = pair_displs_K_%s[K * NC + t];" % cd,
             "    const uint32_t nJ = pair_counts_K_%s[I * NA] / NB;" % ab,
             "    const uint32_t nL = pair_counts_K_%s[K * NC] / ND;" % cd,
             "    double gacc[%d] = {%s};" % (3 * len(cents), ", ".join(["0.0"] * (3 * len(cents)))),
             "    for (uint32_t jc = 0; jc < nJ; jc += 64) {",
             "        const uint32_t jp = jc + lane;",
             "        bool active = jp < nJ;",
             "        double zeta = 1.0, sq_zeta = 1.0, inv_S1 = 1.0, P_x = 0.0, P_y = 0.0, P_z = 0.0;",
             "        double PA_x = 0.0, PA_y = 0.0, PA_z = 0.0, AB_x = 0.0, AB_y = 0.0, AB_z = 0.0;",
             "        double qij = 0.0;",
             "        double tau[NAB], Sab[NAB];",
             "        #pragma unroll",
             "        for (int t = 0; t < NAB; ++t) { tau[t] = 0.0; Sab[t] = 0.0; }",
             "        uint32_t cgJ[NB];",
             "        #pragma unroll",
             "        for (int t = 0; t < NB; ++t) cgJ[t] = 0;",
             "        if (active) {",
             "            const uint32_t jcomb = D_inds_K_%s[displ_i[0] + jp * NB];" % ab,
             "            const uint32_t J = jcomb / NB;",
             "            const double aj = %s_prim_info[J];" % b,
             "            const double B_x = %s_prim_info[J + %s_prim_count * 2], B_y = %s_prim_info[J + %s_prim_count * 3], B_z = %s_prim_info[J + %s_prim_count * 4];" % ((b,) * 6),
             "            zeta = ai + aj; inv_S1 = 1.0 / zeta; sq_zeta = sqrt(zeta);",
             "            P_x = (ai * A_x + aj * B_x) * inv_S1; P_y = (ai * A_y + aj * B_y) * inv_S1; P_z = (ai * A_z + aj * B_z) * inv_S1;",
             "            PA_x = (aj * inv_S1) * (B_x - A_x); PA_y = (aj * inv_S1) * (B_y - A_y); PA_z = (aj * inv_S1) * (B_z - A_z);",
             "            AB_x = A_x - B_x; AB_y = A_y - B_y; AB_z = A_z - B_z;",
             "            #pragma unroll",
             "            for (int ca = 0; ca < NA; ++ca) {",
             "                #pragma unroll",
             "                for (int cb = 0; cb < NB; ++cb) {",
             "                    const uint32_t pos = displ_i[ca] + jp * NB + cb;",
             "                    const double q = fabs(Q_K_%s[pos]);" % ab,
             "                    qij = fmax(qij, q);",
             "                    tau[ca * NB + cb] = thr_eff / q;",
             "                    Sab[ca * NB + cb] = pair_data_K_%s[pos];" % ab,
             "                }",
             "            }",
             "            #pragma unroll",
             "            for (int cb = 0; cb < NB; ++cb) {",
             "                const uint32_t jj = D_inds_K_%s[displ_i[0] + jp * NB + cb];" % ab]
        if b == "s":
            o.append("                cgJ[cb] = %s_prim_aoinds[jj];" % b)
        else:
            o.append("                cgJ[cb] = %s_prim_aoinds[(jj / NB) + %s_prim_count * (jj %% NB)];" % (b, b))
        o += ["            }",
              "        }",
              "        for (uint32_t t0 = 0; t0 < nL; t0 += KT) {",
              "            if (!__any(active)) break;",
              "            __syncthreads();",
              "            {",
              "                const uint32_t lp = t0 + lane;",
              "                double* R = ket + lane * REC;",
              "                if (lane < KT && lp < nL) {",
              "                    const uint32_t lcomb = D_inds_K_%s[displ_k[0] + lp * ND];" % cd,
              "                    const uint32_t Lp = lcomb / ND;",
              "                    const double al = %s_prim_info[Lp];" % d,
              "                    const double D_x = %s_prim_info[Lp + %s_prim_count * 2], D_y = %s_prim_info[Lp + %s_prim_count * 3], D_z = %s_prim_info[Lp + %s_prim_count * 4];" % ((d,) * 6),
              "                    const double et = ak + al, iet = 1.0 / et;",
              "                    R[1] = et; R[2] = iet; R[12] = sqrt(et);",
              "                    R[3] = (ak * C_x + al * D_x) * iet; R[4] = (ak * C_y + al * D_y) * iet; R[5] = (ak * C_z + al * D_z) * iet;",
              "                    R[6] = (al * iet) * (D_x - C_x); R[7] = (al * iet) * (D_y - C_y); R[8] = (al * iet) * (D_z - C_z);",
              "                    R[9] = C_x - D_x; R[10] = C_y - D_y; R[11] = C_z - D_z;",
              "                    double qk = 0.0;",
              "                    #pragma unroll",
              "                    for (int cc = 0; cc < NC; ++cc) {",
              "                        #pragma unroll",
              "                        for (int cdd = 0; cdd < ND; ++cdd) {",
              "                            const uint32_t pos = displ_k[cc] + lp * ND + cdd;",
              "                            const double q = fabs(Q_K_%s[pos]);" % cd,
              "                            qk = fmax(qk, q);",
              "                            R[%d + cc * ND + cdd] = pair_data_K_%s[pos];" % (KREC0, cd),
              "                            R[%d + NCD + cc * ND + cdd] = q;" % KREC0,
              "                        }",
              "                    }",
              "                    R[0] = qk;",
              "                    #pragma unroll",
              "                    for (int cdd = 0; cdd < ND; ++cdd) {",
              "                        const uint32_t ll = D_inds_K_%s[displ_k[0] + lp * ND + cdd];" % cd]
        if d == "s":
            o.append("                        ket_cg[lane * ND + cdd] = %s_prim_aoinds[ll];" % d)
        else:
            o.append("                        ket_cg[lane * ND + cdd] = %s_prim_aoinds[(ll / ND) + %s_prim_count * (ll %% ND)];" % (d, d))
        o += ["                    }",
              "                }",
              "            }",
              "            __syncthreads();",
              "            const uint32_t nt = (nL - t0 < (uint32_t)KT) ? (nL - t0) : (uint32_t)KT;"]

        def body(tyq, ind):
            lines = quartet_prologue(L, omega=True, ind=ind)
            for ibb in range(nb):
                for idd in range(nd):
                    lines.append(ind + "const double Dj%d_%d = Dm[(size_t)cgJ[%d] * nao + ket_cg[s * ND + %d]];" % (ibb, idd, ibb, idd))
            for isl, group in enumerate(groups):
                if isl:
                    lines += [ind + x for x in launder(L)]
                dag = QuartetDAG(la, lb, lc, ld, prefix="q%d_" % isl, rule_vrr="first", rule_hrr="first")
                dig = []
                for tcd in group:
                    ic, idd = divmod(tcd, nd)
                    cc, dd = tri(lc)[ic], tri(ld)[idd]
                    dig.append("const double Sc%d = R[%d], Qk%d = R[%d];" % (tcd, KREC0 + tcd, tcd, KREC0 + ncd + tcd))
                    for ia, aa in enumerate(tri(la)):
                        for ib, bbx in enumerate(tri(lb)):
                            tab = ia * nb + ib
                            tac = ia * nc + ic
                            comp = (aa, bbx, cc, dd)
                            cf = "cf%d_%d" % (tab, tcd)
                            dag.lines.append("const double %s = (Qk%d > tau[%d]) ? (Dac[%d] * Sab[%d] * Sc%d * Dj%d_%d) : 0.0;"
                                             % (cf, tcd, tab, tac, tab, tcd, ib, idd))
                            for g in range(3):
                                for ic_, (pos, two, _) in enumerate(cents):
                                    dag.lines.append("gacc[%d] += %s * (%s);"
                                                     % (3 * ic_ + g, cf, _deriv_terms(dag, pos, comp[pos], comp, g, two)))
                lines += [ind + x for x in dig]
                lines += [ind + x for x in _dag_lines(dag)]
            return lines

        o += ket_loops("            ", "active", "qij * R[0] * maxD * (1.0 + 1e-13) <= thr", "qij * R[0] * maxD",
                       KET_HEADER, body, False, "lo_mode", count=False)
        o += ["        }",
              "    }",
              "    #pragma unroll",
              "    for (int g = 0; g < %d; ++g) {" % (3 * len(cents)),
              "        double v = gacc[g];",
              "        #pragma unroll",
              "        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);",
              "        if (lane == 0) grad_%s[(size_t)blk * 6 + %d + g] += v;" % (ac, cents[0][2]),
              "    }",
              "}",
              ""]
        return o

    # (position, 2 x exponent, first output column) of the centres a kernel covers
    both = [(0, "two_ai", 0), (2, "two_ak", 3)]
    parts = [(name, both)] if not split else [(name + "_A", both[:1]), (name + "_C", both[1:])]
    o = []
    for kname, cents in parts:
        o += kernel(kname, cents)
    lsig = _strip_restrict(sig)
    args = _argnames(sig)
    o += ['extern "C" int vlxmi_launch_%s(%s,\n    void* stream)' % (name, ",\n    ".join(lsig)),
          "{",
          "    if (pair_inds_count_for_K_%s %% %d != 0) return VLXMI_ERR_LAYOUT;" % (ac, nac),
          "    const uint32_t nblk = pair_inds_count_for_K_%s / %d;" % (ac, nac),
          "    if (nblk == 0) return VLXMI_OK;",
          "    if (!grad_%s || !mat_D_full_AO || !boys_func_table) return VLXMI_ERR_ARG;" % ac,
          "    if (vlxmi_list_validation()) {",
          "        int rc = vlxmi_validate_ik_list(%d, %d, pair_inds_i_for_K_%s, pair_inds_k_for_K_%s, pair_inds_count_for_K_%s, stream);"
          % (la, lc, ac, ac, ac),
          "        if (rc == 0) rc = vlxmi_validate_k_csr(%d, %d, %s_prim_count, pair_displs_K_%s, pair_counts_K_%s, Q_K_%s, D_inds_K_%s, stream);"
          % (la, lb, a, ab, ab, ab, ab)]
    if cd != ab:
        o.append("        if (rc == 0) rc = vlxmi_validate_k_csr(%d, %d, %s_prim_count, pair_displs_K_%s, pair_counts_K_%s, Q_K_%s, D_inds_K_%s, stream);"
                 % (lc, ld, c, cd, cd, cd, cd))
    o += ["        if (rc != 0) return rc;",
          "    }"]
    for kn, _ in parts:
        o += ["    hipLaunchKernelGGL(%s, dim3(nblk), dim3(64), 0, (hipStream_t)stream, %s);" % (kn, ", ".join(args)),
              "    {",
              "        const int rc = (int)hipGetLastError();",
              "        if (rc != 0) return rc;",
              "    }"]
    o += ["    return VLXMI_OK;",
          "}"]
    return "\n".join(o) + "\n"
