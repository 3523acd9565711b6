"""HIP (gfx950) source emitter for the J, K and Schwarz kernels.

Kernel names and parameter lists are the reference's:
  computeCoulombFock{AB}{CD}    gen_coulomb.py:74-93
  computeExchangeFock{A}{B}{C}{D}  gen_exchange.py:75-115 (the cd != ab arguments
                                   are only present when the two pair types differ,
                                   gen_exchange.py:93-110)
The bodies are this project's own design (see the module docstrings of
``vlxmi.pairs`` for the list-order contract they rely on):

J: one lane owns one bra primitive-pair block (all na*nb components); the
   wave walks the ket blocks in order, staging per-block records (eta, Q, QC,
   CD, S*D*w per component, |Q*D| per component, block bound) in LDS, one record
   per lane; each lane evaluates the whole primitive quartet by straight-line
   HGP code and digests every component with the reference's per-component
   screen |Q_ij Q_kl D_kl| > thr (as D_kl-weight select against the lane's
   per-component threshold thr/|Q_ij|).  Lanes leave the scan at the first ket
   block whose bound fails (blocks are sorted).  grid.y splits the ket list;
   with more than one split the per-lane sums are added atomically.
K: one workgroup (one wave) owns one (I,K) primitive block of the ik list;
   lanes take partner blocks J of I, the wave walks the partner blocks L of K in
   order (LDS-staged records), each lane evaluates (IJ|KL) and digests with the
   reference's per-component screen |Q_ij Q_kl maxD_bd| > thr against D_AO
   gathers; the per-lane K partial sums are combined by a wavefront xor-shuffle
   reduction and lane 0 adds them to mat_K.
Schwarz: one lane per primitive-pair block computes the (ab|ab) diagonal.
"""

from __future__ import annotations

from .osdag import QuartetDAG, tri

LV = {"s": 0, "p": 1, "d": 2}
NC = {"s": 1, "p": 3, "d": 6}

KET_TILE = 64
# LDS a single-wavefront workgroup may use without lowering the occupancy of a
# kernel that needs all 512 registers (160 KB per CU / 4 SIMDs)
LDS_PER_WAVE_FULL_REGS = 40960
# ket tile of the kernels that keep the coarse Boys rows in LDS (9.7 KB)
LDS_BOYS_KT = 16


def lds_boys_setup(L):
    """Kernel-start copy of the coarse Boys rows of order L into LDS (every
    thread of the workgroup, then a barrier); quartets then evaluate F_L with
    vlxmi::boys_lds instead of gathering a table row per lane."""
    assert L + 11 <= 15, "coarse rows hold F_0..F_15"
    return ["    __shared__ double lboys[VLX_BOYS_CROWS * VLX_BOYS_CTERMS];",
            "    vlxmi::boys_lds_fill<%d>(boys_func_table, lboys);" % L,
            "    __syncthreads();"]
# ket record (LDS, doubles): 0 screening bound, 1 eta, 2 1/eta, 3-5 Q, 6-8 QC,
# 9-11 CD, 12 sqrt(eta), then the per-component values from KREC0
KREC0 = 13


def _types_in(cls):
    return [x for x in "spd" if x in cls]


# --------------------------------------------------------------------------
# common pieces
# --------------------------------------------------------------------------

def quartet_prologue(L, ty_geo="double", omega=False, ind="    ", lds_boys=False):
    """Quartet prefactors.  Expects in scope: zeta, sq_zeta = sqrt(zeta), inv_S1,
    P_x.., PA_x.., AB_x.. (bra), eta, sq_eta = sqrt(eta), inv_eta, Q_x.., QC_x..,
    CD_x.. (ket), boys tables.  One Newton-refined 1/sqrt(zeta + eta) gives both
    1/(zeta + eta) and Lambda = sqrt(4 rho / pi) = (2/sqrt(pi)) sqrt(zeta eta) / sqrt(zeta + eta)."""
    o = []
    o.append("const double S4 = zeta + eta;")
    o.append("const double rS4 = vlxmi::rsq_nr(S4);")
    o.append("const double inv_S4 = rS4 * rS4;")
    o.append("const double rho = zeta * eta * inv_S4;")
    o.append("const double PQ_x = Q_x - P_x, PQ_y = Q_y - P_y, PQ_z = Q_z - P_z;")
    o.append("double Tb = rho * (PQ_x * PQ_x + PQ_y * PQ_y + PQ_z * PQ_z);")
    o.append("double Lambda = (1.12837916709551257390 * rS4) * (sq_zeta * sq_eta);")
    if omega:
        o.append("double d2 = 1.0;")
        o.append("if (omega != 0.0) { d2 = omega * omega / (rho + omega * omega); Tb *= d2; Lambda *= sqrt(d2); }")
    o.append("double Fb[%d];" % (L + 1))
    if lds_boys:
        o.append("vlxmi::boys_lds<%d>(Tb, lboys, Fb);" % L)
    else:
        o.append("vlxmi::boys<%d>(Tb, boys_func_table, Fb);" % L)
    if omega and L > 0:
        o.append("if (omega != 0.0) { double dp = d2; "
                 + " ".join("Fb[%d] *= dp; dp *= d2;" % m for m in range(1, L + 1)) + " }")
    for m in range(L + 1):
        o.append("%s F%d = %s(Fb[%d] * Lambda);" % (ty_geo, m, ty_geo, m))
    o.append("%s WP_x = %s(eta * inv_S4 * PQ_x), WP_y = %s(eta * inv_S4 * PQ_y), WP_z = %s(eta * inv_S4 * PQ_z);"
             % (ty_geo, ty_geo, ty_geo, ty_geo))
    o.append("%s WQ_x = %s(-zeta * inv_S4 * PQ_x), WQ_y = %s(-zeta * inv_S4 * PQ_y), WQ_z = %s(-zeta * inv_S4 * PQ_z);"
             % (ty_geo, ty_geo, ty_geo, ty_geo))
    o.append("%s rz = %s(eta * inv_S4), re = %s(zeta * inv_S4);" % (ty_geo, ty_geo, ty_geo))
    o.append("%s h1 = %s(0.5 * inv_S1), h2 = %s(0.5 * inv_eta), h4 = %s(0.5 * inv_S4);"
             % (ty_geo, ty_geo, ty_geo, ty_geo))
    o.append("%s gPA_x = %s(PA_x), gPA_y = %s(PA_y), gPA_z = %s(PA_z);" % ((ty_geo,) * 4))
    o.append("%s gQC_x = %s(QC_x), gQC_y = %s(QC_y), gQC_z = %s(QC_z);" % ((ty_geo,) * 4))
    o.append("%s gAB_x = %s(AB_x), gAB_y = %s(AB_y), gAB_z = %s(AB_z);" % ((ty_geo,) * 4))
    o.append("%s gCD_x = %s(CD_x), gCD_y = %s(CD_y), gCD_z = %s(CD_z);" % ((ty_geo,) * 4))
    return [ind + x for x in o]


GEO_VARS = ["WP_x", "WP_y", "WP_z", "WQ_x", "WQ_y", "WQ_z", "rz", "re", "h1", "h2", "h4",
            "gPA_x", "gPA_y", "gPA_z", "gQC_x", "gQC_y", "gQC_z", "gAB_x", "gAB_y", "gAB_z",
            "gCD_x", "gCD_y", "gCD_z"]


def launder(L):
    """Empty volatile asm statements that make the quartet inputs opaque, so the
    compiler cannot merge (CSE / hoist) the recomputed intermediates of two
    slices and stretch their live ranges across the whole quartet."""
    names = ["F%d" % m for m in range(L + 1)] + GEO_VARS
    out = []
    for i in range(0, len(names), 10):
        grp = names[i:i + 10]
        out.append('asm volatile("" : %s);' % ", ".join('"+v"(%s)' % n for n in grp))
    return out


def _live_estimate(la, lb, lc, ld, cgroup):
    """Largest number of simultaneously live intermediates of one slice (the
    ket-first-index components in cgroup), in emission order."""
    import re as _re
    g = QuartetDAG(la, lb, lc, ld, rule_vrr="first", rule_hrr="first")
    for ic in cgroup:
        cc = tri(lc)[ic]
        for aa in tri(la):
            for bb in tri(lb):
                for dd in tri(ld):
                    g.lines.append("USE %s;" % g.X(aa, bb, cc, dd))
    last = {}
    for i, ln in enumerate(g.lines):
        rhs = ln.split("=", 1)[1] if ln.startswith("const") else ln
        for u in _re.findall(r"\bq\d+\b", rhs):
            last[u] = i
    live = mx = 0
    ends = {}
    for v, i in last.items():
        ends.setdefault(i, []).append(v)
    for i, ln in enumerate(g.lines):
        m = _re.match(r"const \w+ (q\d+) = ", ln)
        if m and m.group(1) in last:
            live += 1
            mx = max(mx, live)
        live -= sum(1 for v in ends.get(i, []))
    return mx, g.nflop


_SLICE_CACHE = {}


REG_BUDGET = 240          # doubles per lane (512 VGPR+AGPR minus addresses/indices)


def persistent_doubles(kind, la, lb, lc, ld):
    """Per-lane values live across a whole quartet besides the recursion
    intermediates: accumulators, per-component thresholds (and K's S_ab, D_bd),
    prefactors and Boys values."""
    na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc, ld)]
    base = 40 + (la + lb + lc + ld)
    if kind == "J":
        return 2 * na * nb + base
    if kind == "Kheavy":
        return na * nc + nb * nd + base
    return na * nc + 2 * na * nb + nb * nd + base


def plan_slices(la, lb, lc, ld, kind="J", budget=REG_BUDGET):
    """Partition of the ket-first-index components into passes: the fewest
    passes whose largest live set plus the per-lane persistent values fits
    ``budget`` doubles; one component per pass when none fits."""
    key = (la, lb, lc, ld, kind, budget)
    if key in _SLICE_CACHE:
        return _SLICE_CACHE[key]
    nc = (lc + 1) * (lc + 2) // 2
    nd = (ld + 1) * (ld + 2) // 2
    pers = persistent_doubles(kind, la, lb, lc, ld)
    best = [[i] for i in range(nc)]
    for nsl in range(1, nc + 1):
        if nc % nsl:
            continue
        size = nc // nsl
        groups = [list(range(k * size, (k + 1) * size)) for k in range(nsl)]
        need = max(_live_estimate(la, lb, lc, ld, g_)[0] for g_ in groups) + pers + 2 * size * nd
        if need <= budget:
            best = groups
            break
    _SLICE_CACHE[key] = best
    return best


_SLICE32_CACHE = {}


def plan_slices_f32(la, lb, lc, ld, kind="J", budget=REG_BUDGET):
    """Passes for the FP32 recursion of a mixed kernel: as plan_slices, with
    the recursion intermediates counted at half a double each (the
    accumulators, thresholds and prefactors stay FP64)."""
    key = (la, lb, lc, ld, kind, budget)
    if key in _SLICE32_CACHE:
        return _SLICE32_CACHE[key]
    nc = (lc + 1) * (lc + 2) // 2
    nd = (ld + 1) * (ld + 2) // 2
    pers = persistent_doubles(kind, la, lb, lc, ld)
    best = [[i] for i in range(nc)]
    for nsl in range(1, nc + 1):
        if nc % nsl:
            continue
        size = nc // nsl
        groups = [list(range(k * size, (k + 1) * size)) for k in range(nsl)]
        need = max(_live_estimate(la, lb, lc, ld, g_)[0] for g_ in groups) / 2.0 + pers + 2 * size * nd
        if need <= budget:
            best = groups
            break
    _SLICE32_CACHE[key] = best
    return best


def _dag_lines(dag):
    # the DAG refers to PA_/QC_/AB_/CD_ of the geometric type: rename to g*
    out = []
    for ln in dag.mult_decls() + dag.lines:
        ln = ln.replace("PA_", "gPA_").replace("QC_", "gQC_").replace("AB_", "gAB_").replace("CD_", "gCD_")
        out.append(ln)
    return out


def ket_loop(ind, act, screen, bound, header, body, ty, check_lo, lo, count=False):
    """One loop over the staged records from the current ``s`` to nt-1, the
    recursion in type ``ty``; with ``check_lo`` it stops (setting ``lo``) at
    the first record where every live lane's bound is below fp32_threshold.
    With ``count`` every evaluated quartet adds one to the lane's counter of
    its precision (nq64 / nq32)."""
    q = [ind + "for (; s < nt; ++s) {",
         ind + "    const double* R = ket + s * REC;",
         ind + "    if (%s && %s) %s = false;" % (act, screen, act),
         ind + "    if (!__any(%s)) break;" % act]
    if check_lo:
        q.append(ind + "    if (__all(!%s || %s < fp32_threshold)) { %s = true; break; }" % (act, bound, lo))
    q.append(ind + "    if (%s) {" % act)
    if count:
        q.append(ind + "        %s += 1u;" % ("nq32" if ty == "float" else "nq64"))
    q += [ind + "        " + x for x in header]
    q += body(ty, ind + "        ")
    q += [ind + "    }", ind + "}"]
    return q


def ket_loops(ind, act, screen, bound, header, body, mixed, lo, count=None):
    """Loop(s) over the staged ket records s = 0..nt-1 of one tile.

    ``act`` is the lane's live flag, cleared by the screening test ``screen``
    (the loops end when no lane is live); ``header`` reads the ket record;
    ``body(ty, ind)`` is the quartet code with the recursion in type ``ty``.
    Mixed precision: a quartet runs its recursion in FP32 when ``bound`` <
    fp32_threshold on every live lane of the wavefront.  The bound is
    non-increasing along the ket list, so the tile is one FP64 loop up to the
    first record where that holds and one FP32 loop after it; the wave-uniform
    flag ``lo`` carries the switch to later tiles.  Each loop then holds one
    recursion body only."""
    count = mixed if count is None else count

    def loop(ty, check_lo):
        return ket_loop(ind, act, screen, bound, header, body, ty, check_lo, lo, count=count)

    o = [ind + "{", ind + "uint32_t s = 0;"]
    if mixed:
        o.append(ind + "if (!%s) {" % lo)
        o += ["    " + x for x in loop("double", True)]
        o.append(ind + "}")
        o.append(ind + "if (%s) {" % lo)
        o += ["    " + x for x in loop("float", False)]
        o.append(ind + "}")
    else:
        o += loop("double", False)
    o.append(ind + "}")
    return o


def wave_decls(wpg, tile="TILE_SIZE_J"):
    """Lane and wavefront indices; with several wavefronts per workgroup (wpg)
    each wavefront works on its own items and keeps its own LDS ket tile."""
    if wpg == 1:
        return ["    const uint32_t lane = threadIdx.x;"]
    return ["    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;"]


def sync(wpg):
    """The barrier between staging the ket tile in LDS and reading it: the
    workgroup's one wavefront, or (several wavefronts) this wavefront only."""
    return "__syncthreads();" if wpg == 1 else "vlxmi::wave_sync();"


def count_decl():
    """Per-lane quartet counters of a mixed kernel (FP64 and FP32 evaluations)."""
    return ["    uint32_t nq64 = 0, nq32 = 0;"]


def count_flush(ind="    "):
    """Wavefront sums of the counters, added by lane 0 to quartet_counts[0..1]
    (FP64, FP32 primitive quartets) when the caller passed a counter array."""
    return [ind + "if (quartet_counts != nullptr) {",
            ind + "    #pragma unroll",
            ind + "    for (int off = 32; off > 0; off >>= 1) { nq64 += __shfl_xor(nq64, off, 64); nq32 += __shfl_xor(nq32, off, 64); }",
            ind + "    if ((threadIdx.x & 63) == 0) {",
            ind + "        atomicAdd(&quartet_counts[0], (unsigned long long)nq64);",
            ind + "        atomicAdd(&quartet_counts[1], (unsigned long long)nq32);",
            ind + "    }",
            ind + "}"]


MIXED_ARGS = ["const double fp32_threshold", "unsigned long long* __restrict__ quartet_counts"]


def tail_args(mixed, batched):
    """Parameters after the reference list (and the batched work-list ones):
    mixed kernels take fp32_threshold and quartet_counts, plain batched kernels
    quartet_counts (their FP64 quartets are counted too)."""
    if mixed:
        return list(MIXED_ARGS)
    return [MIXED_ARGS[1]] if batched else []


KET_HEADER = ["const double eta = R[1], inv_eta = R[2], sq_eta = R[12];",
              "const double Q_x = R[3], Q_y = R[4], Q_z = R[5];",
              "const double QC_x = R[6], QC_y = R[7], QC_z = R[8];",
              "const double CD_x = R[9], CD_y = R[10], CD_z = R[11];"]


def j_signature(cls):
    a, b, c, d = cls
    ab, cd = a + b, c + d
    p = ["double* __restrict__ mat_J"]
    for x in "spd":
        if x in cls:
            p.append("const double* __restrict__ %s_prim_info" % x)
            p.append("const uint32_t %s_prim_count" % x)
    p += ["const double* __restrict__ %s_mat_D" % cd,
          "const double* __restrict__ %s_mat_Q_local" % ab,
          "const double* __restrict__ %s_mat_Q" % cd,
          "const uint32_t* __restrict__ %s_first_inds_local" % ab,
          "const uint32_t* __restrict__ %s_second_inds_local" % ab,
          "const double* __restrict__ %s_pair_data_local" % ab,
          "const uint32_t %s_prim_pair_count_local" % ab,
          "const uint32_t* __restrict__ %s_first_inds" % cd,
          "const uint32_t* __restrict__ %s_second_inds" % cd,
          "const double* __restrict__ %s_pair_data" % cd,
          "const uint32_t %s_prim_pair_count" % cd,
          "const double* __restrict__ boys_func_table",
          "const double* __restrict__ boys_func_ft",
          "const double eri_threshold"]
    return p


def k_signature(cls):
    a, b, c, d = cls
    ac, ab, cd, bd = a + c, a + b, c + d, b + d
    p = ["double* __restrict__ mat_K",
         "const uint32_t* __restrict__ pair_inds_i_for_K_%s" % ac,
         "const uint32_t* __restrict__ pair_inds_k_for_K_%s" % ac,
         "const uint32_t pair_inds_count_for_K_%s" % ac]
    for x in "spd":
        if x in cls:
            p.append("const double* __restrict__ %s_prim_info" % x)
            p.append("const uint32_t* __restrict__ %s_prim_aoinds" % x)
            p.append("const uint32_t %s_prim_count" % x)
    p += ["const double %s_max_D" % bd,
          "const double* __restrict__ mat_D_full_AO",
          "const uint32_t naos"]
    for nm, ty in (("Q_K", "const double*"), ("D_inds_K", "const uint32_t*"),
                   ("pair_displs_K", "const uint32_t*"), ("pair_counts_K", "const uint32_t*"),
                   ("pair_data_K", "const double*")):
        p.append("%s __restrict__ %s_%s" % (ty, nm, ab))
        if cd != ab:
            p.append("%s __restrict__ %s_%s" % (ty, nm, cd))
    p += ["const double* __restrict__ boys_func_table",
          "const double* __restrict__ boys_func_ft",
          "const double omega",
          "const double eri_threshold"]
    return p


def _argnames(sig):
    return [s.split()[-1] for s in sig]


def _strip_restrict(sig):
    return [s.replace(" __restrict__", "") for s in sig]


# --------------------------------------------------------------------------
# J kernel
# --------------------------------------------------------------------------

def ket_record_bytes(kind, cls):
    """LDS bytes per staged ket record (the ket-tile size unit)."""
    c, d = cls[2], cls[3]
    ncd = NC[c] * NC[d]
    rec = KREC0 + 2 * ncd
    rec += rec % 2
    return rec * 8 + (NC[d] * 4 if kind == "K" else 0)


def emit_j(cls, mixed=False, batched=False, kt=None, layout="reg", lds_boys=False, wpg=1, jpairs=False):
    """Source of computeCoulombFock{CLS} (or its _mixed / _batched variants) and its
    C-ABI launcher.  The batched variant is persistent: workgroup w processes
    work items w, w + gridDim.x, ...; item t = (bra block begin, bra block end
    (<= begin + 64), ket block begin, ket block end) over lists that concatenate
    many molecules; each bra tile must appear in at most one item (plain +=)."""
    a, b, c, d = cls
    la, lb, lc, ld = LV[a], LV[b], LV[c], LV[d]
    na, nb, nc, nd = NC[a], NC[b], NC[c], NC[d]
    ab, cd = a + b, c + d
    nab, ncd = na * nb, nc * nd
    L = la + lb + lc + ld
    rec = KREC0 + 2 * ncd
    rec += rec % 2
    name = "computeCoulombFock" + cls.upper() + ("_batched" if batched else "") + ("_mixed" if mixed else "")
    sig = j_signature(cls) + (["const uint32_t* __restrict__ work", "const uint32_t nwork"] if batched else []) \
        + tail_args(mixed, batched)
    ty = "double"
    o = []
    # Non-batched kernels share one body: with ket_partial == nullptr the
    # chunks of a ket split (grid.y) add their sums to mat_J atomically (the
    # kernel with the reference's parameter list); otherwise chunk y stores its
    # partial sums at ket_partial[y * count_local + e] for the launcher's
    # deterministic reduction.
    KP = ["double* __restrict__ ket_partial"]
    LB = "TILE_SIZE_J" if wpg == 1 else str(64 * wpg)
    if batched:
        o.append("__global__ void __launch_bounds__(%s)" % LB)
        o.append("%s(%s)" % (name, ",\n    ".join(sig)))
    else:
        o.append("static __device__ __forceinline__ void")
        o.append("%s_body(%s)" % (name, ",\n    ".join(sig + KP)))
    o.append("{")
    o.append("    constexpr int NA = %d, NB = %d, NC = %d, ND = %d, NAB = %d, NCD = %d;" % (na, nb, nc, nd, nab, ncd))
    o.append("    constexpr int KT = %d, REC = %d;" % (LDS_BOYS_KT if lds_boys else (kt or KET_TILE), rec))
    if wpg == 1:
        o.append("    __shared__ double ket[KT * REC];")
    else:
        o.append("    constexpr int WPG = %d;" % wpg)
        o.append("    __shared__ double ket_all[WPG * KT * REC];")
    o += wave_decls(wpg)
    if wpg > 1:
        o.append("    double* ket = ket_all + wave * (KT * REC);")
    if lds_boys:
        o += lds_boys_setup(L)
    o.append("    const uint32_t nbra = %s_prim_pair_count_local / NAB;" % ab)
    o.append("    const uint32_t nket = %s_prim_pair_count / NCD;" % cd)
    if batched:
        o += count_decl()
        if wpg == 1:
            o.append("    for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {")
        else:
            o.append("    for (uint32_t w = blockIdx.x * WPG + wave; w < nwork; w += gridDim.x * WPG) {")
        o.append("    const uint32_t bra_end = work[4 * w + 1] < nbra ? work[4 * w + 1] : nbra;")
        o.append("    const uint32_t bb = work[4 * w] + lane;")
        o.append("    const uint32_t k_begin = work[4 * w + 2];")
        o.append("    const uint32_t k_end = work[4 * w + 3] < nket ? work[4 * w + 3] : nket;")
        o.append("    const bool own = bb < bra_end;")
    else:
        o.append("    const uint32_t bb = %s * TILE_SIZE_J + lane;" % ("blockIdx.x" if wpg == 1 else "(blockIdx.x * WPG + wave)"))
        o.append("    const uint32_t chunk = (nket + gridDim.y - 1) / gridDim.y;")
        o.append("    const uint32_t k_begin = blockIdx.y * chunk;")
        o.append("    const uint32_t k_end = (k_begin + chunk < nket) ? (k_begin + chunk) : nket;")
        o.append("    const bool own = bb < nbra;")
    o.append("    const double thr = eri_threshold;")
    o.append("    bool active = own;")
    o.append("    double zeta = 1.0, sq_zeta = 1.0, inv_S1 = 1.0, P_x = 0.0, P_y = 0.0, P_z = 0.0;")
    o.append("    double PA_x = 0.0, PA_y = 0.0, PA_z = 0.0, AB_x = 0.0, AB_y = 0.0, AB_z = 0.0;")
    o.append("    double qmax = 0.0;")
    jsl = plan_slices(la, lb, lc, ld, "J")
    if len(jsl) > 1 and jpairs:
        # one pass per ket component pair (ic, id), recursion recomputed per pair
        assert not mixed, "pair passes are for the FP64 kernels"
        jsl = [[(ic, idd)] for ic in range(nc) for idd in range(nd)]
    # Classes split into passes have two layouts, chosen per class by
    # measurement: "reg" keeps the NAB accumulators and thresholds in registers;
    # "lds" keeps them in LDS ([component][lane], lane-private slots), the
    # digestion adding each term with one LDS f64 add and reading the
    # threshold per use, which frees 4*NAB registers for the recursion.
    jlds = len(jsl) > 1 and layout == "lds"
    assert not (jlds and wpg > 1), "the LDS accumulator layout keeps one wavefront per workgroup"
    if jlds:
        o.append("    __shared__ double ltau[NAB * TILE_SIZE_J], lacc[NAB * TILE_SIZE_J];")
        o.append("    double* tau = ltau;")
        o.append("    #pragma unroll")
        o.append("    for (int t = 0; t < NAB; ++t) { lacc[t * TILE_SIZE_J + lane] = 0.0; tau[t * TILE_SIZE_J + lane] = 0.0; }")
    else:
        o.append("    double tau[NAB];")
        o.append("    double acc[NAB];")
        o.append("    #pragma unroll")
        o.append("    for (int t = 0; t < NAB; ++t) { acc[t] = 0.0; tau[t] = 0.0; }")
    o.append("    if (active) {")
    o.append("        const uint32_t e0 = bb * NAB;")
    o.append("        const uint32_t I = %s_first_inds_local[e0] / NA;" % ab)
    o.append("        const uint32_t J = %s_second_inds_local[e0] / NB;" % ab)
    o.append("        const double ai = %s_prim_info[I];" % a)
    o.append("        const double A_x = %s_prim_info[I + %s_prim_count * 2], A_y = %s_prim_info[I + %s_prim_count * 3], A_z = %s_prim_info[I + %s_prim_count * 4];" % ((a,) * 6))
    o.append("        const double aj = %s_prim_info[J];" % b)
    o.append("        const double B_x = %s_prim_info[J + %s_prim_count * 2], B_y = %s_prim_info[J + %s_prim_count * 3], B_z = %s_prim_info[J + %s_prim_count * 4];" % ((b,) * 6))
    o.append("        zeta = ai + aj; inv_S1 = 1.0 / zeta; sq_zeta = sqrt(zeta);")
    o.append("        P_x = (ai * A_x + aj * B_x) * inv_S1; P_y = (ai * A_y + aj * B_y) * inv_S1; P_z = (ai * A_z + aj * B_z) * inv_S1;")
    o.append("        PA_x = (aj * inv_S1) * (B_x - A_x); PA_y = (aj * inv_S1) * (B_y - A_y); PA_z = (aj * inv_S1) * (B_z - A_z);")
    o.append("        AB_x = A_x - B_x; AB_y = A_y - B_y; AB_z = A_z - B_z;")
    o.append("        #pragma unroll")
    o.append("        for (int t = 0; t < NAB; ++t) {")
    o.append("            const double q = fabs(%s_mat_Q_local[e0 + t]);" % ab)
    o.append("            qmax = fmax(qmax, q);")
    o.append("            tau[%s] = thr / q;" % ("t * TILE_SIZE_J + lane" if jlds else "t"))
    o.append("        }")
    o.append("    }")
    if mixed:
        o.append("    bool lo_mode = false;")
        if not batched:
            o += count_decl()
    o.append("    for (uint32_t t0 = k_begin; t0 < k_end; t0 += KT) {")
    o.append("        if (!__any(active)) break;")
    o.append("        " + sync(wpg))
    o.append("        {")
    o.append("            const uint32_t kb = t0 + lane;")
    o.append("            double* R = ket + lane * REC;")
    o.append("            if (lane < KT && kb < k_end) {")
    o.append("                const uint32_t f0 = kb * NCD;")
    o.append("                const uint32_t k = %s_first_inds[f0] / NC;" % cd)
    o.append("                const uint32_t l = %s_second_inds[f0] / ND;" % cd)
    o.append("                const double ak = %s_prim_info[k];" % c)
    o.append("                const double C_x = %s_prim_info[k + %s_prim_count * 2], C_y = %s_prim_info[k + %s_prim_count * 3], C_z = %s_prim_info[k + %s_prim_count * 4];" % ((c,) * 6))
    o.append("                const double al = %s_prim_info[l];" % d)
    o.append("                const double D_x = %s_prim_info[l + %s_prim_count * 2], D_y = %s_prim_info[l + %s_prim_count * 3], D_z = %s_prim_info[l + %s_prim_count * 4];" % ((d,) * 6))
    o.append("                const double et = ak + al, iet = 1.0 / et;")
    o.append("                R[1] = et; R[2] = iet; R[12] = sqrt(et);")
    o.append("                R[3] = (ak * C_x + al * D_x) * iet; R[4] = (ak * C_y + al * D_y) * iet; R[5] = (ak * C_z + al * D_z) * iet;")
    o.append("                R[6] = (al * iet) * (D_x - C_x); R[7] = (al * iet) * (D_y - C_y); R[8] = (al * iet) * (D_z - C_z);")
    o.append("                R[9] = C_x - D_x; R[10] = C_y - D_y; R[11] = C_z - D_z;")
    o.append("                double beta = 0.0;")
    o.append("                #pragma unroll")
    o.append("                for (int t = 0; t < NCD; ++t) {")
    o.append("                    const double Dkl = %s_mat_D[f0 + t];" % cd)
    if c == d:
        o.append("                    const double w = (%s_first_inds[f0 + t] != %s_second_inds[f0 + t]) ? 2.0 : 1.0;" % (cd, cd))
    else:
        o.append("                    const double w = 2.0;")
    o.append("                    R[%d + t] = %s_pair_data[f0 + t] * Dkl * w;" % (KREC0, cd))
    o.append("                    const double qd = fabs(%s_mat_Q[f0 + t] * Dkl);" % cd)
    o.append("                    R[%d + NCD + t] = qd;" % KREC0)
    o.append("                    beta = fmax(beta, qd);")
    o.append("                }")
    o.append("                R[0] = beta;")
    o.append("            }")
    o.append("        }")
    o.append("        " + sync(wpg))
    o.append("        const uint32_t nt = (k_end - t0 < (uint32_t)KT) ? (k_end - t0) : (uint32_t)KT;")

    def jbody(tyq, ind, groups):
        lines = quartet_prologue(L, ty_geo=tyq, ind=ind, lds_boys=lds_boys)
        if jlds:
            lines += [ind + "uint32_t lid = lane;", ind + "asm volatile(\"\" : \"+v\"(lid));"]
        for isl, group in enumerate(groups):
            if isl:
                lines += [ind + x for x in launder(L)]
            dag = QuartetDAG(la, lb, lc, ld, ty=tyq, rule_vrr="first", rule_hrr="first", prefix="q%d_" % isl)
            dig = []
            for e in group:
                ic, idds = (e[0], [e[1]]) if isinstance(e, tuple) else (e, list(range(nd)))
                cc = tri(lc)[ic]
                for idd in idds:
                    dd = tri(ld)[idd]
                    tcd = ic * nd + idd
                    dig.append("const double Dp%d = R[%d], Qd%d = R[%d];" % (tcd, KREC0 + tcd, tcd, KREC0 + ncd + tcd))
                    for ia, aa in enumerate(tri(la)):
                        for ib, bbb in enumerate(tri(lb)):
                            tab = ia * nb + ib
                            x = dag.X(aa, bbb, cc, dd)
                            xv = x if tyq == "double" else "double(%s)" % x
                            if jlds:
                                dag.lines.append("if (Qd%d > ltau[%d * TILE_SIZE_J + lid]) __hip_atomic_fetch_add(&lacc[%d * TILE_SIZE_J + lid], %s * Dp%d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);"
                                                 % (tcd, tab, tab, xv, tcd))
                            else:
                                dag.lines.append("acc[%d] += (Qd%d > tau[%d]) ? (%s * Dp%d) : 0.0;" % (tab, tcd, tab, xv, tcd))
            lines += [ind + x for x in dig]
            lines += [ind + x for x in _dag_lines(dag)]
        return lines

    # J keeps its passes inside one ket loop (the bra-side state is small);
    # measured faster than one loop per pass
    # the FP32 loop of a mixed kernel may run fewer passes (half-size values)
    jsl32 = plan_slices_f32(la, lb, lc, ld, "J") if mixed else jsl
    o += ket_loops("        ", "active", "qmax * R[0] * (1.0 + 1e-13) <= thr", "qmax * R[0]", KET_HEADER,
                   lambda ty, ind: jbody(ty, ind, jsl if ty == "double" else jsl32), mixed, "lo_mode",
                   count=mixed or batched)
    o.append("    }")
    o.append("    if (own) {")
    o.append("        const uint32_t e0 = bb * NAB;")
    o.append("        #pragma unroll")
    o.append("        for (int t = 0; t < NAB; ++t) {")
    o.append("            const double v = %s_pair_data_local[e0 + t] * %s;" % (ab, "lacc[t * TILE_SIZE_J + lane]" if jlds else "acc[t]"))
    if batched:
        o.append("            mat_J[e0 + t] += v;")
    else:
        o.append("            if (ket_partial != nullptr) ket_partial[(size_t)blockIdx.y * %s_prim_pair_count_local + e0 + t] = v;" % ab)
        o.append("            else if (gridDim.y > 1) atomicAdd(&mat_J[e0 + t], v); else mat_J[e0 + t] += v;")
    o.append("        }")
    o.append("    }")
    if batched:
        o.append("    }")
    if mixed or batched:
        o += count_flush()
    o.append("}")
    o.append("")
    lsig = _strip_restrict(sig)
    args = _argnames(sig)
    if not batched:
        if not mixed:
            o.append("__global__ void __launch_bounds__(%s)" % LB)
            o.append("%s(%s)" % (name, ",\n    ".join(sig)))
            o.append("{")
            o.append("    %s_body(%s, nullptr);" % (name, ", ".join(args)))
            o.append("}")
            o.append("")
        o.append("__global__ void __launch_bounds__(%s)" % LB)
        o.append("%s_ks(%s)" % (name, ",\n    ".join(sig + KP)))
        o.append("{")
        o.append("    %s_body(%s, ket_partial);" % (name, ", ".join(args)))
        o.append("}")
        o.append("")
    # launcher
    checks = ["    if (%s_prim_pair_count_local %% %d != 0 || %s_prim_pair_count %% %d != 0) return VLXMI_ERR_LAYOUT;"
              % (ab, nab, cd, ncd),
              "    const uint32_t nbra = %s_prim_pair_count_local / %d, nket = %s_prim_pair_count / %d;" % (ab, nab, cd, ncd),
              "    if (nbra == 0 || nket == 0) return VLXMI_OK;",
              "    if (!mat_J || !boys_func_table) return VLXMI_ERR_ARG;"]
    if batched:
        o.append('extern "C" int vlxmi_launch_%s(%s,\n    uint32_t n_workgroups, void* stream)' % (name, ",\n    ".join(lsig)))
        o.append("{")
        o += checks
        o.append("    if (nwork == 0) return VLXMI_OK;")
        o.append("    if (!work) return VLXMI_ERR_ARG;")
        o.append("    const uint32_t max_wg = (nwork + %d) / %d;" % (wpg - 1, wpg))
        o.append("    if (n_workgroups == 0 || n_workgroups > max_wg) n_workgroups = max_wg;")
        o.append("    dim3 grid(n_workgroups);")
        o.append("    hipLaunchKernelGGL(%s, grid, dim3(%d), 0, (hipStream_t)stream, %s);" % (name, 64 * wpg, ", ".join(args)))
        o.append("    return (int)hipGetLastError();")
        o.append("}")
        return "\n".join(o) + "\n"
    clamp = ["    if (ket_split == 0) ket_split = 1;",
             "    if (ket_split > nket) ket_split = nket;"]
    # The launch with a caller-provided partial-sum buffer ``part`` of at least
    # ket_split * count_local doubles (used when ket_split > 1): the chunks of
    # the ket list run in parallel (grid.y), each storing its partial sums, and
    # a compensated reduction adds them to mat_J in chunk order.
    o.append('extern "C" int vlxmi_jpart_%s(%s,\n    uint32_t ket_split, double* part, void* stream)'
             % (name, ",\n    ".join(lsig)))
    o.append("{")
    o += checks + clamp
    o.append("    dim3 grid((nbra + %d) / %d, ket_split);" % (64 * wpg - 1, 64 * wpg))
    o.append("    if (ket_split > 1) {")
    o.append("        if (!part) return VLXMI_ERR_ARG;")
    o.append("        hipLaunchKernelGGL(%s_ks, grid, dim3(%d), 0, (hipStream_t)stream, %s, part);" % (name, 64 * wpg, ", ".join(args)))
    o.append("        const int rc = (int)hipGetLastError();")
    o.append("        if (rc != 0) return rc;")
    o.append("        return vlxmi_ket_split_reduce(part, ket_split, %s_prim_pair_count_local, mat_J, stream);" % ab)
    o.append("    }")
    if mixed:
        o.append("    hipLaunchKernelGGL(%s_ks, grid, dim3(%d), 0, (hipStream_t)stream, %s, nullptr);" % (name, 64 * wpg, ", ".join(args)))
    else:
        o.append("    hipLaunchKernelGGL(%s, grid, dim3(%d), 0, (hipStream_t)stream, %s);" % (name, 64 * wpg, ", ".join(args)))
    o.append("    return (int)hipGetLastError();")
    o.append("}")
    o.append("")
    # the public launcher: the partial sums go to the workspace the library
    # keeps for this (device, stream); with list validation switched on, both
    # lists are checked against the block contract first
    o.append('extern "C" int vlxmi_launch_%s(%s,\n    uint32_t ket_split, void* stream)' % (name, ",\n    ".join(lsig)))
    o.append("{")
    o += checks + clamp
    o.append("    if (vlxmi_list_validation()) {")
    o.append("        int rc = vlxmi_validate_j_list(%d, %d, %s_first_inds_local, %s_second_inds_local, %s_mat_Q_local, nullptr, %s_prim_pair_count_local, stream);"
             % (la, lb, ab, ab, ab, ab))
    o.append("        if (rc == 0) rc = vlxmi_validate_j_list(%d, %d, %s_first_inds, %s_second_inds, %s_mat_Q, %s_mat_D, %s_prim_pair_count, stream);"
             % (lc, ld, cd, cd, cd, cd, cd))
    o.append("        if (rc != 0) return rc;")
    o.append("    }")
    o.append("    double* part = nullptr;")
    o.append("    if (ket_split > 1) {")
    o.append("        const int rc = vlxmi_ket_workspace((size_t)ket_split * %s_prim_pair_count_local, stream, &part);" % ab)
    o.append("        if (rc != 0) return rc;")
    o.append("    }")
    o.append("    return vlxmi_jpart_%s(%s, ket_split, part, stream);" % (name, ", ".join(args)))
    o.append("}")
    return "\n".join(o) + "\n"


# --------------------------------------------------------------------------
# K kernel
# --------------------------------------------------------------------------

def emit_k(cls, mixed=False, batched=False, kt=None, layout="reg", dorder="jl", lds_boys=False, wpg=1, kpairs=False):
    """Source of computeExchangeFock{CLS} (or its _mixed / _batched variants) and
    its launcher.  The batched variant is persistent over a list of (i,k) blocks
    of lists that concatenate many molecules; per block, blk_mol gives the
    molecule, whose density lives at mat_D_full_AO + mol_D_off[m] (mol_naos[m]
    Cartesian AOs, AO indices offset by mol_ao_off[m]) and whose maxD_bd is
    mol_maxD[m].

    ``dorder`` is how the density element D_jl of the digestion is read:
    "jl" reads D[j][l] as the reference does (gen_exchange.py:323-342), so the
    64 lanes of a wave (64 partners j, one ket partner l) read 64 rows; "lj"
    reads D[l][j], so the wave reads one row of D at 64 columns.  The two agree
    for a symmetric density (every SCF density; the host symmetrises it
    exactly), and the faster one per class is a measured choice."""
    a, b, c, d = cls
    la, lb, lc, ld = LV[a], LV[b], LV[c], LV[d]
    na, nb, nc, nd = NC[a], NC[b], NC[c], NC[d]
    ac, ab, cd, bd = a + c, a + b, c + d, b + d
    nac, nab, ncd = na * nc, na * nb, nc * nd
    L = la + lb + lc + ld
    rec = KREC0 + 2 * ncd
    rec += rec % 2
    kslices = plan_slices(la, lb, lc, ld, "Kheavy")
    heavy = len(kslices) > 1
    if heavy and kpairs:
        # one pass per ket component pair (ic, id): the recursion is recomputed
        # per pair, and only one pair's intermediates are live at a time
        assert not mixed, "pair passes are for the FP64 kernels"
        kslices = [[(ic, idd)] for ic in range(nc) for idd in range(nd)]
    name = "computeExchangeFock" + cls.upper() + ("_batched" if batched else "") + ("_mixed" if mixed else "")
    bsig = ["const uint32_t* __restrict__ blk_list", "const uint32_t nlist", "const uint32_t* __restrict__ blk_mol",
            "const double* __restrict__ mol_maxD", "const unsigned long long* __restrict__ mol_D_off",
            "const uint32_t* __restrict__ mol_naos", "const uint32_t* __restrict__ mol_ao_off"]
    sig = k_signature(cls) + (bsig if batched else []) + tail_args(mixed, batched)
    ty = "double"
    o = []
    assert not (heavy and wpg > 1), "the heavy layouts keep one wavefront per workgroup"
    o.append("__global__ void __launch_bounds__(%s)" % ("TILE_SIZE_K" if wpg == 1 else str(64 * wpg)))
    o.append("%s(%s)" % (name, ",\n    ".join(sig)))
    o.append("{")
    o.append("    constexpr int NA = %d, NB = %d, NC = %d, ND = %d, NAC = %d, NAB = %d, NCD = %d;" % (na, nb, nc, nd, nac, nab, ncd))
    o.append("    constexpr int KT = %d, REC = %d;" % (LDS_BOYS_KT if lds_boys else (kt or (32 if heavy else KET_TILE)), rec))
    if wpg == 1:
        o.append("    __shared__ double ket[KT * REC];")
        o.append("    __shared__ uint32_t ket_cg[KT * ND];")
    else:
        o.append("    constexpr int WPG = %d;" % wpg)
        o.append("    __shared__ double ket_all[WPG * KT * REC];")
        o.append("    __shared__ uint32_t ket_cg_all[WPG * KT * ND];")
    # heavy classes keep K accumulators in LDS when they fit beside the
    # thresholds, S_ab and an 8-record ket tile within one wave's LDS share
    # Heavy classes have two layouts, chosen per class by measurement:
    #   "reg": K accumulators in registers, a-major digestion, the pass's density
    #          gathers loaded together after the Boys evaluation;
    #   "lds": K accumulators in LDS (a pass keeps its own NA in registers) when
    #          they fit beside the thresholds, S_ab and an 8-record ket tile in one
    #          wave's LDS share, and b-major digestion with each b's density row
    #          loaded just before its block.
    lds_layout = heavy and layout == "lds"
    lacc = lds_layout and (16 * nab + 8 * nac) * 64 + 8 * ket_record_bytes("K", cls) <= LDS_PER_WAVE_FULL_REGS
    if heavy:
        # per-lane thresholds and S_ab live in LDS ([component][lane]) to free registers
        o.append("    __shared__ double ltau[NAB * TILE_SIZE_K], lSab[NAB * TILE_SIZE_K];")
    if lacc:
        # per-lane K accumulators in LDS ([component][lane]); a pass keeps its
        # own NA accumulators in registers for the duration of its ket loop
        o.append("    __shared__ double lacc[NAC * TILE_SIZE_K];")
    o += wave_decls(wpg)
    if wpg > 1:
        o.append("    double* ket = ket_all + wave * (KT * REC);")
        o.append("    uint32_t* ket_cg = ket_cg_all + wave * (KT * ND);")
    if lds_boys:
        o += lds_boys_setup(L)
    if mixed or batched:
        o += count_decl()
    if batched:
        o.append("    (void)%s_max_D;" % bd)
        if wpg == 1:
            o.append("    for (uint32_t w = blockIdx.x; w < nlist; w += gridDim.x) {")
        else:
            o.append("    for (uint32_t w = blockIdx.x * WPG + wave; w < nlist; w += gridDim.x * WPG) {")
        o.append("    const uint32_t blk = blk_list[w];")
        o.append("    if ((blk + 1) * NAC > pair_inds_count_for_K_%s) continue;" % ac)
        o.append("    const uint32_t mol = blk_mol[blk];")
        o.append("    const double maxD = mol_maxD[mol];")
        o.append("    if (!(maxD > 0.0)) continue;")
        o.append("    const double* __restrict__ Dm = mat_D_full_AO + mol_D_off[mol];")
        o.append("    const uint32_t nao = mol_naos[mol], aoff = mol_ao_off[mol];")
    else:
        o.append("    const uint32_t blk = %s;" % ("blockIdx.x" if wpg == 1 else "blockIdx.x * WPG + wave"))
        o.append("    if ((blk + 1) * NAC > pair_inds_count_for_K_%s) return;" % ac)
        o.append("    const double maxD = %s_max_D;" % bd)
        o.append("    if (!(maxD > 0.0)) return;")
        o.append("    const double* __restrict__ Dm = mat_D_full_AO;")
        o.append("    const uint32_t nao = naos, aoff = 0;")
    o.append("    const double thr = eri_threshold;")
    o.append("    const double thr_eff = thr / maxD;")
    o.append("    const uint32_t e0 = blk * NAC;")
    o.append("    const uint32_t I = pair_inds_i_for_K_%s[e0] / NA;" % ac)
    o.append("    const uint32_t K = pair_inds_k_for_K_%s[e0] / NC;" % ac)
    o.append("    const double ai = %s_prim_info[I];" % a)
    o.append("    const double A_x = %s_prim_info[I + %s_prim_count * 2], A_y = %s_prim_info[I + %s_prim_count * 3], A_z = %s_prim_info[I + %s_prim_count * 4];" % ((a,) * 6))
    o.append("    const double ak = %s_prim_info[K];" % c)
    o.append("    const double C_x = %s_prim_info[K + %s_prim_count * 2], C_y = %s_prim_info[K + %s_prim_count * 3], C_z = %s_prim_info[K + %s_prim_count * 4];" % ((c,) * 6))
    o.append("    uint32_t displ_i[NA], displ_k[NC];")
    o.append("    #pragma unroll")
    o.append("    for (int t = 0; t < NA; ++t) displ_i[t] = pair_displs_K_%s[I * NA + t];" % ab)
    o.append("    #pragma unroll")
    o.append("    for (int t = 0; t < NC; ++t) displ_k[t] = pair_displs_K_%s[K * NC + t];" % cd)
    o.append("    const uint32_t nJ = pair_counts_K_%s[I * NA] / NB;" % ab)
    o.append("    const uint32_t nL = pair_counts_K_%s[K * NC] / ND;" % cd)
    if lacc:
        o.append("    #pragma unroll")
        o.append("    for (int t = 0; t < NAC; ++t) lacc[t * TILE_SIZE_K + lane] = 0.0;")
    else:
        o.append("    double acc[NAC];")
        o.append("    #pragma unroll")
        o.append("    for (int t = 0; t < NAC; ++t) acc[t] = 0.0;")
    o.append("    for (uint32_t jc = 0; jc < nJ; jc += TILE_SIZE_K) {")
    o.append("        const uint32_t jp = jc + lane;")
    o.append("        bool active = jp < nJ;")
    o.append("        double zeta = 1.0, sq_zeta = 1.0, inv_S1 = 1.0, P_x = 0.0, P_y = 0.0, P_z = 0.0;")
    o.append("        double PA_x = 0.0, PA_y = 0.0, PA_z = 0.0, AB_x = 0.0, AB_y = 0.0, AB_z = 0.0;")
    o.append("        double qij = 0.0;")
    if heavy:
        o.append("        double* tau = ltau;")
        o.append("        double* Sab = lSab;")
        o.append("        #pragma unroll")
        o.append("        for (int t = 0; t < NAB; ++t) { tau[t * TILE_SIZE_K + lane] = 0.0; Sab[t * TILE_SIZE_K + lane] = 0.0; }")
    else:
        o.append("        double tau[NAB], Sab[NAB];")
        o.append("        #pragma unroll")
        o.append("        for (int t = 0; t < NAB; ++t) { tau[t] = 0.0; Sab[t] = 0.0; }")
    o.append("        uint32_t cgJ[NB];")
    o.append("        #pragma unroll")
    o.append("        for (int t = 0; t < NB; ++t) cgJ[t] = 0;")
    o.append("        if (active) {")
    o.append("            const uint32_t jcomb = D_inds_K_%s[displ_i[0] + jp * NB];" % ab)
    o.append("            const uint32_t J = jcomb / NB;")
    o.append("            const double aj = %s_prim_info[J];" % b)
    o.append("            const double B_x = %s_prim_info[J + %s_prim_count * 2], B_y = %s_prim_info[J + %s_prim_count * 3], B_z = %s_prim_info[J + %s_prim_count * 4];" % ((b,) * 6))
    o.append("            zeta = ai + aj; inv_S1 = 1.0 / zeta; sq_zeta = sqrt(zeta);")
    o.append("            P_x = (ai * A_x + aj * B_x) * inv_S1; P_y = (ai * A_y + aj * B_y) * inv_S1; P_z = (ai * A_z + aj * B_z) * inv_S1;")
    o.append("            PA_x = (aj * inv_S1) * (B_x - A_x); PA_y = (aj * inv_S1) * (B_y - A_y); PA_z = (aj * inv_S1) * (B_z - A_z);")
    o.append("            AB_x = A_x - B_x; AB_y = A_y - B_y; AB_z = A_z - B_z;")
    o.append("            #pragma unroll")
    o.append("            for (int ca = 0; ca < NA; ++ca) {")
    o.append("                #pragma unroll")
    o.append("                for (int cb = 0; cb < NB; ++cb) {")
    o.append("                    const uint32_t pos = displ_i[ca] + jp * NB + cb;")
    o.append("                    const double q = fabs(Q_K_%s[pos]);" % ab)
    o.append("                    qij = fmax(qij, q);")
    if heavy:
        o.append("                    tau[(ca * NB + cb) * TILE_SIZE_K + lane] = thr_eff / q;")
        o.append("                    Sab[(ca * NB + cb) * TILE_SIZE_K + lane] = pair_data_K_%s[pos];" % ab)
    else:
        o.append("                    tau[ca * NB + cb] = thr_eff / q;")
        o.append("                    Sab[ca * NB + cb] = pair_data_K_%s[pos];" % ab)
    o.append("                }")
    o.append("            }")
    o.append("            #pragma unroll")
    o.append("            for (int cb = 0; cb < NB; ++cb) {")
    o.append("                const uint32_t jj = D_inds_K_%s[displ_i[0] + jp * NB + cb];" % ab)
    if b == "s":
        o.append("                cgJ[cb] = %s_prim_aoinds[jj] - aoff;" % b)
    else:
        o.append("                cgJ[cb] = %s_prim_aoinds[(jj / NB) + %s_prim_count * (jj %% NB)] - aoff;" % (b, b))
    o.append("            }")
    o.append("        }")
    if mixed:
        o.append("        bool lo_mode = false;")
    o.append("        for (uint32_t t0 = 0; t0 < nL; t0 += KT) {")
    o.append("            if (!__any(active)) break;")
    o.append("            " + sync(wpg))
    o.append("            {")
    o.append("                const uint32_t lp = t0 + lane;")
    o.append("                double* R = ket + lane * REC;")
    o.append("                if (lane < KT && lp < nL) {")
    o.append("                    const uint32_t lcomb = D_inds_K_%s[displ_k[0] + lp * ND];" % cd)
    o.append("                    const uint32_t Lp = lcomb / ND;")
    o.append("                    const double al = %s_prim_info[Lp];" % d)
    o.append("                    const double D_x = %s_prim_info[Lp + %s_prim_count * 2], D_y = %s_prim_info[Lp + %s_prim_count * 3], D_z = %s_prim_info[Lp + %s_prim_count * 4];" % ((d,) * 6))
    o.append("                    const double et = ak + al, iet = 1.0 / et;")
    o.append("                    R[1] = et; R[2] = iet; R[12] = sqrt(et);")
    o.append("                    R[3] = (ak * C_x + al * D_x) * iet; R[4] = (ak * C_y + al * D_y) * iet; R[5] = (ak * C_z + al * D_z) * iet;")
    o.append("                    R[6] = (al * iet) * (D_x - C_x); R[7] = (al * iet) * (D_y - C_y); R[8] = (al * iet) * (D_z - C_z);")
    o.append("                    R[9] = C_x - D_x; R[10] = C_y - D_y; R[11] = C_z - D_z;")
    o.append("                    double qk = 0.0;")
    o.append("                    #pragma unroll")
    o.append("                    for (int cc = 0; cc < NC; ++cc) {")
    o.append("                        #pragma unroll")
    o.append("                        for (int cdd = 0; cdd < ND; ++cdd) {")
    o.append("                            const uint32_t pos = displ_k[cc] + lp * ND + cdd;")
    o.append("                            const double q = fabs(Q_K_%s[pos]);" % cd)
    o.append("                            qk = fmax(qk, q);")
    o.append("                            R[%d + cc * ND + cdd] = pair_data_K_%s[pos];" % (KREC0, cd))
    o.append("                            R[%d + NCD + cc * ND + cdd] = q;" % KREC0)
    o.append("                        }")
    o.append("                    }")
    o.append("                    R[0] = qk;")
    o.append("                    #pragma unroll")
    o.append("                    for (int cdd = 0; cdd < ND; ++cdd) {")
    o.append("                        const uint32_t ll = D_inds_K_%s[displ_k[0] + lp * ND + cdd];" % cd)
    if d == "s":
        o.append("                        ket_cg[lane * ND + cdd] = %s_prim_aoinds[ll] - aoff;" % d)
    else:
        o.append("                        ket_cg[lane * ND + cdd] = %s_prim_aoinds[(ll / ND) + %s_prim_count * (ll %% ND)] - aoff;" % (d, d))
    o.append("                    }")
    o.append("                }")
    o.append("            }")
    o.append("            " + sync(wpg))
    o.append("            const uint32_t nt = (nL - t0 < (uint32_t)KT) ? (nL - t0) : (uint32_t)KT;")
    assert dorder in ("jl", "lj")

    def dread(jrow, idd):
        """Density element D_jl of partner AO ``jrow`` and the staged ket AO idd."""
        if dorder == "jl":
            return "Dm[(size_t)%s * nao + ket_cg[s * ND + %d]]" % (jrow, idd)
        return "Dm[(size_t)ket_cg[s * ND + %d] * nao + %s]" % (idd, jrow)

    def comps(group):
        """(ic, [id...]) of a pass: an element ic stands for all id."""
        out = {}
        for e in group:
            ic, ids = (e[0], [e[1]]) if isinstance(e, tuple) else (e, list(range(nd)))
            out.setdefault(ic, []).extend(ids)
        return list(out.items())

    def kbody(tyq, ind, groups):
        # the density gathers follow the Boys evaluation: vector loads complete
        # in order, so gathers issued before the Boys table loads only lengthen
        # the wait for the table and the live ranges
        lines = quartet_prologue(L, ty_geo=tyq, omega=True, ind=ind, lds_boys=lds_boys)
        TAU = (lambda t: "tau[%d * TILE_SIZE_K + lid]" % t) if heavy else (lambda t: "tau[%d]" % t)
        SAB = (lambda t: "Sab[%d * TILE_SIZE_K + lid]" % t) if heavy else (lambda t: "Sab[%d]" % t)
        if heavy:
            # opaque copies of the lane index and the D row offsets: the loads of
            # one pass can be neither hoisted out of its ket loop nor shared
            # with another pass
            lines += [ind + "uint32_t lid = lane;", ind + "asm volatile(\"\" : \"+v\"(lid));"]
            for ibb in range(nb):
                lines += [ind + "uint32_t rowJ%d = cgJ[%d];" % (ibb, ibb), ind + "asm volatile(\"\" : \"+v\"(rowJ%d));" % ibb]
        # b-major digestion: each b's density row loaded just before its block
        # (an opaque row index keeps the loads from being hoisted)
        bouter = lds_layout
        ACC = (lambda t: "pacc%d" % t) if lacc else (lambda t: "acc[%d]" % t)
        used = sorted({(ic_, d_) for g_ in groups for ic_, ds in comps(g_) for d_ in ds})
        if not bouter:
            for ibb in range(nb):
                for idd in sorted({d_ for _, d_ in used}):
                    lines.append(ind + "const double Dj%d_%d = %s;"
                                 % (ibb, idd, dread(("rowJ%d" % ibb) if heavy else ("cgJ[%d]" % ibb), idd)))
        # digestion: acc[a,c] += Sab * sum_{b,d} mask * X * Sc * D  (FP64 accumulation)
        for icc, idd in used:
            tcd = icc * nd + idd
            lines.append(ind + "const double Sc%d = R[%d], Qk%d = R[%d];" % (tcd, KREC0 + tcd, tcd, KREC0 + ncd + tcd))
        for isl, group in enumerate(groups):
            if isl:
                lines += [ind + x for x in launder(L)]
            dag = QuartetDAG(la, lb, lc, ld, ty=tyq, rule_vrr="first", rule_hrr="first", prefix="q%d_" % isl)
            for ic, idds in comps(group):
                cc = tri(lc)[ic]
                order = ([(ia, ib) for ib in range(nb) for ia in range(na)] if bouter
                         else [(ia, ib) for ia in range(na) for ib in range(nb)])
                for ia, ib in order:
                    aa, bbx = tri(la)[ia], tri(lb)[ib]
                    tac = ia * nc + ic
                    tab = ia * nb + ib
                    dname = (lambda b_, d_: "Dj%d_%d_c%d" % (b_, d_, ic)) if bouter else (lambda b_, d_: "Dj%d_%d" % (b_, d_))
                    if bouter and ia == 0:
                        # this b's density row, loaded just before its block
                        dag.lines.append("uint32_t rowB%d_%d_c%d = cgJ[%d];" % (isl, ib, ic, ib))
                        dag.lines.append("asm volatile(\"\" : \"+v\"(rowB%d_%d_c%d));" % (isl, ib, ic))
                        for idd in idds:
                            dag.lines.append("const double %s = %s;"
                                             % (dname(ib, idd), dread("rowB%d_%d_c%d" % (isl, ib, ic), idd)))
                    terms = []
                    for idd in idds:
                        dd = tri(ld)[idd]
                        tcd = ic * nd + idd
                        x = dag.X(aa, bbx, cc, dd)
                        xv = x if tyq == "double" else "double(%s)" % x
                        terms.append("((Qk%d > %s) ? (%s * Sc%d * %s) : 0.0)" % (tcd, TAU(tab), xv, tcd, dname(ib, idd)))
                    dag.lines.append("%s += %s * (%s);" % (ACC(tac), SAB(tab), " + ".join(terms)))
            lines += [ind + x for x in _dag_lines(dag)]
        return lines

    kscreen = "qij * R[0] * maxD * (1.0 + 1e-13) <= thr"
    kbound = "qij * R[0] * maxD"
    def pacc_io(group, store):
        out = []
        if lacc:
            for ic, _ in comps(group):
                for ia in range(na):
                    t = ia * nc + ic
                    if store:
                        out.append("            lacc[%d * TILE_SIZE_K + lane] = pacc%d;" % (t, t))
                    else:
                        out.append("            double pacc%d = lacc[%d * TILE_SIZE_K + lane];" % (t, t))
        return out

    kslices32 = plan_slices_f32(la, lb, lc, ld, "Kheavy") if mixed else kslices
    if heavy and len(kslices32) == len(kslices):
        # heavy classes: one loop over the tile per pass of ket-first-index
        # components, so that only one pass's recursion is live at a time; the
        # screening sequence is the same in every pass, so each pass stops at
        # the same record (mixed: each pass's FP64 loop, then its FP32 loop)
        o.append("            bool act_end = active;")
        if mixed:
            o.append("            bool lo_end = lo_mode;")
        for isl, group in enumerate(kslices):
            o.append("            {")
            o.append("            bool act%d = active;" % isl)
            if mixed:
                o.append("            bool lo%d = lo_mode;" % isl)
            o += pacc_io(group, False)
            o += ket_loops("            ", "act%d" % isl, kscreen, kbound, KET_HEADER,
                           lambda ty, ind, g=group: kbody(ty, ind, [g]), mixed, "lo%d" % isl,
                           count=(mixed or batched) and isl == 0)
            o += pacc_io(group, True)
            o.append("            act_end = act%d;" % isl)
            if mixed:
                o.append("            lo_end = lo%d;" % isl)
            o.append("            }")
        o.append("            active = act_end;")
        if mixed:
            o.append("            lo_mode = lo_end;")
    elif heavy:
        # mixed heavy classes whose FP32 recursion fits in fewer passes: the
        # FP64 passes (one loop each) run up to the first record where the
        # whole wavefront is below fp32_threshold (the same record in every
        # pass: identical screening), then the FP32 passes run from there
        o.append("            uint32_t s_sw = 0;")
        o.append("            bool act_sw = active, lo_sw = lo_mode;")
        for isl, group in enumerate(kslices):
            o.append("            {")
            o.append("            bool act%d = active, lo%d = lo_mode;" % (isl, isl))
            o += pacc_io(group, False)
            o.append("            uint32_t s = 0;")
            o.append("            if (!lo%d) {" % isl)
            o += ["    " + x for x in ket_loop("            ", "act%d" % isl, kscreen, kbound, KET_HEADER,
                                             lambda ty, ind, g=group: kbody(ty, ind, [g]), "double", True,
                                             "lo%d" % isl, count=isl == 0)]
            o.append("            }")
            o += pacc_io(group, True)
            o.append("            s_sw = s; act_sw = act%d; lo_sw = lo%d;" % (isl, isl))
            o.append("            }")
        o.append("            bool act_end = act_sw;")
        o.append("            if (lo_sw) {")
        for isl, group in enumerate(kslices32):
            o.append("            {")
            o.append("            bool actf%d = act_sw;" % isl)
            o += pacc_io(group, False)
            o.append("            uint32_t s = s_sw;")
            o += ket_loop("            ", "actf%d" % isl, kscreen, kbound, KET_HEADER,
                          lambda ty, ind, g=group: kbody(ty, ind, [g]), "float", False, "lo_sw", count=isl == 0)
            o += pacc_io(group, True)
            o.append("            act_end = actf%d;" % isl)
            o.append("            }")
        o.append("            }")
        o.append("            active = act_end;")
        o.append("            lo_mode = lo_sw;")
    else:
        o += ket_loops("            ", "active", kscreen, kbound, KET_HEADER,
                       lambda ty, ind: kbody(ty, ind, [list(range(nc))]), mixed, "lo_mode",
                       count=mixed or batched)
    o.append("        }")
    o.append("    }")
    o.append("    #pragma unroll")
    o.append("    for (int t = 0; t < NAC; ++t) {")
    o.append("        double v = %s;" % ("lacc[t * TILE_SIZE_K + lane]" if lacc else "acc[t]"))
    o.append("        #pragma unroll")
    o.append("        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);")
    o.append("        if (lane == 0) mat_K[e0 + t] += v;")
    o.append("    }")
    if batched:
        o.append("    " + sync(wpg))
        o.append("    }")
    if mixed or batched:
        o += count_flush()
    o.append("}")
    o.append("")
    lsig = _strip_restrict(sig)
    args = _argnames(sig)
    if batched:
        o.append('extern "C" int vlxmi_launch_%s(%s,\n    uint32_t n_workgroups, void* stream)' % (name, ",\n    ".join(lsig)))
    else:
        o.append('extern "C" int vlxmi_launch_%s(%s,\n    void* stream)' % (name, ",\n    ".join(lsig)))
    o.append("{")
    o.append("    if (pair_inds_count_for_K_%s %% %d != 0) return VLXMI_ERR_LAYOUT;" % (ac, nac))
    o.append("    const uint32_t nblk = pair_inds_count_for_K_%s / %d;" % (ac, nac))
    o.append("    if (nblk == 0) return VLXMI_OK;")
    o.append("    if (!mat_K || !mat_D_full_AO || !boys_func_table) return VLXMI_ERR_ARG;")
    if not batched:
        # with list validation switched on: the (i,k) list and both CSRs
        o.append("    if (vlxmi_list_validation()) {")
        o.append("        int rc = vlxmi_validate_ik_list(%d, %d, pair_inds_i_for_K_%s, pair_inds_k_for_K_%s, pair_inds_count_for_K_%s, stream);"
                 % (la, lc, ac, ac, ac))
        o.append("        if (rc == 0) rc = vlxmi_validate_k_csr(%d, %d, %s_prim_count, pair_displs_K_%s, pair_counts_K_%s, Q_K_%s, D_inds_K_%s, stream);"
                 % (la, lb, a, ab, ab, ab, ab))
        if cd != ab:
            o.append("        if (rc == 0) rc = vlxmi_validate_k_csr(%d, %d, %s_prim_count, pair_displs_K_%s, pair_counts_K_%s, Q_K_%s, D_inds_K_%s, stream);"
                     % (lc, ld, c, cd, cd, cd, cd))
        o.append("        if (rc != 0) return rc;")
        o.append("    }")
    if batched:
        o.append("    if (nlist == 0) return VLXMI_OK;")
        o.append("    if (!blk_list || !blk_mol || !mol_maxD || !mol_D_off || !mol_naos || !mol_ao_off) return VLXMI_ERR_ARG;")
        o.append("    const uint32_t max_wg = (nlist + %d) / %d;" % (wpg - 1, wpg))
        o.append("    if (n_workgroups == 0 || n_workgroups > max_wg) n_workgroups = max_wg;")
        o.append("    hipLaunchKernelGGL(%s, dim3(n_workgroups), dim3(%d), 0, (hipStream_t)stream, %s);" % (name, 64 * wpg, ", ".join(args)))
    else:
        o.append("    hipLaunchKernelGGL(%s, dim3((nblk + %d) / %d), dim3(%d), 0, (hipStream_t)stream, %s);" % (name, wpg - 1, wpg, 64 * wpg, ", ".join(args)))
    o.append("    return (int)hipGetLastError();")
    o.append("}")
    return "\n".join(o) + "\n"


# --------------------------------------------------------------------------
# Schwarz kernel
# --------------------------------------------------------------------------

def emit_schwarz(pc):
    a, b = pc
    la, lb = LV[a], LV[b]
    na, nb = NC[a], NC[b]
    nab = na * nb
    L = 2 * (la + lb)
    name = "vlxmi_schwarz_%s" % pc
    o = []
    o.append("__global__ void __launch_bounds__(64)")
    o.append("%s_kernel(const double* __restrict__ a_info, const uint32_t a_count, const double* __restrict__ b_info, const uint32_t b_count," % name)
    o.append("    const uint32_t* __restrict__ Iblk, const uint32_t* __restrict__ Jblk, const double* __restrict__ S, const uint32_t nblk,")
    o.append("    double* __restrict__ Qout, const double* __restrict__ boys_func_table)")
    o.append("{")
    o.append("    const uint32_t t = blockIdx.x * 64 + threadIdx.x;")
    o.append("    if (t >= nblk) return;")
    o.append("    const uint32_t I = Iblk[t], J = Jblk[t];")
    o.append("    const double ai = a_info[I], aj = b_info[J];")
    o.append("    const double A_x = a_info[I + a_count * 2], A_y = a_info[I + a_count * 3], A_z = a_info[I + a_count * 4];")
    o.append("    const double B_x = b_info[J + b_count * 2], B_y = b_info[J + b_count * 3], B_z = b_info[J + b_count * 4];")
    o.append("    const double zeta = ai + aj, inv_S1 = 1.0 / zeta, sq_zeta = sqrt(zeta);")
    o.append("    const double P_x = (ai * A_x + aj * B_x) * inv_S1, P_y = (ai * A_y + aj * B_y) * inv_S1, P_z = (ai * A_z + aj * B_z) * inv_S1;")
    o.append("    const double PA_x = (aj * inv_S1) * (B_x - A_x), PA_y = (aj * inv_S1) * (B_y - A_y), PA_z = (aj * inv_S1) * (B_z - A_z);")
    o.append("    const double AB_x = A_x - B_x, AB_y = A_y - B_y, AB_z = A_z - B_z;")
    o.append("    const double eta = zeta, inv_eta = inv_S1, sq_eta = sq_zeta;")
    o.append("    const double Q_x = P_x, Q_y = P_y, Q_z = P_z;")
    o.append("    const double QC_x = PA_x, QC_y = PA_y, QC_z = PA_z;")
    o.append("    const double CD_x = AB_x, CD_y = AB_y, CD_z = AB_z;")
    o += quartet_prologue(L, ind="    ")
    dag = QuartetDAG(la, lb, la, lb, rule_vrr="first", rule_hrr="first")
    for ia, aa in enumerate(tri(la)):
        for ib, bbx in enumerate(tri(lb)):
            tab = ia * nb + ib
            x = dag.X(aa, bbx, aa, bbx)
            dag.lines.append("{ const double s = S[t * %d + %d]; Qout[t * %d + %d] = sqrt(fabs(%s * s * s)); }" % (nab, tab, nab, tab, x))
    o += ["    " + x for x in _dag_lines(dag)]
    o.append("}")
    o.append("")
    o.append('extern "C" int %s(const double* a_info, uint32_t a_count, const double* b_info, uint32_t b_count,' % name)
    o.append("    const uint32_t* Iblk, const uint32_t* Jblk, const double* S, uint32_t nblk, double* Qout,")
    o.append("    const double* boys_func_table, void* stream)")
    o.append("{")
    o.append("    if (nblk == 0) return VLXMI_OK;")
    o.append("    if (!a_info || !b_info || !Iblk || !Jblk || !S || !Qout || !boys_func_table) return VLXMI_ERR_ARG;")
    o.append("    hipLaunchKernelGGL(%s_kernel, dim3((nblk + 63) / 64), dim3(64), 0, (hipStream_t)stream," % name)
    o.append("        a_info, a_count, b_info, b_count, Iblk, Jblk, S, nblk, Qout, boys_func_table);")
    o.append("    return (int)hipGetLastError();")
    o.append("}")
    return "\n".join(o) + "\n"
