/*
 * FP64 CPU oracle for the Obara-Saika J/K Fock path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity reference: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (vlx-twoe-ints_amd/) never links or calls it.
 *
 * It restates, with hand-written generic loops, the arithmetic that the
 * reference's emitted kernels perform:
 *   - pair prefactors, PA/PB/QC/QD, PQ = Q - P (note the sign), Lambda =
 *     sqrt(4 rho / pi), Boys argument rho |PQ|^2          gen_coulomb.py:182-201,244-268
 *   - erf-attenuated variant (omega != 0): d2 = w^2/(rho+w^2), T = rho d2 r2,
 *     Lambda = sqrt(4 rho d2 / pi), F_m *= d2^m          gen_exchange.py:278-298
 *   - VRR on a                                            eri.py:297-355
 *   - VRR on c (with the a-1 cross term)                  eri.py:237-295
 *   - HRR on b and d (AB = PB - PA, CD = QD - QC)         eri.py:177-235
 *   - recursion order HRR-d, HRR-b, VRR-c, VRR-a (top-down), i.e. VRR-a,
 *     VRR-c, HRR-b, HRR-d bottom-up                        eriwriter.py:70-101
 *   - max Boys order L = la+lb+lc+ld                       eriwriter.py:116-119
 *   - J digestion mat_J[ij] += eri * D_kl * w, w = 2 if c != d else (k!=l)+1,
 *     screened by |Q_ij Q_kl D_kl| > thr                   gen_coulomb.py:205-212,276-296
 *   - K digestion mat_K[ik] += sum_j sum_l eri * D_AO[cg(j), cg(l)],
 *     screened by |Q_ij Q_kl maxD_bd| > thr                gen_exchange.py:189-232,323-341
 *   - d Cartesian order xx,xy,xz,yy,yz,zz                  gen_coulomb.py:143-148
 *
 * The Boys function (VeloxChem's gpu::computeBoysFunction, not part of the
 * reference tree) is evaluated here from its definition: a convergent series
 * for T <= 30 followed by downward recursion, and erf + upward recursion for
 * T > 30, both in long double.
 *
 * Two screening modes are provided:
 *   mode 0 ("component"): the reference rule, per Cartesian component pair.
 *   mode 1 ("block"): the rule of the HIP path, per primitive-pair block:
 *     J includes a whole quartet block iff Qmax_IJ * beta_KL > thr with
 *     Qmax_IJ = max |Q_ij| over the bra block and beta_KL = max |Q_kl D_kl| over
 *     the ket block; K iff QKmax_IJ * QKmax_KL * maxD_bd > thr.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Sizes cover the derivative classes too: one shell raised to f (l = 3) for
 * the geometric derivatives, so la+lb, lc+ld <= 5 and L <= 9. */
#define MAXL_SIDE 5           /* la+lb, lc+ld <= 5 */
#define NTRI 56               /* triples with |e| <= 5 */
#define MAXM 10               /* m = 0..10 */
#define NSH 10                /* Cartesian components of one shell, l <= 3 */

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------------ */
/* Cartesian component bookkeeping                                          */
/* ------------------------------------------------------------------------ */

static int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

/* canonical order within shell n: lx descending, then ly descending */
static int tri_pos(int ex, int ey, int ez) {
    int n = ex + ey + ez;
    return (n - ex) * (n - ex + 1) / 2 + ez;
}
static int tri_index(int ex, int ey, int ez) {
    int n = ex + ey + ez;
    return n * (n + 1) * (n + 2) / 6 + tri_pos(ex, ey, ez);
}
static void tri_comp(int l, int pos, int* e) {
    int p = 0;
    e[0] = l; e[1] = 0; e[2] = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            if (p == pos) { e[0] = lx; e[1] = ly; e[2] = l - lx - ly; return; }
            ++p;
        }
}

/* ------------------------------------------------------------------------ */
/* Boys function                                                            */
/* ------------------------------------------------------------------------ */

void orc_boys(int mmax, double T, double* F) {
    long double t = (long double)T;
    if (t < 1e-300L) {
        for (int m = 0; m <= mmax; ++m) F[m] = 1.0 / (2.0 * m + 1.0);
        return;
    }
    if (t <= 30.0L) {
        /* F_M(T) = exp(-T) sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)) */
        int M = mmax;
        long double term = 1.0L / (2.0L * M + 1.0L), sum = term;
        for (int k = 1; k < 2000; ++k) {
            term *= 2.0L * t / (2.0L * M + 2.0L * k + 1.0L);
            sum += term;
            if (term < sum * 1e-21L) break;
        }
        long double et = expl(-t);
        long double Fm = et * sum;
        F[M] = (double)Fm;
        for (int m = M - 1; m >= 0; --m) {
            Fm = (2.0L * t * Fm + et) / (2.0L * m + 1.0L);
            F[m] = (double)Fm;
        }
        return;
    }
    long double et = expl(-t);
    long double st = sqrtl(t);
    long double Fm = 0.5L * sqrtl((long double)ORC_PI / t) * erfl(st);
    F[0] = (double)Fm;
    for (int m = 0; m < mmax; ++m) {
        Fm = ((2.0L * m + 1.0L) * Fm - et) / (2.0L * t);
        F[m + 1] = (double)Fm;
    }
}

/* ------------------------------------------------------------------------ */
/* One primitive quartet: Lambda * polynomial for every Cartesian component */
/* out[((ia*nb + ib)*nc + ic)*nd + id]; no contraction/normalisation (S)    */
/* ------------------------------------------------------------------------ */

void orc_eri_block(int la, int lb, int lc, int ld,
                   double ai, const double* A, double aj, const double* B,
                   double ak, const double* C, double al, const double* D,
                   double omega, double* out)
{
    const int L = la + lb + lc + ld;
    const int Le = la + lb, Lf = lc + ld;
    const double S1 = ai + aj, S2 = ak + al, S4 = S1 + S2;
    const double inv_S1 = 1.0 / S1, inv_S2 = 1.0 / S2, inv_S4 = 1.0 / S4;
    double P[3], Q[3], PA[3], QC[3], PQ[3], AB[3], CD[3];
    for (int x = 0; x < 3; ++x) {
        P[x] = (ai * A[x] + aj * B[x]) * inv_S1;
        Q[x] = (ak * C[x] + al * D[x]) * inv_S2;
        PA[x] = (aj * inv_S1) * (B[x] - A[x]);
        QC[x] = (al * inv_S2) * (D[x] - C[x]);
        PQ[x] = Q[x] - P[x];
        AB[x] = A[x] - B[x];
        CD[x] = C[x] - D[x];
    }
    const double r2 = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
    const double rho = S1 * S2 * inv_S4;
    double d2 = 1.0;
    if (omega != 0.0) d2 = omega * omega / (rho + omega * omega);
    const double Lambda = sqrt(4.0 * rho * d2 / ORC_PI);
    double F[MAXM + 1];
    orc_boys(L, rho * d2 * r2, F);
    {
        double dp = 1.0;
        for (int m = 0; m <= L; ++m) { F[m] *= Lambda * dp; dp *= d2; }
    }

    /* V[e][f][m] = [e0|f0]^(m) */
    static _Thread_local double V[NTRI][NTRI][MAXM + 1];
    const double WP_f = S2 * inv_S4;      /* WP = (S2/S4) PQ  */
    const double WQ_f = -S1 * inv_S4;     /* WQ = -(S1/S4) PQ */
    const double r_z = S2 * inv_S4;       /* rho/zeta */
    const double r_e = S1 * inv_S4;       /* rho/eta  */
    const double h1 = 0.5 * inv_S1, h2 = 0.5 * inv_S2, h4 = 0.5 * inv_S4;

    for (int m = 0; m <= L; ++m) V[0][0][m] = F[m];
    /* VRR on a: [e0|00]^(m) */
    for (int n = 1; n <= Le; ++n) {
        for (int p = 0; p < ncart_of(n); ++p) {
            int e[3]; tri_comp(n, p, e);
            int i = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
            int e1[3] = {e[0], e[1], e[2]}; e1[i] -= 1;
            int ie = tri_index(e[0], e[1], e[2]);
            int ie1 = tri_index(e1[0], e1[1], e1[2]);
            int ie2 = -1;
            if (e1[i] > 0) { int e2[3] = {e1[0], e1[1], e1[2]}; e2[i] -= 1; ie2 = tri_index(e2[0], e2[1], e2[2]); }
            for (int m = 0; m <= L - n; ++m) {
                double v = PA[i] * V[ie1][0][m] + WP_f * PQ[i] * V[ie1][0][m + 1];
                if (ie2 >= 0) v += (double)e1[i] * h1 * (V[ie2][0][m] - r_z * V[ie2][0][m + 1]);
                V[ie][0][m] = v;
            }
        }
    }
    /* VRR on c: [e0|f0]^(m) */
    for (int nf = 1; nf <= Lf; ++nf) {
        for (int pf = 0; pf < ncart_of(nf); ++pf) {
            int f[3]; tri_comp(nf, pf, f);
            int i = f[0] > 0 ? 0 : (f[1] > 0 ? 1 : 2);
            int f1[3] = {f[0], f[1], f[2]}; f1[i] -= 1;
            int jf = tri_index(f[0], f[1], f[2]);
            int jf1 = tri_index(f1[0], f1[1], f1[2]);
            int jf2 = -1;
            if (f1[i] > 0) { int f2[3] = {f1[0], f1[1], f1[2]}; f2[i] -= 1; jf2 = tri_index(f2[0], f2[1], f2[2]); }
            for (int ne = 0; ne <= Le; ++ne) {
                for (int pe = 0; pe < ncart_of(ne); ++pe) {
                    int e[3]; tri_comp(ne, pe, e);
                    int ie = tri_index(e[0], e[1], e[2]);
                    int ie1 = -1;
                    if (e[i] > 0) { int e1[3] = {e[0], e[1], e[2]}; e1[i] -= 1; ie1 = tri_index(e1[0], e1[1], e1[2]); }
                    for (int m = 0; m <= L - ne - nf; ++m) {
                        double v = QC[i] * V[ie][jf1][m] + WQ_f * PQ[i] * V[ie][jf1][m + 1];
                        if (jf2 >= 0) v += (double)f1[i] * h2 * (V[ie][jf2][m] - r_e * V[ie][jf2][m + 1]);
                        if (ie1 >= 0) v += (double)e[i] * h4 * V[ie1][jf1][m + 1];
                        V[ie][jf][m] = v;
                    }
                }
            }
        }
    }

    /* HRR on d (ket): W[e][c][d] for e with la <= |e| <= Le */
    const int nc = ncart_of(lc), nd = ncart_of(ld), na = ncart_of(la), nb = ncart_of(lb);
    /* X[lvl][e][c'][d'] with |c'| from lc .. Lf-lvl, |d'| = lvl */
    static _Thread_local double Xk[NTRI][NTRI][NSH]; /* [e][c' triple index][d' pos] current level */
    static _Thread_local double Xn[NTRI][NTRI][NSH];
    static _Thread_local double W[NTRI][NSH][NSH];   /* [e][c][d] */
    for (int ie = 0; ie < NTRI; ++ie) {
        int ne;
        { int cnt = 0; ne = 0; while (cnt + ncart_of(ne) <= ie) { cnt += ncart_of(ne); ++ne; } }
        if (ne < la || ne > Le) continue;
        /* level 0: X(e, f, d'=0) = V(e, f, 0) for |f| in [lc, Lf] */
        for (int nf = lc; nf <= Lf; ++nf)
            for (int pf = 0; pf < ncart_of(nf); ++pf) {
                int f[3]; tri_comp(nf, pf, f);
                Xk[ie][tri_index(f[0], f[1], f[2])][0] = V[ie][tri_index(f[0], f[1], f[2])][0];
            }
        for (int lvl = 1; lvl <= ld; ++lvl) {
            for (int ncp = lc; ncp <= Lf - lvl; ++ncp)
                for (int pc = 0; pc < ncart_of(ncp); ++pc) {
                    int c[3]; tri_comp(ncp, pc, c);
                    int jc = tri_index(c[0], c[1], c[2]);
                    for (int pd = 0; pd < ncart_of(lvl); ++pd) {
                        int d[3]; tri_comp(lvl, pd, d);
                        int i = d[0] > 0 ? 0 : (d[1] > 0 ? 1 : 2);
                        int dm[3] = {d[0], d[1], d[2]}; dm[i] -= 1;
                        int cp[3] = {c[0], c[1], c[2]}; cp[i] += 1;
                        int pdm = tri_pos(dm[0], dm[1], dm[2]);
                        double v = Xk[ie][tri_index(cp[0], cp[1], cp[2])][pdm] + CD[i] * Xk[ie][jc][pdm];
                        Xn[ie][jc][pd] = v;
                    }
                }
            for (int ncp = lc; ncp <= Lf - lvl; ++ncp)
                for (int pc = 0; pc < ncart_of(ncp); ++pc) {
                    int c[3]; tri_comp(ncp, pc, c);
                    int jc = tri_index(c[0], c[1], c[2]);
                    for (int pd = 0; pd < ncart_of(lvl); ++pd) Xk[ie][jc][pd] = Xn[ie][jc][pd];
                }
        }
        for (int pc = 0; pc < nc; ++pc) {
            int c[3]; tri_comp(lc, pc, c);
            for (int pd = 0; pd < nd; ++pd) W[ie][pc][pd] = Xk[ie][tri_index(c[0], c[1], c[2])][pd];
        }
    }
    /* HRR on b (bra), per (c,d) */
    static _Thread_local double Yk[NTRI][NSH], Yn[NTRI][NSH];
    for (int pc = 0; pc < nc; ++pc)
        for (int pd = 0; pd < nd; ++pd) {
            for (int ne = la; ne <= Le; ++ne)
                for (int pe = 0; pe < ncart_of(ne); ++pe) {
                    int e[3]; tri_comp(ne, pe, e);
                    int ie = tri_index(e[0], e[1], e[2]);
                    Yk[ie][0] = W[ie][pc][pd];
                }
            for (int lvl = 1; lvl <= lb; ++lvl) {
                for (int nap = la; nap <= Le - lvl; ++nap)
                    for (int pa = 0; pa < ncart_of(nap); ++pa) {
                        int a[3]; tri_comp(nap, pa, a);
                        int ia = tri_index(a[0], a[1], a[2]);
                        for (int pb = 0; pb < ncart_of(lvl); ++pb) {
                            int b[3]; tri_comp(lvl, pb, b);
                            int i = b[0] > 0 ? 0 : (b[1] > 0 ? 1 : 2);
                            int bm[3] = {b[0], b[1], b[2]}; bm[i] -= 1;
                            int ap[3] = {a[0], a[1], a[2]}; ap[i] += 1;
                            int pbm = tri_pos(bm[0], bm[1], bm[2]);
                            Yn[ia][pb] = Yk[tri_index(ap[0], ap[1], ap[2])][pbm] + AB[i] * Yk[ia][pbm];
                        }
                    }
                for (int nap = la; nap <= Le - lvl; ++nap)
                    for (int pa = 0; pa < ncart_of(nap); ++pa) {
                        int a[3]; tri_comp(nap, pa, a);
                        int ia = tri_index(a[0], a[1], a[2]);
                        for (int pb = 0; pb < ncart_of(lvl); ++pb) Yk[ia][pb] = Yn[ia][pb];
                    }
            }
            for (int pa = 0; pa < na; ++pa) {
                int a[3]; tri_comp(la, pa, a);
                int ia = tri_index(a[0], a[1], a[2]);
                for (int pb = 0; pb < nb; ++pb)
                    out[((pa * nb + pb) * nc + pc) * nd + pd] = Yk[ia][pb];
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Geometric derivative of one primitive quartet                            */
/* ------------------------------------------------------------------------ */

static size_t blk_index(const int* p, const int* n) {
    return (((size_t)p[0] * n[1] + p[1]) * n[2] + p[2]) * n[3] + p[3];
}

/*
 * Derivative of one primitive quartet with respect to the centre of the
 * function in position pos (0: a, 1: b, 2: c, 3: d): Obara-Saika, JCP 84, 3963
 * (1986), eq. (5), the operator the reference applies symbolically in
 * eri.py:85-175 (apply_gradient_a..d) and eriwriter.py:72-86:
 *   d/dR_g ( .. r .. ) = 2 alpha_r ( .. r+1_g .. ) - N_g(r) ( .. r-1_g .. )
 * with alpha_r the exponent of that function and N_g(r) its exponent of
 * direction g.  The raised and lowered classes are evaluated by orc_eri_block
 * with the same exponents and centres (so, like orc_eri_block, without the
 * pair factors S, which do not depend on the angular momentum).
 * out[g * n + t], g = x, y, z, t = ((ia*nb + ib)*nc + ic)*nd + id,
 * n = na nb nc nd.  Each side (la+lb, lc+ld) may be at most 5 after raising.
 */
void orc_eri_deriv_block(int pos, int la, int lb, int lc, int ld,
                         double ai, const double* A, double aj, const double* B,
                         double ak, const double* C, double al, const double* D,
                         double omega, double* out)
{
    const int l[4] = {la, lb, lc, ld};
    const double ex[4] = {ai, aj, ak, al};
    int n[4], lu[4], nu[4], lm[4], nm[4];
    for (int k = 0; k < 4; ++k) { lu[k] = lm[k] = l[k]; n[k] = ncart_of(l[k]); }
    lu[pos] += 1;
    lm[pos] -= 1;
    for (int k = 0; k < 4; ++k) { nu[k] = ncart_of(lu[k]); nm[k] = lm[k] >= 0 ? ncart_of(lm[k]) : 0; }
    static _Thread_local double up[4096], dn[4096];
    orc_eri_block(lu[0], lu[1], lu[2], lu[3], ai, A, aj, B, ak, C, al, D, omega, up);
    if (l[pos] > 0) orc_eri_block(lm[0], lm[1], lm[2], lm[3], ai, A, aj, B, ak, C, al, D, omega, dn);
    const size_t ntot = (size_t)n[0] * n[1] * n[2] * n[3];
    int p[4];
    for (p[0] = 0; p[0] < n[0]; ++p[0])
    for (p[1] = 0; p[1] < n[1]; ++p[1])
    for (p[2] = 0; p[2] < n[2]; ++p[2])
    for (p[3] = 0; p[3] < n[3]; ++p[3]) {
        const size_t t = blk_index(p, n);
        int e[3];
        tri_comp(l[pos], p[pos], e);
        for (int g = 0; g < 3; ++g) {
            int q[4] = {p[0], p[1], p[2], p[3]};
            int eu[3] = {e[0], e[1], e[2]};
            eu[g] += 1;
            q[pos] = tri_pos(eu[0], eu[1], eu[2]);
            double v = 2.0 * ex[pos] * up[blk_index(q, nu)];
            if (e[g] > 0) {
                int em[3] = {e[0], e[1], e[2]};
                em[g] -= 1;
                q[pos] = tri_pos(em[0], em[1], em[2]);
                v -= (double)e[g] * dn[blk_index(q, nm)];
            }
            out[(size_t)g * ntot + t] = v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Helpers over the reference data layout                                   */
/* prim_info: [5][prim_count] SoA (row 0 exponent, rows 2..4 centre)        */
/* combined prim-Cartesian index = prim * ncart + cart                      */
/* ------------------------------------------------------------------------ */

static void prim_of(const double* info, uint32_t count, uint32_t prim, double* a, double* r) {
    *a = info[prim];
    r[0] = info[prim + (size_t)count * 2];
    r[1] = info[prim + (size_t)count * 3];
    r[2] = info[prim + (size_t)count * 4];
}

/*
 * J for one class (la lb | lc ld) over the reference-layout lists.  Lists are
 * sequences of complete primitive-pair blocks (na*nb entries, component order
 * a-major).  bra_blocks selects the bra blocks to evaluate (NULL: all).
 * mat_J is accumulated (+=) like the reference kernels.  *n_eval (if not
 * NULL) is increased by the number of primitive quartets evaluated.
 */
int orc_coulomb_class(int la, int lb, int lc, int ld,
                      const double* const* prim_info, const uint32_t* prim_count,
                      const double* cd_mat_D, const double* ab_mat_Q, const double* cd_mat_Q,
                      const uint32_t* ab_first, const uint32_t* ab_second,
                      const double* ab_pair_data, uint32_t ab_count,
                      const uint32_t* cd_first, const uint32_t* cd_second,
                      const double* cd_pair_data, uint32_t cd_count,
                      double thr, int mode,
                      const uint32_t* bra_blocks, uint32_t n_bra_blocks,
                      double* mat_J, unsigned long long* n_eval)
{
    unsigned long long nev = 0;
    const int na = ncart_of(la), nb = ncart_of(lb), nc = ncart_of(lc), nd = ncart_of(ld);
    const int nab = na * nb, ncd = nc * nd;
    if (ab_count % nab || cd_count % ncd) return -1;
    const uint32_t nbra = ab_count / nab, nket = cd_count / ncd;
    const uint32_t nsel = bra_blocks ? n_bra_blocks : nbra;
    const int same_cd = (lc == ld);

    /* per ket block: beta = max |Q_kl D_kl| */
    double* beta = (double*)malloc(sizeof(double) * (nket ? nket : 1));
    for (uint32_t kb = 0; kb < nket; ++kb) {
        double m = 0.0;
        for (int t = 0; t < ncd; ++t) {
            double v = fabs(cd_mat_Q[(size_t)kb * ncd + t] * cd_mat_D[(size_t)kb * ncd + t]);
            if (v > m) m = v;
        }
        beta[kb] = m;
    }

    #pragma omp parallel reduction(+:nev)
    {
        double* blk = (double*)malloc(sizeof(double) * 1296);
        #pragma omp for schedule(dynamic, 4)
        for (uint32_t s = 0; s < nsel; ++s) {
            const uint32_t bb = bra_blocks ? bra_blocks[s] : s;
            const size_t e0 = (size_t)bb * nab;
            const uint32_t I = ab_first[e0] / na, J = ab_second[e0] / nb;
            double ai, aj, A[3], B[3];
            prim_of(prim_info[la], prim_count[la], I, &ai, A);
            prim_of(prim_info[lb], prim_count[lb], J, &aj, B);
            double qmax = 0.0;
            for (int t = 0; t < nab; ++t) { double q = fabs(ab_mat_Q[e0 + t]); if (q > qmax) qmax = q; }
            /* extended-precision sums: the oracle's own rounding stays far below
               the parity bar even for sums of millions of quartets */
            long double acc[36];
            for (int t = 0; t < nab; ++t) acc[t] = 0.0L;
            for (uint32_t kb = 0; kb < nket; ++kb) {
                if (mode == 1) {
                    if (!(qmax * beta[kb] > thr)) continue;
                } else {
                    if (qmax * beta[kb] * (1.0 + 1e-13) <= thr) continue;
                }
                const size_t f0 = (size_t)kb * ncd;
                const uint32_t K = cd_first[f0] / nc, Lp = cd_second[f0] / nd;
                double ak, al, C[3], D[3];
                prim_of(prim_info[lc], prim_count[lc], K, &ak, C);
                prim_of(prim_info[ld], prim_count[ld], Lp, &al, D);
                orc_eri_block(la, lb, lc, ld, ai, A, aj, B, ak, C, al, D, 0.0, blk);
                ++nev;
                for (int tab = 0; tab < nab; ++tab) {
                    const double Qij = ab_mat_Q[e0 + tab];
                    const double Sij = ab_pair_data[e0 + tab];
                    long double s_acc = 0.0L;
                    for (int tcd = 0; tcd < ncd; ++tcd) {
                        const size_t kl = f0 + tcd;
                        const double Dkl = cd_mat_D[kl];
                        if (mode == 0 && !(fabs(Qij * cd_mat_Q[kl] * Dkl) > thr)) continue;
                        double w;
                        if (!same_cd) w = 2.0;
                        else w = (cd_first[kl] != cd_second[kl]) ? 2.0 : 1.0;
                        s_acc += (long double)(blk[tab * ncd + tcd] * Sij * cd_pair_data[kl] * Dkl * w);
                    }
                    acc[tab] += s_acc;
                }
            }
            for (int t = 0; t < nab; ++t) mat_J[e0 + t] = (double)((long double)mat_J[e0 + t] + acc[t]);
        }
        free(blk);
    }
    free(beta);
    if (n_eval) *n_eval += nev;
    return 0;
}

/*
 * K for one class (la lb | lc ld) over the reference-layout CSR lists.  The
 * ik list holds complete (I,K) blocks (na*nc entries, a-major); every CSR
 * segment of a combined index i = I*na + ca lists the partner primitive blocks
 * of I (nb entries each) in the same order for every ca.  max_j > 0 limits
 * each block to its first max_j partner blocks of I (CPU-timing samples only);
 * *n_eval (if not NULL) is increased by the number of quartets evaluated.
 */
int orc_exchange_class(int la, int lb, int lc, int ld,
                       const uint32_t* ik_i, const uint32_t* ik_k, uint32_t ik_count,
                       const double* const* prim_info, const uint32_t* const* prim_aoinds,
                       const uint32_t* prim_count,
                       double max_D_bd, const double* D_AO, uint32_t naos,
                       const double* Q_ab, const double* Q_cd,
                       const uint32_t* Dinds_ab, const uint32_t* Dinds_cd,
                       const uint32_t* displs_ab, const uint32_t* displs_cd,
                       const uint32_t* counts_ab, const uint32_t* counts_cd,
                       const double* pd_ab, const double* pd_cd,
                       double omega, double thr, int mode,
                       const uint32_t* ik_blocks, uint32_t n_ik_blocks,
                       double* mat_K, uint32_t max_j, unsigned long long* n_eval)
{
    unsigned long long nev = 0;
    const int na = ncart_of(la), nb = ncart_of(lb), nc = ncart_of(lc), nd = ncart_of(ld);
    const int nac = na * nc;
    if (ik_count % nac) return -1;
    const uint32_t nblk = ik_count / nac;
    const uint32_t nsel = ik_blocks ? n_ik_blocks : nblk;

    #pragma omp parallel reduction(+:nev)
    {
        double* blk = (double*)malloc(sizeof(double) * 1296);
        #pragma omp for schedule(dynamic, 2)
        for (uint32_t s = 0; s < nsel; ++s) {
            const uint32_t ib = ik_blocks ? ik_blocks[s] : s;
            const size_t e0 = (size_t)ib * nac;
            const uint32_t i0 = ik_i[e0], k0 = ik_k[e0];
            const uint32_t I = i0 / na, K = k0 / nc;
            double ai, ak, A[3], C[3];
            prim_of(prim_info[la], prim_count[la], I, &ai, A);
            prim_of(prim_info[lc], prim_count[lc], K, &ak, C);
            /* segments of I's components and K's components */
            const uint32_t cnt_i = counts_ab[I * na], cnt_k = counts_cd[K * nc];
            uint32_t nJ = cnt_i / nb;
            const uint32_t nL = cnt_k / nd;
            if (max_j && nJ > max_j) nJ = max_j;
            long double acc[36];
            for (int t = 0; t < nac; ++t) acc[t] = 0.0L;
            for (uint32_t jb = 0; jb < nJ; ++jb) {
                const uint32_t J = Dinds_ab[displs_ab[I * na] + jb * nb] / nb;
                double aj, B[3];
                prim_of(prim_info[lb], prim_count[lb], J, &aj, B);
                double qij_max = 0.0;
                for (int ca = 0; ca < na; ++ca)
                    for (int cb = 0; cb < nb; ++cb) {
                        double q = fabs(Q_ab[displs_ab[I * na + ca] + jb * nb + cb]);
                        if (q > qij_max) qij_max = q;
                    }
                for (uint32_t lbk = 0; lbk < nL; ++lbk) {
                    const uint32_t Lp = Dinds_cd[displs_cd[K * nc] + lbk * nd] / nd;
                    double qkl_max = 0.0;
                    for (int cc = 0; cc < nc; ++cc)
                        for (int cd = 0; cd < nd; ++cd) {
                            double q = fabs(Q_cd[displs_cd[K * nc + cc] + lbk * nd + cd]);
                            if (q > qkl_max) qkl_max = q;
                        }
                    if (mode == 1) {
                        if (!(qij_max * qkl_max * max_D_bd > thr)) continue;
                    } else {
                        if (qij_max * qkl_max * max_D_bd * (1.0 + 1e-13) <= thr) continue;
                    }
                    double al, Dc[3];
                    prim_of(prim_info[ld], prim_count[ld], Lp, &al, Dc);
                    orc_eri_block(la, lb, lc, ld, ai, A, aj, B, ak, C, al, Dc, omega, blk);
                    ++nev;
                    for (int ca = 0; ca < na; ++ca)
                        for (int cc = 0; cc < nc; ++cc) {
                            long double s_acc = 0.0L;
                            for (int cb = 0; cb < nb; ++cb) {
                                const size_t pj = displs_ab[I * na + ca] + jb * nb + cb;
                                const uint32_t jprim = Dinds_ab[pj];
                                const uint32_t jcg = prim_aoinds[lb][(jprim / nb) + prim_count[lb] * (jprim % nb)];
                                const double Qij = Q_ab[pj], Sij = pd_ab[pj];
                                for (int cd = 0; cd < nd; ++cd) {
                                    const size_t pl = displs_cd[K * nc + cc] + lbk * nd + cd;
                                    const uint32_t lprim = Dinds_cd[pl];
                                    const uint32_t lcg = prim_aoinds[ld][(lprim / nd) + prim_count[ld] * (lprim % nd)];
                                    if (mode == 0 && !(fabs(Qij * Q_cd[pl] * max_D_bd) > thr)) continue;
                                    s_acc += (long double)(blk[((ca * nb + cb) * nc + cc) * nd + cd] * Sij * pd_cd[pl]
                                                           * D_AO[(size_t)jcg * naos + lcg]);
                                }
                            }
                            acc[ca * nc + cc] += s_acc;
                        }
                }
            }
            for (int t = 0; t < nac; ++t) mat_K[e0 + t] = (double)((long double)mat_K[e0 + t] + acc[t]);
        }
        free(blk);
    }
    if (n_eval) *n_eval += nev;
    return 0;
}

/*
 * Schwarz factors Q_ab = sqrt(|(ab|ab)|) for every component pair of a list of
 * primitive-pair blocks; S holds the pair data (contraction x normalisation x
 * Gaussian overlap factor) per component pair, layout (nblocks, na, nb).
 */
void orc_schwarz(int la, int lb, const double* info_a, uint32_t count_a,
                 const double* info_b, uint32_t count_b,
                 const uint32_t* I, const uint32_t* J, uint32_t nblocks,
                 const double* S, double* Q)
{
    const int na = ncart_of(la), nb = ncart_of(lb);
    #pragma omp parallel
    {
        double* blk = (double*)malloc(sizeof(double) * 1296);
        #pragma omp for schedule(dynamic, 64)
        for (uint32_t t = 0; t < nblocks; ++t) {
            double ai, aj, A[3], B[3];
            prim_of(info_a, count_a, I[t], &ai, A);
            prim_of(info_b, count_b, J[t], &aj, B);
            orc_eri_block(la, lb, la, lb, ai, A, aj, B, ai, A, aj, B, 0.0, blk);
            const int nab = na * nb;
            for (int x = 0; x < nab; ++x) {
                double s = S[(size_t)t * nab + x];
                Q[(size_t)t * nab + x] = sqrt(fabs(blk[x * nab + x] * s * s));
            }
        }
        free(blk);
    }
}

/*
 * Full Cartesian AO two-electron tensor (mu nu | la si) for a small basis.
 * Shells: l, nprim, centre, exponents, primitive factors (cnorm c N(x^l)),
 * Cartesian AO offset.  comp_ratio[l][c] scales component c.  Output naos^4.
 */
void orc_ao_eri(int nshell, const int* shl_l, const int* shl_nprim, const double* shl_center,
                const int* shl_prim_off, const double* prim_exp, const double* prim_fac,
                const int* shl_ao_off, const double* comp_ratio, int naos, double* out)
{
    #pragma omp parallel
    {
        double* blk = (double*)malloc(sizeof(double) * 1296);
        #pragma omp for schedule(dynamic, 1) collapse(2)
        for (int s1 = 0; s1 < nshell; ++s1)
            for (int s2 = 0; s2 < nshell; ++s2)
                for (int s3 = 0; s3 < nshell; ++s3)
                    for (int s4 = 0; s4 < nshell; ++s4) {
                        int l1 = shl_l[s1], l2 = shl_l[s2], l3 = shl_l[s3], l4 = shl_l[s4];
                        int n1 = ncart_of(l1), n2 = ncart_of(l2), n3 = ncart_of(l3), n4 = ncart_of(l4);
                        for (int c1 = 0; c1 < n1; ++c1) for (int c2 = 0; c2 < n2; ++c2)
                        for (int c3 = 0; c3 < n3; ++c3) for (int c4 = 0; c4 < n4; ++c4)
                            out[(((size_t)(shl_ao_off[s1] + c1) * naos + shl_ao_off[s2] + c2) * naos
                                 + shl_ao_off[s3] + c3) * naos + shl_ao_off[s4] + c4] = 0.0;
                        for (int p1 = 0; p1 < shl_nprim[s1]; ++p1)
                        for (int p2 = 0; p2 < shl_nprim[s2]; ++p2)
                        for (int p3 = 0; p3 < shl_nprim[s3]; ++p3)
                        for (int p4 = 0; p4 < shl_nprim[s4]; ++p4) {
                            int q1 = shl_prim_off[s1] + p1, q2 = shl_prim_off[s2] + p2;
                            int q3 = shl_prim_off[s3] + p3, q4 = shl_prim_off[s4] + p4;
                            double a1 = prim_exp[q1], a2 = prim_exp[q2], a3 = prim_exp[q3], a4 = prim_exp[q4];
                            const double *A = shl_center + 3 * s1, *B = shl_center + 3 * s2;
                            const double *C = shl_center + 3 * s3, *D = shl_center + 3 * s4;
                            /* S factors: c_i c_j N_i N_j (pi/zeta)^1.5 exp(-a_i a_j |AB|^2 / zeta) */
                            double z = a1 + a2, e = a3 + a4;
                            double ab2 = 0, cd2 = 0;
                            for (int x = 0; x < 3; ++x) { ab2 += (A[x]-B[x])*(A[x]-B[x]); cd2 += (C[x]-D[x])*(C[x]-D[x]); }
                            double Sab = prim_fac[q1] * prim_fac[q2] * pow(ORC_PI / z, 1.5) * exp(-a1 * a2 / z * ab2);
                            double Scd = prim_fac[q3] * prim_fac[q4] * pow(ORC_PI / e, 1.5) * exp(-a3 * a4 / e * cd2);
                            orc_eri_block(l1, l2, l3, l4, a1, A, a2, B, a3, C, a4, D, 0.0, blk);
                            int t = 0;
                            for (int c1 = 0; c1 < n1; ++c1) for (int c2 = 0; c2 < n2; ++c2)
                            for (int c3 = 0; c3 < n3; ++c3) for (int c4 = 0; c4 < n4; ++c4, ++t) {
                                double f = Sab * Scd * comp_ratio[l1 * 6 + c1] * comp_ratio[l2 * 6 + c2]
                                         * comp_ratio[l3 * 6 + c3] * comp_ratio[l4 * 6 + c4];
                                out[(((size_t)(shl_ao_off[s1] + c1) * naos + shl_ao_off[s2] + c2) * naos
                                     + shl_ao_off[s3] + c3) * naos + shl_ao_off[s4] + c4] += f * blk[t];
                            }
                        }
                    }
        free(blk);
    }
}

/* ------------------------------------------------------------------------ */
/* Energy-gradient contributions of the J and K classes                      */
/* ------------------------------------------------------------------------ */

/*
 * Coulomb-energy gradient of one J class over the lists of orc_coulomb_class,
 * with its screening (a component quartet counts iff |Q_ij Q_kl D_kl| > thr,
 * gen_coulomb.py:209-212).  For E_J = 1/2 sum_ab D_ab J_ab with D symmetric,
 * the bra-ket and a<->b symmetries of the ordered sum give
 *   dE_J/dX = sum_{bra entries ab} w_ab D_ab sum_{ket entries cd} w_cd D_cd
 *             [ dA (ab|cd) if a sits on X  +  dB (ab|cd) if b sits on X ]
 * with w the J weight (2, or 1 for an entry with first == second of a
 * same-type list; gen_coulomb.py:276-281) and dA, dB the derivatives of
 * orc_eri_deriv_block times the pair factors.  ab_mat_D holds the bra
 * entries' D in the ket role's convention (0 below the diagonal of a diagonal
 * block).  grad[6 bb + 0..2] += the A-centre part and grad[6 bb + 3..5] +=
 * the B-centre part of every selected bra block bb.
 */
int orc_coulomb_grad_class(int la, int lb, int lc, int ld,
                           const double* const* prim_info, const uint32_t* prim_count,
                           const double* ab_mat_D, const double* cd_mat_D,
                           const double* ab_mat_Q, const double* cd_mat_Q,
                           const uint32_t* ab_first, const uint32_t* ab_second,
                           const double* ab_pair_data, uint32_t ab_count,
                           const uint32_t* cd_first, const uint32_t* cd_second,
                           const double* cd_pair_data, uint32_t cd_count,
                           double thr, const uint32_t* bra_blocks, uint32_t n_bra_blocks,
                           double* grad, unsigned long long* n_eval)
{
    unsigned long long nev = 0;
    const int na = ncart_of(la), nb = ncart_of(lb), nc = ncart_of(lc), nd = ncart_of(ld);
    const int nab = na * nb, ncd = nc * nd, n = nab * ncd;
    if (ab_count % nab || cd_count % ncd) return -1;
    const uint32_t nbra = ab_count / nab, nket = cd_count / ncd;
    const uint32_t nsel = bra_blocks ? n_bra_blocks : nbra;
    const int same_ab = (la == lb), same_cd = (lc == ld);
    double* beta = (double*)malloc(sizeof(double) * (nket ? nket : 1));
    for (uint32_t kb = 0; kb < nket; ++kb) {
        double m = 0.0;
        for (int t = 0; t < ncd; ++t) {
            double v = fabs(cd_mat_Q[(size_t)kb * ncd + t] * cd_mat_D[(size_t)kb * ncd + t]);
            if (v > m) m = v;
        }
        beta[kb] = m;
    }
    #pragma omp parallel reduction(+:nev)
    {
        double* dA = (double*)malloc(sizeof(double) * 3 * 1296);
        double* dB = (double*)malloc(sizeof(double) * 3 * 1296);
        #pragma omp for schedule(dynamic, 4)
        for (uint32_t s = 0; s < nsel; ++s) {
            const uint32_t bb = bra_blocks ? bra_blocks[s] : s;
            const size_t e0 = (size_t)bb * nab;
            const uint32_t I = ab_first[e0] / na, J = ab_second[e0] / nb;
            double ai, aj, A[3], B[3];
            prim_of(prim_info[la], prim_count[la], I, &ai, A);
            prim_of(prim_info[lb], prim_count[lb], J, &aj, B);
            double qmax = 0.0;
            for (int t = 0; t < nab; ++t) { double q = fabs(ab_mat_Q[e0 + t]); if (q > qmax) qmax = q; }
            long double acc[6] = {0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L};
            for (uint32_t kb = 0; kb < nket; ++kb) {
                if (qmax * beta[kb] * (1.0 + 1e-13) <= thr) continue;
                const size_t f0 = (size_t)kb * ncd;
                const uint32_t K = cd_first[f0] / nc, Lp = cd_second[f0] / nd;
                double ak, al, C[3], D[3];
                prim_of(prim_info[lc], prim_count[lc], K, &ak, C);
                prim_of(prim_info[ld], prim_count[ld], Lp, &al, D);
                orc_eri_deriv_block(0, la, lb, lc, ld, ai, A, aj, B, ak, C, al, D, 0.0, dA);
                orc_eri_deriv_block(1, la, lb, lc, ld, ai, A, aj, B, ak, C, al, D, 0.0, dB);
                ++nev;
                for (int tab = 0; tab < nab; ++tab) {
                    const double Qij = ab_mat_Q[e0 + tab];
                    const double wab = (same_ab && ab_first[e0 + tab] == ab_second[e0 + tab]) ? 1.0 : 2.0;
                    const double fab = ab_pair_data[e0 + tab] * ab_mat_D[e0 + tab] * wab;
                    for (int tcd = 0; tcd < ncd; ++tcd) {
                        const size_t kl = f0 + tcd;
                        const double Dkl = cd_mat_D[kl];
                        if (!(fabs(Qij * cd_mat_Q[kl] * Dkl) > thr)) continue;
                        const double wcd = (same_cd && cd_first[kl] == cd_second[kl]) ? 1.0 : 2.0;
                        const double f = fab * cd_pair_data[kl] * Dkl * wcd;
                        const int t = tab * ncd + tcd;
                        for (int g = 0; g < 3; ++g) {
                            acc[g] += (long double)(f * dA[g * n + t]);
                            acc[3 + g] += (long double)(f * dB[g * n + t]);
                        }
                    }
                }
            }
            for (int g = 0; g < 6; ++g) grad[6 * (size_t)bb + g] = (double)((long double)grad[6 * (size_t)bb + g] + acc[g]);
        }
        free(dA);
        free(dB);
    }
    free(beta);
    if (n_eval) *n_eval += nev;
    return 0;
}

/*
 * Exchange-energy gradient of one K class over the lists of
 * orc_exchange_class, with its screening (a component quartet counts iff
 * |Q_ij Q_kl maxD_bd| > thr, gen_exchange.py:229-232).  For
 * E_K = 1/2 sum_ac D_ac K_ac with D symmetric, the symmetries of the ordered
 * sum give
 *   dE_K/dX = sum_{(I,K) blocks} m_IK sum_{a in I, c in K} D_ac sum_{b,d} D_bd
 *             [ dA (ab|cd) if a sits on X  +  dC (ab|cd) if c sits on X ]
 * with m_IK = 1 for a block I = K of a same-type list and 2 otherwise (the
 * list holds one of (I,K), (K,I)).  grad[6 ib + 0..2] += the A-centre part,
 * grad[6 ib + 3..5] += the C-centre part of every selected (i,k) block ib.
 */
int orc_exchange_grad_class(int la, int lb, int lc, int ld,
                            const uint32_t* ik_i, const uint32_t* ik_k, uint32_t ik_count,
                            const double* const* prim_info, const uint32_t* const* prim_aoinds,
                            const uint32_t* prim_count,
                            double max_D_bd, const double* D_AO, uint32_t naos,
                            const double* Q_ab, const double* Q_cd,
                            const uint32_t* Dinds_ab, const uint32_t* Dinds_cd,
                            const uint32_t* displs_ab, const uint32_t* displs_cd,
                            const uint32_t* counts_ab, const uint32_t* counts_cd,
                            const double* pd_ab, const double* pd_cd,
                            double omega, double thr,
                            const uint32_t* ik_blocks, uint32_t n_ik_blocks,
                            double* grad, unsigned long long* n_eval)
{
    unsigned long long nev = 0;
    const int na = ncart_of(la), nb = ncart_of(lb), nc = ncart_of(lc), nd = ncart_of(ld);
    const int nac = na * nc, n = na * nb * nc * nd;
    if (ik_count % nac) return -1;
    const uint32_t nblk = ik_count / nac;
    const uint32_t nsel = ik_blocks ? n_ik_blocks : nblk;
    #pragma omp parallel reduction(+:nev)
    {
        double* dA = (double*)malloc(sizeof(double) * 3 * 1296);
        double* dC = (double*)malloc(sizeof(double) * 3 * 1296);
        #pragma omp for schedule(dynamic, 2)
        for (uint32_t s = 0; s < nsel; ++s) {
            const uint32_t ib = ik_blocks ? ik_blocks[s] : s;
            const size_t e0 = (size_t)ib * nac;
            const uint32_t I = ik_i[e0] / na, K = ik_k[e0] / nc;
            double ai, ak, A[3], C[3];
            prim_of(prim_info[la], prim_count[la], I, &ai, A);
            prim_of(prim_info[lc], prim_count[lc], K, &ak, C);
            const double mIK = (la == lc && I == K) ? 1.0 : 2.0;
            double Dac[36];
            for (int ca = 0; ca < na; ++ca)
                for (int cc = 0; cc < nc; ++cc) {
                    const uint32_t ao_a = prim_aoinds[la][I + prim_count[la] * ca];
                    const uint32_t ao_c = prim_aoinds[lc][K + prim_count[lc] * cc];
                    Dac[ca * nc + cc] = D_AO[(size_t)ao_a * naos + ao_c];
                }
            const uint32_t nJ = counts_ab[I * na] / nb, nL = counts_cd[K * nc] / nd;
            long double acc[6] = {0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L};
            for (uint32_t jb = 0; jb < nJ; ++jb) {
                const uint32_t J = Dinds_ab[displs_ab[I * na] + jb * nb] / nb;
                double aj, B[3];
                prim_of(prim_info[lb], prim_count[lb], J, &aj, B);
                double qij_max = 0.0;
                for (int ca = 0; ca < na; ++ca)
                    for (int cb = 0; cb < nb; ++cb) {
                        double q = fabs(Q_ab[displs_ab[I * na + ca] + jb * nb + cb]);
                        if (q > qij_max) qij_max = q;
                    }
                for (uint32_t lbk = 0; lbk < nL; ++lbk) {
                    const uint32_t Lp = Dinds_cd[displs_cd[K * nc] + lbk * nd] / nd;
                    double qkl_max = 0.0;
                    for (int cc = 0; cc < nc; ++cc)
                        for (int cd = 0; cd < nd; ++cd) {
                            double q = fabs(Q_cd[displs_cd[K * nc + cc] + lbk * nd + cd]);
                            if (q > qkl_max) qkl_max = q;
                        }
                    if (qij_max * qkl_max * max_D_bd * (1.0 + 1e-13) <= thr) continue;
                    double al, Dc[3];
                    prim_of(prim_info[ld], prim_count[ld], Lp, &al, Dc);
                    orc_eri_deriv_block(0, la, lb, lc, ld, ai, A, aj, B, ak, C, al, Dc, omega, dA);
                    orc_eri_deriv_block(2, la, lb, lc, ld, ai, A, aj, B, ak, C, al, Dc, omega, dC);
                    ++nev;
                    for (int ca = 0; ca < na; ++ca)
                        for (int cb = 0; cb < nb; ++cb) {
                            const size_t pj = displs_ab[I * na + ca] + jb * nb + cb;
                            const uint32_t jprim = Dinds_ab[pj];
                            const uint32_t jcg = prim_aoinds[lb][(jprim / nb) + prim_count[lb] * (jprim % nb)];
                            const double Qij = Q_ab[pj], Sij = pd_ab[pj];
                            for (int cc = 0; cc < nc; ++cc)
                                for (int cd = 0; cd < nd; ++cd) {
                                    const size_t pl = displs_cd[K * nc + cc] + lbk * nd + cd;
                                    if (!(fabs(Qij * Q_cd[pl] * max_D_bd) > thr)) continue;
                                    const uint32_t lprim = Dinds_cd[pl];
                                    const uint32_t lcg = prim_aoinds[ld][(lprim / nd) + prim_count[ld] * (lprim % nd)];
                                    const double f = mIK * Dac[ca * nc + cc] * Sij * pd_cd[pl]
                                                     * D_AO[(size_t)jcg * naos + lcg];
                                    const int t = ((ca * nb + cb) * nc + cc) * nd + cd;
                                    for (int g = 0; g < 3; ++g) {
                                        acc[g] += (long double)(f * dA[g * n + t]);
                                        acc[3 + g] += (long double)(f * dC[g * n + t]);
                                    }
                                }
                        }
                }
            }
            for (int g = 0; g < 6; ++g) grad[6 * (size_t)ib + g] = (double)((long double)grad[6 * (size_t)ib + g] + acc[g]);
        }
        free(dA);
        free(dC);
    }
    if (n_eval) *n_eval += nev;
    return 0;
}
