"""Concrete-component Obara-Saika / Head-Gordon-Pople expression DAG.

The reference expands every Cartesian component of a class into one flat
polynomial over symbolic directions with run-time Kronecker deltas
(eriwriter.py:35-160, intswriter.py:30-408), and evaluates it once per component
lane.  This generator instead works on concrete Cartesian exponent triples known
at generation time, builds the recursion DAG of one *primitive quartet* with
memoisation (every intermediate computed once and shared by all components),
prunes the terms whose integer factors vanish, and emits straight-line code.

Recursions (same forms as the reference; PQ = Q - P):
  VRR-a  eri.py:297-355
    [e+1_i 0|f 0]^m = PA_i [e|f]^m + WP_i [e|f]^{m+1}
                      + e_i/(2 zeta) ([e-1_i|f]^m - (rho/zeta) [e-1_i|f]^{m+1})
                      + f_i/(2 S4) [e|f-1_i]^{m+1}
    with WP = (eta/S4) PQ.  It is only applied with f = 0 here (VRR-c runs
    after it bottom-up), so the last term never occurs.
  VRR-c  eri.py:237-295
    [e|f+1_i]^m = QC_i [e|f]^m + WQ_i [e|f]^{m+1}
                  + f_i/(2 eta) ([e|f-1_i]^m - (rho/eta) [e|f-1_i]^{m+1})
                  + e_i/(2 S4) [e-1_i|f]^{m+1}
    with WQ = -(zeta/S4) PQ.
  HRR-b  eri.py:207-235   (a, b+1_i| = (a+1_i, b| + AB_i (a, b|,  AB = A - B
  HRR-d  eri.py:177-205   |c, d+1_i) = |c+1_i, d) + CD_i |c, d),  CD = C - D
  base   [00|00]^m = Lambda F_m(T)                       gen_coulomb.py:255-270
"""

from __future__ import annotations

AX = "xyz"


def tri(l):
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def tri_upto(l):
    out = []
    for n in range(l + 1):
        out += tri(n)
    return out


def tname(t):
    return "%d%d%d" % t


def add(t, i, k=1):
    t = list(t)
    t[i] += k
    return tuple(t)


def choose_dir(t, rule):
    nz = [i for i in range(3) if t[i] > 0]
    if rule == "first":
        return nz[0]
    if rule == "last":
        return nz[-1]
    # "max": largest exponent, ties to the first
    best = nz[0]
    for i in nz:
        if t[i] > t[best]:
            best = i
    return best


class QuartetDAG:
    """Emits code for one primitive quartet of class (la lb | lc ld).

    Inputs expected in scope of the emitted code (type ``ty``):
      F0..FL (Lambda-scaled Boys values), PA_x/y/z, WP_x/y/z, QC_x/y/z,
      WQ_x/y/z, AB_x/y/z, CD_x/y/z, h1 = 1/(2 zeta), h2 = 1/(2 eta),
      h4 = 1/(2 S4), rz = rho/zeta, re = rho/eta.
    Multiples k*h1 etc. are declared on demand as h1_k.
    """

    def __init__(self, la, lb, lc, ld, ty="double", rule_vrr="max", rule_hrr="max", prefix="q"):
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.L = la + lb + lc + ld
        self.ty = ty
        self.rule_vrr = rule_vrr
        self.rule_hrr = rule_hrr
        self.prefix = prefix
        self.lines = []
        self.memo = {}
        self.mults = set()
        self.nflop = 0

    # ---------------------------------------------------------------- helpers
    def _new(self, key, expr, flops):
        name = "%s%d" % (self.prefix, len(self.memo))
        self.memo[key] = name
        self.lines.append("const %s %s = %s;" % (self.ty, name, expr))
        self.nflop += flops
        return name

    def _mult(self, base, k):
        if k == 1:
            return base
        self.mults.add((base, k))
        return "%sm_%s_%d" % (self.prefix, base, k)

    # ---------------------------------------------------------------- VRR
    def V(self, e, f, m):
        key = ("V", e, f, m)
        if key in self.memo:
            return self.memo[key]
        if sum(e) == 0 and sum(f) == 0:
            self.memo[key] = "F%d" % m
            return self.memo[key]
        if sum(f) > 0:
            i = choose_dir(f, self.rule_vrr)
            f1 = add(f, i, -1)
            a = self.V(e, f1, m)
            b = self.V(e, f1, m + 1)
            terms = ["QC_%s * %s" % (AX[i], a), "WQ_%s * %s" % (AX[i], b)]
            fl = 3
            if f1[i] > 0:
                f2 = add(f1, i, -1)
                c = self.V(e, f2, m)
                d = self.V(e, f2, m + 1)
                terms.append("%s * (%s - re * %s)" % (self._mult("h2", f1[i]), c, d))
                fl += 4
            if e[i] > 0:
                g = self.V(add(e, i, -1), f1, m + 1)
                terms.append("%s * %s" % (self._mult("h4", e[i]), g))
                fl += 2
            return self._new(key, " + ".join(terms), fl)
        i = choose_dir(e, self.rule_vrr)
        e1 = add(e, i, -1)
        a = self.V(e1, f, m)
        b = self.V(e1, f, m + 1)
        terms = ["PA_%s * %s" % (AX[i], a), "WP_%s * %s" % (AX[i], b)]
        fl = 3
        if e1[i] > 0:
            e2 = add(e1, i, -1)
            c = self.V(e2, f, m)
            d = self.V(e2, f, m + 1)
            terms.append("%s * (%s - rz * %s)" % (self._mult("h1", e1[i]), c, d))
            fl += 4
        return self._new(key, " + ".join(terms), fl)

    # ---------------------------------------------------------------- HRR
    def W(self, e, c, d):
        """(e 0 | c d) from [e0|f0] by HRR on d."""
        key = ("W", e, c, d)
        if key in self.memo:
            return self.memo[key]
        if sum(d) == 0:
            v = self.V(e, c, 0)
            self.memo[key] = v
            return v
        i = choose_dir(d, self.rule_hrr)
        d1 = add(d, i, -1)
        a = self.W(e, add(c, i), d1)
        b = self.W(e, c, d1)
        return self._new(key, "%s + CD_%s * %s" % (a, AX[i], b), 2)

    def X(self, a, b, c, d):
        """(a b | c d) from (e0|cd) by HRR on b."""
        key = ("X", a, b, c, d)
        if key in self.memo:
            return self.memo[key]
        if sum(b) == 0:
            v = self.W(a, c, d)
            self.memo[key] = v
            return v
        i = choose_dir(b, self.rule_hrr)
        b1 = add(b, i, -1)
        x = self.X(add(a, i), b1, c, d)
        y = self.X(a, b1, c, d)
        return self._new(key, "%s + AB_%s * %s" % (x, AX[i], y), 2)

    def mult_decls(self):
        out = []
        for base, k in sorted(self.mults):
            out.append("const %s %sm_%s_%d = %s(%d) * %s;" % (self.ty, self.prefix, base, k, self.ty, k, base))
        return out


def count_nodes(la, lb, lc, ld, rule_vrr="max", rule_hrr="max"):
    g = QuartetDAG(la, lb, lc, ld, rule_vrr=rule_vrr, rule_hrr=rule_hrr)
    for c in tri(lc):
        for d in tri(ld):
            for a in tri(la):
                for b in tri(lb):
                    g.X(a, b, c, d)
    return len(g.lines), g.nflop
