#!/usr/bin/env python3
"""Generate the RELACS_TUV device functions for the RODAS3 kernels.

Reads brams_amd/mechanism/relacs_tuv.json and writes
brams_amd/csrc/generated/relacs_tuv_gen.h with straight-line HIP device code:

  mech_kinetic  rk[128] from T, p, rv, J[17]       (chem_spack_kinetic.f90)
  mech_fun      f[47] = S w(rk, y)                  (rates.f90 + fexchem.f90)
  mech_mask     value-based sparsity mask of one cell's matrix
                1/(gamma h) I - J                   (rodas3:449-482)
  mech_factor   masked matrix assembled from rk, y0 and factored in place
                (Crout, reciprocal pivots, as spFactor.c:336-401)
  mech_solve    forward / backward substitution     (spSolve.c:183-218)

Sums keep the reference's term order (reactions ascending).  The LU uses one
pivot order fixed here: Markowitz on the structural pattern with diagonal
pivots, ties to the lowest species index (the oracle uses the same rule).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
JSON = os.path.join(PKG, "mechanism", "relacs_tuv.json")
OUT = os.path.join(PKG, "csrc", "generated", "relacs_tuv_gen.h")


def lit(x):
    x = float(x)
    s = repr(x)
    if "e" not in s and "." not in s and "inf" not in s and "nan" not in s:
        s += ".0"
    return s


def load():
    m = json.load(open(JSON))
    ns, nr = m["nspecies"], m["nr"]
    react = [[x - 1 for x in r["reactants"]] for r in m["reactions"]]
    prod = {sp - 1: [(c, r - 1) for c, r in terms] for sp, terms in m["production"]}
    return m, ns, nr, react, prod


def jac_terms(ns, react, prod):
    present = {}
    for sp, terms in prod.items():
        for c, r in terms:
            present[(sp, r)] = c
    jt = {}
    for i in range(ns):
        for j in range(ns):
            terms = []
            for r, rs in enumerate(react):
                if (i, r) not in present:
                    continue
                for slot, s in enumerate(rs):
                    if s == j:
                        other = rs[1 - slot] if len(rs) == 2 else -1
                        terms.append((present[(i, r)], r, other))
            if terms:
                jt[(i, j)] = terms
    return jt


def markowitz(ns, struct):
    """Diagonal Markowitz order; returns perm (internal -> species) and the
    filled pattern in species indices."""
    P = [row[:] for row in struct]
    active = [True] * ns
    perm = []
    for _ in range(ns):
        best, bestprod = -1, None
        for c in range(ns):
            if not active[c]:
                continue
            rc = sum(1 for j in range(ns) if active[j] and P[c][j]) - 1
            cc = sum(1 for i in range(ns) if active[i] and P[i][c]) - 1
            prod = rc * cc
            if best < 0 or prod < bestprod:
                best, bestprod = c, prod
        perm.append(best)
        active[best] = False
        for i in range(ns):
            if active[i] and P[i][best]:
                for j in range(ns):
                    if active[j] and P[best][j]:
                        P[i][j] = True
    return perm, P


# rate-law families -> device helpers in mech_families.h
def rate_expr(fam, p, r):
    if fam == "unset":
        return "0.0"
    if fam == "photolysis":
        return "J[%d]" % (int(p[0]) - 1)
    if fam == "constant":
        return lit(p[0])
    if fam == "arrhenius":
        return "rl_arr(T, %s, %s)" % (lit(p[0]), lit(p[1]))
    if fam == "arrhenius_tn":
        return "rl_arr_tn(T, lnT, %s, %s, %s)" % (lit(p[0]), lit(p[1]), lit(p[2]))
    if fam == "exp_sum":
        c0 = p[0]
        terms = [p[1 + 3 * t: 4 + 3 * t] for t in range((len(p) - 1) // 3)]
        parts = []
        for c, n, b in terms:
            if n != 0.0:
                parts.append("%s * exp(%s * lnT - %s / T)" % (lit(c), lit(n), lit(b)))
            else:
                parts.append("%s * exp(-(%s / T))" % (lit(c), lit(b)))
        return "(" + " + ".join(parts) + ") + %s" % lit(c0)
    if fam == "m_power":
        return "M * %s * pow(T / 3.0e2, %s)" % (lit(p[0]), lit(p[1]))
    if fam == "exp_plus_m_exp":
        return "%s * exp(%s / T) + %s * M * exp(%s / T)" % tuple(lit(x) for x in p)
    if fam == "troe":
        return "rl_troe(T, M, %s, %s, %s, %s, %s)" % tuple(lit(x) for x in p)
    if fam == "hno3_oh":
        return "rl_hno3_oh(T, M, %s, %s, %s, %s, %s, %s)" % tuple(lit(x) for x in p)
    if fam == "linear_m":
        return "%s * (1.0 + %s * M)" % (lit(p[0]), lit(p[1]))
    raise ValueError(fam)


def main():
    m, ns, nr, react, prod = load()
    jt = jac_terms(ns, react, prod)
    struct = [[False] * ns for _ in range(ns)]
    for (i, j) in jt:
        struct[i][j] = True
    for i in range(ns):
        struct[i][i] = True
    # structural entries in row-major order: these carry the mask bits
    sentries = [(i, j) for i in range(ns) for j in range(ns) if struct[i][j]]
    sidx = {e: k for k, e in enumerate(sentries)}
    perm, filled = markowitz(ns, struct)
    pos = {s: k for k, s in enumerate(perm)}
    F = [[filled[perm[I]][perm[J]] for J in range(ns)] for I in range(ns)]
    # LU storage slot of each (I,J) in the filled pattern (internal indices)
    slot = {}
    for J in range(ns):
        for I in range(ns):
            if F[I][J]:
                slot[(I, J)] = len(slot)
    nlu = len(slot)
    nwords = (len(sentries) + 31) // 32

    L = []
    w = L.append
    w("// Generated by brams_amd/codegen/gen_mech.py from brams_amd/mechanism/relacs_tuv.json.")
    w("// RELACS_TUV: %d species, %d reactions, %d Jacobian entries, LU nnz %d." % (ns, nr, len(jt), nlu))
    w("#pragma once")
    w("#include \"mech_families.h\"")
    w("#define MECH_NS %d" % ns)
    w("#define MECH_NR %d" % nr)
    w("#define MECH_NPHOT %d" % m["nr_photo"])
    w("#define MECH_NSTRUCT %d" % len(sentries))
    w("#define MECH_MASK_WORDS %d" % nwords)
    w("#define MECH_NLU %d" % nlu)
    w("// pivot order (internal -> species, 0-based): %s" % perm)
    w("static const int mech_perm_host[MECH_NS] = {%s};" % ", ".join(map(str, perm)))
    w("static const int mech_struct_i_host[MECH_NSTRUCT] = {%s};" % ", ".join(str(i) for i, _ in sentries))
    w("static const int mech_struct_j_host[MECH_NSTRUCT] = {%s};" % ", ".join(str(j) for _, j in sentries))
    w("")
    # ---------------- kinetic
    w("// chem_spack_kinetic.f90:7-435")
    w("__device__ __forceinline__ void mech_kinetic(double T, double P, double vapp, const double* J, double* rk) {")
    w("  const double M = P * 7.243e16 / T;")
    w("  const double H2O = 29.0 * M * vapp / (18.0 + 11.0 * vapp);")
    w("  const double lnT = log(T);")
    w("  (void)H2O; (void)lnT;")
    for r, rx in enumerate(m["reactions"]):
        law = rx["rate"]
        e = rate_expr(law["family"], law["params"], r)
        post = law["post"]
        if post:
            kind = post[0][0]
            if kind == "mul_M_const":
                w("  rk[%d] = (%s) * M * %s;" % (r, e, lit(post[0][1])))
            elif kind == "mul_H2O":
                w("  rk[%d] = (%s) * H2O;" % (r, e))
            elif kind == "mul_troe":
                w("  rk[%d] = rl_troe(T, M, %s) * (%s);" % (r, ", ".join(lit(x) for x in post[0][1:]), e))
            else:
                raise ValueError(kind)
        else:
            w("  rk[%d] = %s;" % (r, e))
    w("}")
    w("")
    # ---------------- rates + fun
    def wexpr(r):
        rs = react[r]
        s = "rk[%d]" % r
        for x in rs:
            s = "%s * y[%d]" % (s, x)
        return s

    def sum_terms(terms):
        out = ""
        for k, (c, ref) in enumerate(terms):
            if c == 1.0:
                t = ref if k == 0 else " + " + ref
            elif c == -1.0:
                t = "-" + ref if k == 0 else " - " + ref
            elif c < 0:
                t = ("%s * %s" % (lit(c), ref)) if k == 0 else " - %s * %s" % (lit(-c), ref)
            else:
                t = ("%s * %s" % (lit(c), ref)) if k == 0 else " + %s * %s" % (lit(c), ref)
            out += t
        return out

    w("// chem_spack_rates.f90:63-192 and chem_spack_fexchem.f90:73-676")
    w("__device__ __forceinline__ void mech_fun(const double* rk, const double* y, double* f) {")
    used = sorted({r for terms in prod.values() for _, r in terms})
    for r in used:
        w("  const double w%d = %s;" % (r, wexpr(r)))
    for s in range(ns):
        terms = prod.get(s, [])
        if not terms:
            w("  f[%d] = 0.0;" % s)
        else:
            w("  f[%d] = %s;" % (s, sum_terms([(c, "w%d" % r) for c, r in terms])))
    w("}")
    w("")

    def jexpr(i, j):
        terms = []
        for c, r, o in jt[(i, j)]:
            ref = "(rk[%d] * y[%d])" % (r, o) if o >= 0 else "rk[%d]" % r
            terms.append((c, ref))
        return sum_terms(terms)

    def mval(i, j):
        """entry of 1/(gamma h) I - J for species (i,j)"""
        if (i, j) in jt:
            if i == j:
                return "igdt - (%s)" % jexpr(i, j)
            return "-(%s)" % jexpr(i, j)
        return "igdt" if i == j else None

    def mbit(e):
        return "((mw[%d] >> %d) & 1u)" % (e // 32, e % 32)

    # ---------------- mask of one cell
    w("// rodas3:470-482: keep the entries of DLmat that are nonzero in this cell")
    w("__device__ __forceinline__ void mech_mask(const double* rk, const double* y, double igdt, unsigned* mw) {")
    w("  unsigned acc[MECH_MASK_WORDS] = {0};")
    for e, (i, j) in enumerate(sentries):
        w("  acc[%d] |= ((%s) != 0.0 ? 1u : 0u) << %d;" % (e // 32, mval(i, j), e % 32))
    w("  for (int k = 0; k < MECH_MASK_WORDS; ++k) mw[k] = acc[k];")
    w("}")
    w("")
    # ---------------- factor
    w("// spFactor.c:336-401 on the fixed pivot order; lu[] keeps L (reciprocal")
    w("// diagonal) and unit-diagonal U.  Returns nonzero on a zero pivot.")
    w("__device__ __forceinline__ int mech_factor(const double* rk, const double* y, double igdt,")
    w("                                           const unsigned* mw, double* lu) {")
    w("  int bad = 0;")
    for J in range(ns):
        rows = [I for I in range(ns) if F[I][J]]
        w("  { // column %d (species %d)" % (J, perm[J]))
        for I in rows:
            i, j = perm[I], perm[J]
            v = mval(i, j)
            if v is None:
                w("    double d%d = 0.0;" % I)
            else:
                w("    double d%d = %s ? (%s) : 0.0;" % (I, mbit(sidx[(i, j)]), v))
        for K in [K for K in rows if K < J]:
            w("    { const double u = d%d * lu[%d]; lu[%d] = u;" % (K, slot[(K, K)], slot[(K, J)]))
            for I in range(K + 1, ns):
                if F[I][K]:
                    w("      d%d -= u * lu[%d];" % (I, slot[(I, K)]))
            w("    }")
        for I in rows:
            if I > J:
                w("    lu[%d] = d%d;" % (slot[(I, J)], I))
        w("    bad |= (d%d == 0.0); lu[%d] = 1.0 / d%d;" % (J, slot[(J, J)], J))
        w("  }")
    w("  return bad;")
    w("}")
    w("")
    # ---------------- solve
    w("// spSolve.c:183-218: x (species order) overwritten by the solution")
    w("__device__ __forceinline__ void mech_solve(const double* lu, double* x) {")
    for I in range(ns):
        w("  double c%d = x[%d];" % (I, perm[I]))
    for I in range(ns):
        w("  c%d *= lu[%d];" % (I, slot[(I, I)]))
        for R in range(I + 1, ns):
            if F[R][I]:
                w("  c%d -= c%d * lu[%d];" % (R, I, slot[(R, I)]))
    for I in range(ns - 1, -1, -1):
        for Jc in range(I + 1, ns):
            if F[I][Jc]:
                w("  c%d -= lu[%d] * c%d;" % (I, slot[(I, Jc)], Jc))
    for I in range(ns):
        w("  x[%d] = c%d;" % (perm[I], I))
    w("}")
    # ---------------- algorithmic FP64 operation counts of the generated code
    def nflop_sum(coefs):
        return (len(coefs) - 1) + sum(1 for c in coefs if abs(c) != 1.0)

    fl_fun = sum(len(react[r]) for r in used) + sum(
        nflop_sum([c for c, _ in prod[s]]) for s in prod)
    fl_jac = 0
    for (i, j), terms in jt.items():
        fl_jac += sum(1 for _, _, o in terms if o >= 0) + nflop_sum([c for c, _, _ in terms])
        fl_jac += 1 if i == j else 0
    fl_elim = 0
    for J in range(ns):
        for K in range(J):
            if F[K][J]:
                fl_elim += 1 + 2 * sum(1 for I in range(K + 1, ns) if F[I][K])
        fl_elim += 1                                    # reciprocal pivot
    nl = sum(1 for (I, J) in slot if I > J)
    nu = sum(1 for (I, J) in slot if I < J)
    fl_solve = ns + 2 * nl + 2 * nu
    tr = {"exp": 0, "log": 1, "log10": 0, "pow": 0}
    ar_kin = 7
    for rx in m["reactions"]:
        fam, p, post = rx["rate"]["family"], rx["rate"]["params"], rx["rate"]["post"]
        if fam == "arrhenius":
            tr["exp"] += 1; ar_kin += 2
        elif fam == "arrhenius_tn":
            tr["exp"] += 1; ar_kin += 4
        elif fam == "exp_sum":
            nt = (len(p) - 1) // 3
            tr["exp"] += nt; ar_kin += 4 * nt + 1
        elif fam == "m_power":
            tr["pow"] += 1; ar_kin += 3
        elif fam == "exp_plus_m_exp":
            tr["exp"] += 2; ar_kin += 6
        elif fam == "troe":
            tr["pow"] += 3; tr["log10"] += 1; ar_kin += 14
        elif fam == "hno3_oh":
            tr["exp"] += 3; ar_kin += 10
        elif fam == "linear_m":
            ar_kin += 3
        for q in post:
            if q[0] == "mul_M_const":
                ar_kin += 2
            elif q[0] == "mul_H2O":
                ar_kin += 1
            elif q[0] == "mul_troe":
                tr["pow"] += 3; tr["log10"] += 1; ar_kin += 15
    flops = {"fun": fl_fun, "jacobian_assembly": fl_jac, "elimination": fl_elim,
             "factor": fl_jac + fl_elim, "solve": fl_solve, "nlu": nlu, "nl": nl, "nu": nu,
             "kinetic_arith": ar_kin, "kinetic_transcendentals": tr,
             "note": "FP64 add/sub/mul/div counted once each (an FMA is 2), "
                     "transcendentals listed separately"}
    for k in ("fun", "factor", "solve"):
        w("#define MECH_FLOPS_%s %d" % (k.upper(), flops[k]))

    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(os.path.join(os.path.dirname(OUT), "relacs_tuv_flops.json"), "w") as fh:
        json.dump(flops, fh, indent=1)
    text = "\n".join(L) + "\n"
    out = sys.argv[1] if len(sys.argv) > 1 else OUT
    old = open(out).read() if os.path.exists(out) else None
    if old != text:
        with open(out, "w") as fh:
            fh.write(text)
    print("generated %s: nstruct=%d nlu=%d" % (out, len(sentries), nlu))


if __name__ == "__main__":
    main()
